// v0.hip -- the V0 basis and its Galerkin Laplacians on the GPU:
//   merge_v0d1 (src/fdtd.hpp:1823-2774)  -> V0d, V0da   (dielectric node columns, then one
//                                           aggregate column per non-reference conductor)
//   generateAd (src/fdtd.hpp:2778-2862)  -> Ad = V0da^T D_eps V0d  (|v| > 1e-8 kept)
//   merge_v0c  (src/fdtd.hpp:2866-3479)  -> V0c, V0ca   (one column per conductor node;
//                                           floating/excited conductors omit their last node)
//   generateAc (src/fdtd.hpp:3481-3575)  -> Ac = V0ca^T D_sig V0c over conductor edges
//
// The reference's block sizes are zero (src/matrixCon.cpp:32-40), so every node
// group is a singleton and the column layout reduces to prefix sums over node
// flags.  Entry order, values and the accumulation order of every Ad/Ac cell
// match the reference: contributions are expanded in entry order, radix-sorted
// stably by (row, column) and each cell is summed sequentially in that order.
// Built with FP contraction off.
#pragma clang fp contract(off)

#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "v0.hpp"

namespace gp {

void destroy_v0(V0Data *v) { delete v; }

namespace {

struct Geo {
  const double *xn, *yn, *zn;
  Grid g;
  int64_t nx, ny, nz;
};

__device__ __forceinline__ double avgl(const double *c, int64_t i, int64_t n) {
  if (i == 0) return c[1] - c[0];
  if (i == n - 1) return c[i] - c[i - 1];
  return (c[i + 1] - c[i - 1]) / 2;
}

// The six incident edges of node v in the reference order: lower z, upper z,
// left x, right x, front y, back y.  val = +-1/h, axis 0=x 1=y 2=z, sgn +-1.
struct Inc { int64_t e, other; double val; int axis, sgn; };

__device__ int incident(const Geo &G, int64_t v, Inc out[6]) {
  const Grid &g = G.g;
  int64_t iz = v / g.Nns, ix = (v % g.Nns) / (g.Ny + 1), iy = (v % g.Nns) % (g.Ny + 1);
  int k = 0;
  if (iz != 0) out[k++] = {g.ez(iz - 1, ix, iy), g.node(iz - 1, ix, iy), -1 / (G.zn[iz] - G.zn[iz - 1]), 2, -1};
  if (iz != G.nz - 1) out[k++] = {g.ez(iz, ix, iy), g.node(iz + 1, ix, iy), 1 / (G.zn[iz + 1] - G.zn[iz]), 2, 1};
  if (ix != 0) out[k++] = {g.ex(iz, ix - 1, iy), g.node(iz, ix - 1, iy), -1 / (G.xn[ix] - G.xn[ix - 1]), 0, -1};
  if (ix != G.nx - 1) out[k++] = {g.ex(iz, ix, iy), g.node(iz, ix + 1, iy), 1 / (G.xn[ix + 1] - G.xn[ix]), 0, 1};
  if (iy != 0) out[k++] = {g.ey(iz, ix, iy - 1), g.node(iz, ix, iy - 1), -1 / (G.yn[iy] - G.yn[iy - 1]), 1, -1};
  if (iy != G.ny - 1) out[k++] = {g.ey(iz, ix, iy), g.node(iz, ix, iy + 1), 1 / (G.yn[iy + 1] - G.yn[iy]), 1, 1};
  return k;
}

// ------------------------------------------------------------------ V0d
__global__ void k_diel_flags(Geo G, const int32_t *__restrict__ mark_node, const uint8_t *__restrict__ bd,
                             int64_t *__restrict__ elig, int64_t *__restrict__ cnt) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= G.g.Nn) return;
  int64_t iz = v / G.g.Nns;
  bool ok = mark_node[v] == 0 && !((iz == 0 || iz == G.nz - 1) && bd[v]);
  elig[v] = ok;
  if (!ok) { cnt[v] = 0; return; }
  Inc in[6];
  cnt[v] = incident(G, v, in);
}

__global__ void k_diel_fill(Geo G, const int64_t *__restrict__ elig, const int64_t *__restrict__ colscan,
                            const int64_t *__restrict__ off, int64_t *__restrict__ row, int64_t *__restrict__ col,
                            double *__restrict__ val, double *__restrict__ aval, int64_t *__restrict__ map,
                            int64_t *__restrict__ node_of_col) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= G.g.Nn || !elig[v]) return;
  const Grid &g = G.g;
  int64_t iz = v / g.Nns, ix = (v % g.Nns) / (g.Ny + 1), iy = (v % g.Nns) % (g.Ny + 1);
  const double lx = avgl(G.xn, ix, G.nx), ly = avgl(G.yn, iy, G.ny), lz = avgl(G.zn, iz, G.nz);
  const int64_t c = colscan[v];
  map[v] = c + 1;
  node_of_col[c] = v;
  Inc in[6];
  int k = incident(G, v, in);
  int64_t o = off[v];
  for (int q = 0; q < k; ++q, ++o) {
    const double l = in[q].axis == 0 ? lx : in[q].axis == 1 ? ly : lz;
    row[o] = in[q].e;
    col[o] = c;
    val[o] = in[q].val;
    aval[o] = in[q].sgn > 0 ? 1 / l : -1 / l;
  }
}

// conductor nodes sorted by (conductor, node): count dielectric-facing edges
__global__ void k_cond_count(Geo G, int64_t ncn, const int64_t *__restrict__ nodes, const int32_t *__restrict__ mark_node,
                             const int32_t *__restrict__ mark_edge, const int32_t *__restrict__ mark_port,
                             int64_t *__restrict__ cnt) {
  int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ncn) return;
  const int64_t v = nodes[a];
  const int32_t c = mark_node[v];
  if (mark_port[c - 1] == -1) { cnt[a] = 0; return; }
  Inc in[6];
  int k = incident(G, v, in);
  int64_t n = 0;
  for (int q = 0; q < k; ++q)
    if (mark_edge[in[q].e] == 0 && mark_node[in[q].other] != c) ++n;
  cnt[a] = n;
}

__global__ void k_cond_fill(Geo G, int64_t ncn, const int64_t *__restrict__ nodes, const int32_t *__restrict__ mark_node,
                            const int32_t *__restrict__ mark_edge, const int32_t *__restrict__ mark_port,
                            const int64_t *__restrict__ off, const int64_t *__restrict__ fillcol,
                            const int64_t *__restrict__ mapcol, double lxw, double lyw, double lzw,
                            int64_t *__restrict__ row, int64_t *__restrict__ col, double *__restrict__ val,
                            double *__restrict__ aval, int64_t *__restrict__ map) {
  int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ncn) return;
  const int64_t v = nodes[a];
  const int32_t c = mark_node[v];
  if (mark_port[c - 1] == -1) return;
  map[v] = mapcol[c - 1];
  const Grid &g = G.g;
  int64_t iz = v / g.Nns, ix = (v % g.Nns) / (g.Ny + 1), iy = (v % g.Nns) % (g.Ny + 1);
  const double lx = avgl(G.xn, ix, G.nx), ly = avgl(G.yn, iy, G.ny), lz = avgl(G.zn, iz, G.nz);
  Inc in[6];
  int k = incident(G, v, in);
  int64_t o = off[a];
  const int64_t cc = fillcol[c - 1];
  for (int q = 0; q < k; ++q) {
    if (!(mark_edge[in[q].e] == 0 && mark_node[in[q].other] != c)) continue;
    double av;
    // src/fdtd.hpp:2687-2757: -lx_avg * ly_avg / (lx_whole_avg * ly_whole_avg * lz_whole_avg) etc.
    if (in[q].axis == 2) av = in[q].sgn > 0 ? lx * ly / (lxw * lyw * lzw) : -lx * ly / (lxw * lyw * lzw);
    else if (in[q].axis == 0) av = in[q].sgn > 0 ? ly * lz / (lxw * lyw * lzw) : -ly * lz / (lxw * lyw * lzw);
    else av = in[q].sgn > 0 ? lx * lz / (lxw * lyw * lzw) : -lx * lz / (lxw * lyw * lzw);
    row[o] = in[q].e;
    col[o] = cc;
    val[o] = in[q].val;
    aval[o] = av;
    ++o;
  }
}

// ------------------------------------------------------------------ V0c
__global__ void k_v0c_flags(Geo G, int64_t ncn, const int64_t *__restrict__ nodes, const int64_t *__restrict__ rank,
                            const int64_t *__restrict__ csize, const int32_t *__restrict__ mark_node,
                            const int32_t *__restrict__ mark_port, const uint8_t *__restrict__ bd,
                            int64_t *__restrict__ elig, int64_t *__restrict__ cnt) {
  int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ncn) return;
  const int64_t v = nodes[a];
  const int32_t c = mark_node[v];
  const int32_t mp = mark_port[c - 1];
  const int64_t n = mp <= -1 ? csize[c - 1] : csize[c - 1] - 1;
  bool ok = rank[a] < n && !(mp <= 0 && bd[v]);
  elig[a] = ok;
  if (!ok) { cnt[a] = 0; return; }
  Inc in[6];
  cnt[a] = incident(G, v, in);
}

__global__ void k_v0c_fill(Geo G, int64_t ncn, const int64_t *__restrict__ nodes, const int64_t *__restrict__ elig,
                           const int64_t *__restrict__ colscan, const int64_t *__restrict__ off,
                           int64_t *__restrict__ row, int64_t *__restrict__ col, double *__restrict__ val,
                           double *__restrict__ aval, int64_t *__restrict__ map, int64_t *__restrict__ node_of_col) {
  int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ncn || !elig[a]) return;
  const int64_t v = nodes[a];
  const Grid &g = G.g;
  int64_t iz = v / g.Nns, ix = (v % g.Nns) / (g.Ny + 1), iy = (v % g.Nns) % (g.Ny + 1);
  const double lx = avgl(G.xn, ix, G.nx), ly = avgl(G.yn, iy, G.ny), lz = avgl(G.zn, iz, G.nz);
  const int64_t c = colscan[a];
  map[v] = c + 1;
  node_of_col[c] = v;
  Inc in[6];
  int k = incident(G, v, in);
  int64_t o = off[a];
  for (int q = 0; q < k; ++q, ++o) {
    const double l = in[q].axis == 0 ? lx : in[q].axis == 1 ? ly : lz;
    row[o] = in[q].e;
    col[o] = c;
    val[o] = in[q].val;
    aval[o] = in[q].sgn > 0 ? 1 / l : -1 / l;
  }
}

// ------------------------------------------------------------------ Galerkin (Ad / Ac)
// Expansion of generateAd / generateAc: each entry yields 1 or 2 contributions
// (row = entry column, key = other column or the diagonal).  ncon[i] = count.
__global__ void k_gal_count(Geo G, int64_t nent, const int64_t *__restrict__ row, const int64_t *__restrict__ col,
                            const int64_t *__restrict__ map, const int32_t *__restrict__ mark_edge, int sigma_mode,
                            int64_t *__restrict__ ncon) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nent) return;
  int64_t n1, n2;
  G.g.ends(row[i], n1, n2);
  const bool use = sigma_mode ? mark_edge[row[i]] != 0 : true;
  const int64_t c1 = col[i] + 1;
  int64_t n = 0;
  if (use) n = ((map[n1] != c1 && map[n1] != 0) || (map[n2] != c1 && map[n2] != 0)) ? 2 : 1;
  ncon[i] = n;
}

__global__ void k_gal_expand(Geo G, int64_t nent, const int64_t *__restrict__ row, const int64_t *__restrict__ col,
                             const double *__restrict__ aval, const int64_t *__restrict__ map,
                             const int32_t *__restrict__ mark_edge, const double *__restrict__ stack, int sigma_mode,
                             const int64_t *__restrict__ off, int64_t ncols, uint64_t *__restrict__ key,
                             double *__restrict__ v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nent) return;
  const int64_t e = row[i];
  int d; int64_t iz, ix, iy;
  G.g.decode(e, d, iz, ix, iy);
  const double h = d == 2 ? G.zn[iz + 1] - G.zn[iz] : d == 1 ? G.xn[ix + 1] - G.xn[ix] : G.yn[iy + 1] - G.yn[iy];
  int64_t n1, n2;
  G.g.ends(e, n1, n2);
  const bool use = sigma_mode ? mark_edge[e] != 0 : true;
  if (!use) return;
  const double w = stack[iz], av = aval[i];
  const int64_t c = col[i], c1 = c + 1;
  int64_t o = off[i];
  // generateAd: v0d1aval * (+-1) / h * stackEpsn * EPSILON0;  generateAc: v0caval * (+-1) / h * stackSign
#define TERM(S) (sigma_mode ? av * (S) / h * w : av * (S) / h * w * kEps0)
  if (map[n1] != c1 && map[n1] != 0) {
    key[o] = (uint64_t)c * (uint64_t)ncols + (uint64_t)(map[n1] - 1); v[o] = TERM(1); ++o;
    key[o] = (uint64_t)c * (uint64_t)ncols + (uint64_t)c; v[o] = TERM(-1);
  } else if (map[n2] != c1 && map[n2] != 0) {
    key[o] = (uint64_t)c * (uint64_t)ncols + (uint64_t)(map[n2] - 1); v[o] = TERM(-1); ++o;
    key[o] = (uint64_t)c * (uint64_t)ncols + (uint64_t)c; v[o] = TERM(1);
  } else {
    key[o] = (uint64_t)c * (uint64_t)ncols + (uint64_t)c; v[o] = fabs(TERM(1));
  }
#undef TERM
}

__global__ void k_iota32(int64_t n, uint32_t *__restrict__ a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = (uint32_t)i;
}
__global__ void k_heads(int64_t n, const uint64_t *__restrict__ key, int64_t *__restrict__ head) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) head[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}
__global__ void k_seg_sum(int64_t n, const int64_t *__restrict__ head, const int64_t *__restrict__ hscan,
                          const uint64_t *__restrict__ key, const uint32_t *__restrict__ perm,
                          const double *__restrict__ v, int64_t ncols, int filter, int64_t *__restrict__ keep,
                          int64_t *__restrict__ srow, int64_t *__restrict__ scol, double *__restrict__ sval) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !head[i]) return;
  double s = 0.;
  int64_t j = i;
  for (; j < n && (j == i || !head[j]); ++j) s += v[perm[j]];
  const int64_t seg = hscan[i];
  srow[seg] = (int64_t)(key[i] / (uint64_t)ncols);
  scol[seg] = (int64_t)(key[i] % (uint64_t)ncols);
  sval[seg] = s;
  keep[seg] = filter ? (fabs(s) > 1e-8) : 1;
}
__global__ void k_compact(int64_t n, const int64_t *__restrict__ keep, const int64_t *__restrict__ kscan,
                          const int64_t *__restrict__ a, const int64_t *__restrict__ b, const double *__restrict__ c,
                          int64_t *__restrict__ oa, int64_t *__restrict__ ob, double *__restrict__ oc) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !keep[i]) return;
  const int64_t o = kscan[i];
  oa[o] = a[i]; ob[o] = b[i]; oc[o] = c[i];
}

__global__ void k_cond_key(int64_t nn, const int32_t *__restrict__ mark_node, uint32_t *__restrict__ key,
                           int64_t *__restrict__ idx, int64_t *__restrict__ cnt) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  key[v] = (uint32_t)mark_node[v];
  idx[v] = v;
  (void)cnt;
}

__global__ void k_rank(int64_t ncn, const int64_t *__restrict__ nodes, const int32_t *__restrict__ mark_node,
                       const int64_t *__restrict__ cstart, int64_t *__restrict__ rank) {
  int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ncn) return;
  rank[a] = a - cstart[mark_node[nodes[a]] - 1];
}

}  // namespace

// Galerkin product of a V0 basis, shared by Ad and Ac.
static gp_status galerkin(gp_layout *lay, const Geo &G, const V0Coo &V, int64_t ncols, const int64_t *map,
                          int sigma_mode, int filter, V0Coo *out, int64_t *len_unfiltered, Arena &A) {
  hipStream_t s = lay->stream;
  Arena tmp;
  gp_status st = GP_OK;
  int64_t *ncon = nullptr, *off = nullptr;
  uint64_t *key = nullptr, *key2 = nullptr;
  double *v = nullptr;
  uint32_t *perm = nullptr, *perm2 = nullptr;
  int64_t *head = nullptr, *hscan = nullptr, *keep = nullptr, *kscan = nullptr, *srow = nullptr, *scol = nullptr;
  double *sval = nullptr;
  *out = V0Coo();
  *len_unfiltered = 0;
  do {
    if (V.n == 0) break;
    if ((st = tmp.alloc(&ncon, V.n)) != GP_OK) break;
    if ((st = tmp.alloc(&off, V.n + 1)) != GP_OK) break;
    k_gal_count<<<blocks_for(V.n), 256, 0, s>>>(G, V.n, V.row, V.col, map, lay->mark_edge, sigma_mode, ncon);
    if ((st = exclusive_scan_i64(ncon, off, V.n, s, tmp)) != GP_OK) break;
    int64_t np = 0;
    if (hipMemcpyAsync(&np, off + V.n, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "count"); break; }
    if (np == 0) break;
    if (np > 2147483647LL) { st = fail(GP_ERR_NOMEM, "V0 Galerkin expansion too large"); break; }
    if ((st = tmp.alloc(&key, np)) != GP_OK) break;
    if ((st = tmp.alloc(&key2, np)) != GP_OK) break;
    if ((st = tmp.alloc(&v, np)) != GP_OK) break;
    if ((st = tmp.alloc(&perm, np)) != GP_OK) break;
    if ((st = tmp.alloc(&perm2, np)) != GP_OK) break;
    k_gal_expand<<<blocks_for(V.n), 256, 0, s>>>(G, V.n, V.row, V.col, V.aval, map, lay->mark_edge,
                                                 sigma_mode ? lay->sign : lay->epsn, sigma_mode, off, ncols, key, v);
    k_iota32<<<blocks_for(np), 256, 0, s>>>(np, perm);
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, key, key2, perm, perm2, (int)np, 0, 64, s) != hipSuccess) { st = fail(GP_ERR_HIP, "sort size"); break; }
    uint8_t *buf = nullptr;
    if ((st = tmp.alloc(&buf, (int64_t)bytes + 16)) != GP_OK) break;
    if (hipcub::DeviceRadixSort::SortPairs(buf, bytes, key, key2, perm, perm2, (int)np, 0, 64, s) != hipSuccess) { st = fail(GP_ERR_HIP, "sort"); break; }
    if ((st = tmp.alloc(&head, np)) != GP_OK) break;
    if ((st = tmp.alloc(&hscan, np + 1)) != GP_OK) break;
    k_heads<<<blocks_for(np), 256, 0, s>>>(np, key2, head);
    if ((st = exclusive_scan_i64(head, hscan, np, s, tmp)) != GP_OK) break;
    int64_t nseg = 0;
    if (hipMemcpyAsync(&nseg, hscan + np, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "segments"); break; }
    if ((st = tmp.alloc(&keep, nseg)) != GP_OK) break;
    if ((st = tmp.alloc(&kscan, nseg + 1)) != GP_OK) break;
    if ((st = tmp.alloc(&srow, nseg)) != GP_OK) break;
    if ((st = tmp.alloc(&scol, nseg)) != GP_OK) break;
    if ((st = tmp.alloc(&sval, nseg)) != GP_OK) break;
    k_seg_sum<<<blocks_for(np), 256, 0, s>>>(np, head, hscan, key2, perm2, v, ncols, filter, keep, srow, scol, sval);
    if ((st = exclusive_scan_i64(keep, kscan, nseg, s, tmp)) != GP_OK) break;
    int64_t nk = 0;
    if (hipMemcpyAsync(&nk, kscan + nseg, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "kept"); break; }
    out->n = nk;
    if ((st = A.alloc(&out->row, std::max<int64_t>(nk, 1))) != GP_OK) break;
    if ((st = A.alloc(&out->col, std::max<int64_t>(nk, 1))) != GP_OK) break;
    if ((st = A.alloc(&out->val, std::max<int64_t>(nk, 1))) != GP_OK) break;
    k_compact<<<blocks_for(nseg), 256, 0, s>>>(nseg, keep, kscan, srow, scol, sval, out->row, out->col, out->val);
    *len_unfiltered = nseg;
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) { st = fail(GP_ERR_HIP, "galerkin kernels"); break; }
  } while (0);
  (void)hipStreamSynchronize(s);
  tmp.free_all();
  return st;
}

gp_status build_v0(gp_layout *lay) {
  if (lay->v0) return GP_OK;
  const gp_mesh &m = lay->hm.view;
  const Grid &g = lay->g;
  hipStream_t s = lay->stream;
  Arena &A = lay->arena;
  Geo G{lay->xn, lay->yn, lay->zn, g, m.nx, m.ny, m.nz};
  auto *V = new V0Data();
  Arena tmp;
  gp_status st = GP_OK;
  const int32_t ncdt = m.num_cdt;
  do {
    // conductor nodes sorted by (conductor, node)
    uint32_t *ck = nullptr, *ck2 = nullptr;
    int64_t *cidx = nullptr, *cidx2 = nullptr;
    if ((st = tmp.alloc(&ck, g.Nn)) != GP_OK) break;
    if ((st = tmp.alloc(&ck2, g.Nn)) != GP_OK) break;
    if ((st = tmp.alloc(&cidx, g.Nn)) != GP_OK) break;
    if ((st = tmp.alloc(&cidx2, g.Nn)) != GP_OK) break;
    k_cond_key<<<blocks_for(g.Nn), 256, 0, s>>>(g.Nn, lay->mark_node, ck, cidx, nullptr);
    size_t bytes = 0;
    int endbit = 1;
    while (endbit < 32 && ((uint64_t)ncdt >> endbit)) ++endbit;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, ck, ck2, cidx, cidx2, (int)g.Nn, 0, endbit, s) != hipSuccess) { st = fail(GP_ERR_HIP, "sort size"); break; }
    uint8_t *buf = nullptr;
    if ((st = tmp.alloc(&buf, (int64_t)bytes + 16)) != GP_OK) break;
    if (hipcub::DeviceRadixSort::SortPairs(buf, bytes, ck, ck2, cidx, cidx2, (int)g.Nn, 0, endbit, s) != hipSuccess) { st = fail(GP_ERR_HIP, "sort"); break; }
    // per-conductor sizes from the host copy of mark_node
    std::vector<int64_t> csize(std::max(ncdt, 1), 0), cstart(std::max(ncdt, 1), 0);
    int64_t ndnodes = 0;
    for (int64_t v = 0; v < g.Nn; ++v) {
      int32_t c = m.mark_node[v];
      if (c > 0) csize[c - 1]++;
      else ndnodes++;
    }
    int64_t ncn = g.Nn - ndnodes, acc = 0;
    for (int32_t c = 0; c < ncdt; ++c) { cstart[c] = acc; acc += csize[c]; }
    const int64_t *cnodes = cidx2 + ndnodes;  // conductor nodes follow the dielectric ones (key 0)
    int64_t *d_csize = nullptr, *d_cstart = nullptr, *rank = nullptr;
    int32_t *d_mp = nullptr;
    if ((st = tmp.alloc(&d_csize, std::max(ncdt, 1))) != GP_OK) break;
    if ((st = tmp.alloc(&d_cstart, std::max(ncdt, 1))) != GP_OK) break;
    if ((st = tmp.alloc(&d_mp, std::max(ncdt, 1))) != GP_OK) break;
    if ((st = tmp.alloc(&rank, std::max<int64_t>(ncn, 1))) != GP_OK) break;
    if (ncdt) {
      if (hipMemcpyAsync(d_csize, csize.data(), sizeof(int64_t) * ncdt, hipMemcpyHostToDevice, s) != hipSuccess ||
          hipMemcpyAsync(d_cstart, cstart.data(), sizeof(int64_t) * ncdt, hipMemcpyHostToDevice, s) != hipSuccess ||
          hipMemcpyAsync(d_mp, m.cdt_mark_port, sizeof(int32_t) * ncdt, hipMemcpyHostToDevice, s) != hipSuccess) { st = fail(GP_ERR_HIP, "upload"); break; }
    }
    if (ncn) k_rank<<<blocks_for(ncn), 256, 0, s>>>(ncn, cnodes, lay->mark_node, d_cstart, rank);

    // ---- V0d: dielectric singleton columns
    int64_t *elig = nullptr, *cnt = nullptr, *colscan = nullptr, *off = nullptr;
    if ((st = tmp.alloc(&elig, g.Nn)) != GP_OK) break;
    if ((st = tmp.alloc(&cnt, g.Nn)) != GP_OK) break;
    if ((st = tmp.alloc(&colscan, g.Nn + 1)) != GP_OK) break;
    if ((st = tmp.alloc(&off, g.Nn + 1)) != GP_OK) break;
    k_diel_flags<<<blocks_for(g.Nn), 256, 0, s>>>(G, lay->mark_node, lay->bd_node, elig, cnt);
    if ((st = exclusive_scan_i64(elig, colscan, g.Nn, s, tmp)) != GP_OK) break;
    if ((st = exclusive_scan_i64(cnt, off, g.Nn, s, tmp)) != GP_OK) break;
    int64_t ndiel = 0, ndent = 0;
    if (hipMemcpyAsync(&ndiel, colscan + g.Nn, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&ndent, off + g.Nn, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "counts"); break; }
    // conductor columns: entries per node, per-conductor totals on the host
    int64_t *ccnt = nullptr, *coff = nullptr;
    if ((st = tmp.alloc(&ccnt, std::max<int64_t>(ncn, 1))) != GP_OK) break;
    if ((st = tmp.alloc(&coff, ncn + 1)) != GP_OK) break;
    if (ncn) k_cond_count<<<blocks_for(ncn), 256, 0, s>>>(G, ncn, cnodes, lay->mark_node, lay->mark_edge, d_mp, ccnt);
    if ((st = exclusive_scan_i64(ccnt, coff, ncn, s, tmp)) != GP_OK) break;
    std::vector<int64_t> hcoff(ncn + 1);
    if (hipMemcpyAsync(hcoff.data(), coff, sizeof(int64_t) * (ncn + 1), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "cond offsets"); break; }
    // count pass: map column = running count, advancing only for conductors with
    // entries (src/fdtd.hpp:2430-2500); fill pass: `mark` is not reset between
    // conductors (src/fdtd.hpp:2673-2760)
    std::vector<int64_t> mapcol(std::max(ncdt, 1), 0), fillcol(std::max(ncdt, 1), 0);
    int64_t count = ndiel + 1;
    int mark = 0;
    for (int32_t c = 0; c < ncdt; ++c) {
      if (m.cdt_mark_port[c] == -1) continue;
      int64_t ents = hcoff[cstart[c] + csize[c]] - hcoff[cstart[c]];
      mapcol[c] = count;
      mark = ents > 0 ? 1 : 0;
      if (mark == 1) count++;
    }
    int64_t leng = ndiel;
    for (int32_t c = 0; c < ncdt; ++c) {
      if (m.cdt_mark_port[c] == -1) continue;
      int64_t ents = hcoff[cstart[c] + csize[c]] - hcoff[cstart[c]];
      fillcol[c] = leng;
      if (ents > 0) mark = 1;
      if (mark == 1) leng++;
    }
    V->v0d_cols = leng;
    V->ndiel = ndiel;
    const int64_t ncent = hcoff[ncn];
    V->v0d.n = ndent + ncent;
    if ((st = A.alloc(&V->v0d.row, std::max<int64_t>(V->v0d.n, 1))) != GP_OK) break;
    if ((st = A.alloc(&V->v0d.col, std::max<int64_t>(V->v0d.n, 1))) != GP_OK) break;
    if ((st = A.alloc(&V->v0d.val, std::max<int64_t>(V->v0d.n, 1))) != GP_OK) break;
    if ((st = A.alloc(&V->v0d.aval, std::max<int64_t>(V->v0d.n, 1))) != GP_OK) break;
    if ((st = A.alloc(&V->mapd, g.Nn)) != GP_OK) break;
    if ((st = A.alloc(&V->dnode_of_col, std::max<int64_t>(ndiel, 1))) != GP_OK) break;
    if (hipMemsetAsync(V->mapd, 0, sizeof(int64_t) * g.Nn, s) != hipSuccess) { st = fail(GP_ERR_HIP, "memset"); break; }
    k_diel_fill<<<blocks_for(g.Nn), 256, 0, s>>>(G, elig, colscan, off, V->v0d.row, V->v0d.col, V->v0d.val, V->v0d.aval, V->mapd,
                                                  V->dnode_of_col);
    if (ncn) {
      int64_t *d_fill = nullptr, *d_mapc = nullptr, *coff2 = nullptr;
      if ((st = tmp.alloc(&d_fill, std::max(ncdt, 1))) != GP_OK) break;
      if ((st = tmp.alloc(&d_mapc, std::max(ncdt, 1))) != GP_OK) break;
      if ((st = tmp.alloc(&coff2, ncn + 1)) != GP_OK) break;
      std::vector<int64_t> hoff2(ncn + 1);
      for (int64_t a = 0; a <= ncn; ++a) hoff2[a] = hcoff[a] + ndent;
      if (hipMemcpyAsync(d_fill, fillcol.data(), sizeof(int64_t) * ncdt, hipMemcpyHostToDevice, s) != hipSuccess ||
          hipMemcpyAsync(d_mapc, mapcol.data(), sizeof(int64_t) * ncdt, hipMemcpyHostToDevice, s) != hipSuccess ||
          hipMemcpyAsync(coff2, hoff2.data(), sizeof(int64_t) * (ncn + 1), hipMemcpyHostToDevice, s) != hipSuccess) { st = fail(GP_ERR_HIP, "upload"); break; }
      const double lxw = (m.xn[m.nx - 1] - m.xn[0]) / (m.nx - 1);
      const double lyw = (m.yn[m.ny - 1] - m.yn[0]) / (m.ny - 1);
      const double lzw = (m.zn[m.nz - 1] - m.zn[0]) / (m.nz - 1);
      k_cond_fill<<<blocks_for(ncn), 256, 0, s>>>(G, ncn, cnodes, lay->mark_node, lay->mark_edge, d_mp, coff2, d_fill, d_mapc,
                                                  lxw, lyw, lzw, V->v0d.row, V->v0d.col, V->v0d.val, V->v0d.aval, V->mapd);
      if (hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "fill"); break; }
    }
    // Ad
    if ((st = galerkin(lay, G, V->v0d, V->v0d_cols, V->mapd, 0, 1, &V->ad, &V->ad_len, A)) != GP_OK) break;
    // the reference sizes the Ad arrays to the unfiltered count (src/fdtd.hpp:2839-2845);
    // ad_len keeps that count, the trailing (0, 0, 0) slots are implicit

    // ---- V0c
    int64_t *celig = nullptr, *ccnt2 = nullptr, *ccol = nullptr, *coff3 = nullptr;
    if ((st = tmp.alloc(&celig, std::max<int64_t>(ncn, 1))) != GP_OK) break;
    if ((st = tmp.alloc(&ccnt2, std::max<int64_t>(ncn, 1))) != GP_OK) break;
    if ((st = tmp.alloc(&ccol, ncn + 1)) != GP_OK) break;
    if ((st = tmp.alloc(&coff3, ncn + 1)) != GP_OK) break;
    if (ncn) k_v0c_flags<<<blocks_for(ncn), 256, 0, s>>>(G, ncn, cnodes, rank, d_csize, lay->mark_node, d_mp, lay->bd_node, celig, ccnt2);
    if ((st = exclusive_scan_i64(celig, ccol, ncn, s, tmp)) != GP_OK) break;
    if ((st = exclusive_scan_i64(ccnt2, coff3, ncn, s, tmp)) != GP_OK) break;
    int64_t nvc = 0, nvcent = 0;
    if (hipMemcpyAsync(&nvc, ccol + ncn, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&nvcent, coff3 + ncn, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "v0c counts"); break; }
    V->v0c_cols = nvc;
    V->v0c.n = nvcent;
    if ((st = A.alloc(&V->v0c.row, std::max<int64_t>(nvcent, 1))) != GP_OK) break;
    if ((st = A.alloc(&V->v0c.col, std::max<int64_t>(nvcent, 1))) != GP_OK) break;
    if ((st = A.alloc(&V->v0c.val, std::max<int64_t>(nvcent, 1))) != GP_OK) break;
    if ((st = A.alloc(&V->v0c.aval, std::max<int64_t>(nvcent, 1))) != GP_OK) break;
    if ((st = A.alloc(&V->mapc, g.Nn)) != GP_OK) break;
    if ((st = A.alloc(&V->cnode_of_col, std::max<int64_t>(nvc, 1))) != GP_OK) break;
    if (hipMemsetAsync(V->mapc, 0, sizeof(int64_t) * g.Nn, s) != hipSuccess) { st = fail(GP_ERR_HIP, "memset"); break; }
    if (ncn) k_v0c_fill<<<blocks_for(ncn), 256, 0, s>>>(G, ncn, cnodes, celig, ccol, coff3, V->v0c.row, V->v0c.col, V->v0c.val, V->v0c.aval, V->mapc,
                                                         V->cnode_of_col);
    if ((st = galerkin(lay, G, V->v0c, V->v0c_cols, V->mapc, 1, 0, &V->ac, &V->ac_len, A)) != GP_OK) break;
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) { st = fail(GP_ERR_HIP, "v0 kernels"); break; }
  } while (0);
  (void)hipStreamSynchronize(s);
  tmp.free_all();
  if (st != GP_OK) { delete V; return st; }
  lay->v0 = V;
  return GP_OK;
}

}  // namespace gp

using namespace gp;

extern "C" gp_status gp_build_v0(gp_layout *lay) {
  if (!lay) return fail(GP_ERR_ARG, "null layout");
  return build_v0(lay);
}

extern "C" gp_status gp_v0_sizes_get(const gp_layout *lay, gp_v0_sizes *out) {
  if (!lay || !out) return fail(GP_ERR_ARG, "null argument");
  if (!lay->v0) return fail(GP_ERR_STATE, "V0 basis not built");
  const V0Data &V = *lay->v0;
  out->v0d_nnz = V.v0d.n; out->v0d_cols = V.v0d_cols;
  out->ad_nnz = V.ad.n; out->ad_len = V.ad_len;
  out->v0c_nnz = V.v0c.n; out->v0c_cols = V.v0c_cols;
  out->ac_nnz = V.ac.n;
  return GP_OK;
}

static gp_status copy_coo(const V0Coo &c, int64_t *row, int64_t *col, double *val, double *aval) {
  if (c.n == 0) return GP_OK;
  if (row) GP_HIP(hipMemcpy(row, c.row, sizeof(int64_t) * c.n, hipMemcpyDeviceToHost));
  if (col) GP_HIP(hipMemcpy(col, c.col, sizeof(int64_t) * c.n, hipMemcpyDeviceToHost));
  if (val) GP_HIP(hipMemcpy(val, c.val, sizeof(double) * c.n, hipMemcpyDeviceToHost));
  if (aval && c.aval) GP_HIP(hipMemcpy(aval, c.aval, sizeof(double) * c.n, hipMemcpyDeviceToHost));
  return GP_OK;
}

extern "C" gp_status gp_v0d_copy(const gp_layout *lay, int64_t *row, int64_t *col, double *val, double *aval) {
  if (!lay || !lay->v0) return fail(GP_ERR_STATE, "V0 basis not built");
  return copy_coo(lay->v0->v0d, row, col, val, aval);
}
extern "C" gp_status gp_ad_copy(const gp_layout *lay, int64_t *row, int64_t *col, double *val) {
  if (!lay || !lay->v0) return fail(GP_ERR_STATE, "V0 basis not built");
  return copy_coo(lay->v0->ad, row, col, val, nullptr);
}
extern "C" gp_status gp_v0c_copy(const gp_layout *lay, int64_t *row, int64_t *col, double *val, double *aval) {
  if (!lay || !lay->v0) return fail(GP_ERR_STATE, "V0 basis not built");
  return copy_coo(lay->v0->v0c, row, col, val, aval);
}
extern "C" gp_status gp_ac_copy(const gp_layout *lay, int64_t *row, int64_t *col, double *val) {
  if (!lay || !lay->v0) return fail(GP_ERR_STATE, "V0 basis not built");
  return copy_coo(lay->v0->ac, row, col, val, nullptr);
}

// aux.hip -- auxiliary nodal space of the full-wave preconditioner.
//
// The space is spanned by the columns of Gx = [V0d | V0c] (the reference's V0
// basis, src/fdtd.hpp:1823-3479): gradients of dielectric node potentials, one
// aggregate per non-reference conductor, and gradients of conductor node
// potentials with one node per floating/excited conductor omitted.  V0d lives
// on dielectric edges only, so conductor potentials never appear as
// differences of large floating values, and together the columns span every
// free node potential.  The Galerkin operator
//   N(w) = Gx^T K(w) Gx = -w^2 Gx^T W eps Gx + i w Gx^T W sigma Gx
// has a contrast of ~1e7-1e10 between conductor and dielectric couplings, so
// its AMG aggregates never mix classes: dielectric nodes, the nodes of each
// conductor, and the conductor aggregates (by centroid) are blocked 2x2x2
// separately; the tentative prolongation is smoothed once with the
// class-filtered operator (smoothed aggregation).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <complex>

#include "mg.hpp"
#include "v0.hpp"

namespace gp {

namespace {

constexpr int32_t kSuperClass = 1 << 14;
// Entries of the smoothed prolongation below this fraction of their row's
// largest are dropped (0: none).  Truncation thins the coarse operators (at
// 200x200x24, level 2 of 4.28M entries falls to 2.25M / 1.41M / 0.59M at
// 0.05 / 0.1 / 0.2) but weakens the correction: the mixed batched solve ran
// 22.6 / 21.9 / 12.1 solves/s against 23.5 untruncated (median iterations 48 /
// 47 / 80 against 43), so none is applied.
constexpr double kAuxTrunc = 0.0;

__global__ void k_aux_triplets(int64_t n, const int64_t *__restrict__ row, const int64_t *__restrict__ col,
                               const double *__restrict__ val, const int64_t *__restrict__ map_edge, int64_t coff,
                               int64_t *__restrict__ r, int64_t *__restrict__ c, double *__restrict__ v,
                               int64_t *__restrict__ bad) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t m = map_edge[row[i]];
  if (m < 0) { atomicAdd((unsigned long long *)bad, 1ULL); r[i] = 0; c[i] = coff + col[i]; v[i] = 0.; return; }
  r[i] = m; c[i] = coff + col[i]; v[i] = val[i];
}

__global__ void k_rowscale2(int64_t nr, const int64_t *__restrict__ rp, const double *__restrict__ g,
                            const double *__restrict__ W, const double *__restrict__ epsr, const double *__restrict__ sig,
                            double *__restrict__ o0, double *__restrict__ o1) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  const double a = W[i] * epsr[i] * kEps0, b = W[i] * sig[i];
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) { o0[p] = a * g[p]; o1[p] = b * g[p]; }
}

// class and integer coordinates of every auxiliary column
__global__ void k_aux_class(int64_t naux, int64_t ndiel, int64_t nd, const int64_t *__restrict__ dnode,
                            const int64_t *__restrict__ cnode, const int32_t *__restrict__ mark_node, Grid g,
                            int32_t *__restrict__ cls, int32_t *__restrict__ crd) {
  int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= naux) return;
  int64_t v = -1;
  int32_t c = 0;
  if (a < ndiel) { v = dnode[a]; c = 0; }
  else if (a < nd) { v = -1; c = kSuperClass; }
  else { v = cnode[a - nd]; c = mark_node[v]; }
  cls[a] = c;
  if (v >= 0) {
    crd[3 * a] = (int32_t)(v / g.Nns);
    crd[3 * a + 1] = (int32_t)((v % g.Nns) / (g.Ny + 1));
    crd[3 * a + 2] = (int32_t)((v % g.Nns) % (g.Ny + 1));
  }
}

// centroid of each conductor aggregate column (super node) from its V0d entries
__global__ void k_super_centroid(int64_t nent, const int64_t *__restrict__ row, const int64_t *__restrict__ col,
                                 int64_t ndiel, Grid g, unsigned long long *__restrict__ acc) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nent) return;
  const int64_t c = col[i];
  if (c < ndiel) return;
  int d; int64_t iz, ix, iy;
  g.decode(row[i], d, iz, ix, iy);
  unsigned long long *p = acc + 4 * (c - ndiel);
  atomicAdd(p, (unsigned long long)iz);
  atomicAdd(p + 1, (unsigned long long)ix);
  atomicAdd(p + 2, (unsigned long long)iy);
  atomicAdd(p + 3, 1ULL);
}

__global__ void k_super_crd(int64_t nsup, int64_t ndiel, const unsigned long long *__restrict__ acc,
                            int32_t *__restrict__ crd) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsup) return;
  const unsigned long long n = acc[4 * s + 3] ? acc[4 * s + 3] : 1;
  for (int k = 0; k < 3; ++k) crd[3 * (ndiel + s) + k] = (int32_t)(acc[4 * s + k] / n);
}

__global__ void k_agg_keys(int64_t n, const int32_t *__restrict__ cls, const int32_t *__restrict__ crd,
                           uint64_t *__restrict__ key, int64_t *__restrict__ idx) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t c = (uint64_t)cls[i];
  key[i] = (c << 48) | ((uint64_t)(crd[3 * i] >> 1) << 32) | ((uint64_t)(crd[3 * i + 1] >> 1) << 16) |
           (uint64_t)(crd[3 * i + 2] >> 1);
  idx[i] = i;
}

__global__ void k_heads64(int64_t n, const uint64_t *__restrict__ key, int64_t *__restrict__ head) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) head[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

__global__ void k_agg_assign(int64_t n, const int64_t *__restrict__ sidx, const int64_t *__restrict__ hscan,
                             const int64_t *__restrict__ head, const uint64_t *__restrict__ skey,
                             int64_t *__restrict__ agg, int32_t *__restrict__ ccls, int32_t *__restrict__ ccrd) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  // hscan is the exclusive scan of head: the aggregate of sorted position p is
  // hscan[p] when p heads a run, else the aggregate of the run's head = hscan[p] - 1
  const int64_t g = head[p] ? hscan[p] : hscan[p] - 1;
  agg[sidx[p]] = g;
  if (head[p]) {
    const uint64_t k = skey[p];
    ccls[g] = (int32_t)(k >> 48);
    ccrd[3 * g] = (int32_t)((k >> 32) & 0xffff);
    ccrd[3 * g + 1] = (int32_t)((k >> 16) & 0xffff);
    ccrd[3 * g + 2] = (int32_t)(k & 0xffff);
  }
}

// Connected components inside each box: label propagation over the couplings
// of the class block (N_eps between dielectric nodes, N_sig within a
// conductor) restricted to pairs in the same box; super nodes stay alone.
// A box may hold disconnected pieces of a conductor (parallel wires one cell
// apart), which one constant would tie together and so spoil the coarse space.
__global__ void k_cc_init(int64_t n, int64_t *__restrict__ lab) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) lab[i] = i;
}

__global__ void k_cc_prop(int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                          const double *__restrict__ ne, const double *__restrict__ ns, const int32_t *__restrict__ cls,
                          const int64_t *__restrict__ box, const int64_t *__restrict__ lin, int64_t *__restrict__ lout,
                          int *__restrict__ changed) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t c = cls[i];
  int64_t best = lin[i];
  if (c != kSuperClass)
    for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
      const int32_t j = ci[p];
      if (j == (int32_t)i || box[j] != box[i] || cls[j] != c) continue;
      const double a = c == 0 ? ne[p] : ns[p];
      if (a != 0. && lin[j] < best) best = lin[j];
    }
  lout[i] = best;
  if (best != lin[i]) atomicAdd(changed, 1);
}

__global__ void k_cc_keys(int64_t n, const int64_t *__restrict__ lab, uint64_t *__restrict__ key, int64_t *__restrict__ idx) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key[i] = (uint64_t)lab[i];
  idx[i] = i;
}

// aggregate of each node from the sorted component labels; the coarse node
// takes the class of its component and the box coordinates (super nodes keep
// their own coordinates)
__global__ void k_cc_assign(int64_t n, const int64_t *__restrict__ sidx, const int64_t *__restrict__ hscan,
                            const int64_t *__restrict__ head, const uint64_t *__restrict__ skey,
                            const int32_t *__restrict__ cls, const int32_t *__restrict__ crd,
                            int64_t *__restrict__ agg, int32_t *__restrict__ ccls, int32_t *__restrict__ ccrd) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int64_t g = head[p] ? hscan[p] : hscan[p] - 1;
  agg[sidx[p]] = g;
  if (head[p]) {
    const int64_t lab = (int64_t)skey[p];
    const int32_t c = cls[lab];
    ccls[g] = c;
    const int sh = c == kSuperClass ? 0 : 1;
    for (int d = 0; d < 3; ++d) ccrd[3 * g + d] = crd[3 * lab + d] >> sh;
  }
}

__global__ void k_ptent(int64_t n, const int64_t *__restrict__ agg, int64_t *__restrict__ rp, int32_t *__restrict__ ci,
                        double *__restrict__ v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  rp[i] = i;
  if (i < n) { ci[i] = (int32_t)agg[i]; v[i] = 1.0; }
}

// class-filtered smoothing operator: omega/D_F * A_F, A_F from N_eps (dielectric
// pairs) or N_sig (same-conductor pairs); super nodes and cross-class pairs dropped
__global__ void k_filter(int64_t nr, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                         const double *__restrict__ ne, const double *__restrict__ ns, const int32_t *__restrict__ cls,
                         double *__restrict__ out, double *__restrict__ dF) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  const int32_t c = cls[i];
  double d = 0.;
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
    const int32_t j = ci[p];
    double a = 0.;
    if (c != kSuperClass && cls[j] == c) a = c == 0 ? ne[p] : ns[p];
    out[p] = a;
    if (j == (int32_t)i) d = a;
  }
  dF[i] = d > 0 ? d : 0.;
}

__global__ void k_scale_rows_inv(int64_t nr, const int64_t *__restrict__ rp, const double *__restrict__ dF, double s,
                                 double *__restrict__ v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  const double f = dF[i] > 0 ? s / dF[i] : 0.;
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) v[p] *= f;
}

// power iteration step for lambda_max(D^-1 A): y = D^-1 A x, returns partial norms
// y = D^-1 A x, one wavefront per row: the conductor aggregate rows hold up
// to ~1e5 entries, which one thread per row walked alone (~10 ms a launch)
__global__ void k_dinv_spmv(int64_t nr, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                            const double *__restrict__ v, const double *__restrict__ dF, const double *__restrict__ x,
                            double *__restrict__ y) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (i >= nr) return;
  double s = 0.;
  for (int64_t p = rp[i] + lane; p < rp[i + 1]; p += 64) s += v[p] * x[ci[p]];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) y[i] = dF[i] > 0 ? s / dF[i] : 0.;
}

__global__ void k_fill_hash(int64_t n, double *__restrict__ x) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.5 + 0.5 * (double)((i * 2654435761ULL) % 1000) / 1000.0;
}

double dev_norm(const double *x, int64_t n, hipStream_t s) {
  std::vector<double> h(n);
  (void)hipMemcpyAsync(h.data(), x, sizeof(double) * n, hipMemcpyDeviceToHost, s);
  (void)hipStreamSynchronize(s);
  double a = 0;
  for (double v : h) a += v * v;
  return std::sqrt(a);
}

__global__ void k_scal(int64_t n, double f, double *__restrict__ x) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= f;
}

gp_status dense_host_aux(const Csr &M, int which, int64_t n, std::vector<double> *out) {
  HostCsr h;
  GP_TRY(csr_host_copy(M, &h, which));
  out->assign(n * n, 0.0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t p = h.rp[i]; p < h.rp[i + 1]; ++p) (*out)[i * n + h.ci[p]] += h.v[p];
  return GP_OK;
}

}  // namespace

gp_status build_aux(gp_layout *lay, Arena &A, Hierarchy *H) {
  GP_TRY(build_v0(lay));
  const V0Data &V = *lay->v0;
  hipStream_t s = lay->stream;
  AuxAMG &X = H->aux;
  const int64_t n = lay->n;
  const int64_t nd = V.v0d_cols, nc = V.v0c_cols, naux = nd + nc;
  X.naux = naux;
  X.ndiel = V.ndiel;
  Arena tmp;
  gp_status st = GP_OK;
  do {
    // Gx = [V0d | V0c] on the kept edges
    const int64_t ne = V.v0d.n + V.v0c.n;
    int64_t *r = nullptr, *c = nullptr, *bad = nullptr;
    double *v = nullptr;
    if ((st = tmp.alloc(&r, ne)) != GP_OK) break;
    if ((st = tmp.alloc(&c, ne)) != GP_OK) break;
    if ((st = tmp.alloc(&v, ne)) != GP_OK) break;
    if ((st = tmp.alloc(&bad, 1)) != GP_OK) break;
    if (hipMemsetAsync(bad, 0, sizeof(int64_t), s) != hipSuccess) { st = fail(GP_ERR_HIP, "memset"); break; }
    k_aux_triplets<<<blocks_for(V.v0d.n), 256, 0, s>>>(V.v0d.n, V.v0d.row, V.v0d.col, V.v0d.val, lay->map_edge, 0, r, c, v, bad);
    if (V.v0c.n)
      k_aux_triplets<<<blocks_for(V.v0c.n), 256, 0, s>>>(V.v0c.n, V.v0c.row, V.v0c.col, V.v0c.val, lay->map_edge, nd,
                                                         r + V.v0d.n, c + V.v0d.n, v + V.v0d.n, bad);
    const double *vals[1] = {v};
    if ((st = csr_from_coo(A, n, naux, ne, r, c, vals, 1, &X.Gx, s)) != GP_OK) break;
    if ((st = transpose(A, X.Gx, &X.GxT, s)) != GP_OK) break;
    if ((st = csr_alloc(A, n, naux, X.Gx.nnz, 2, &X.Bx)) != GP_OK) break;
    if (hipMemcpyAsync(X.Bx.rp, X.Gx.rp, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipMemcpyAsync(X.Bx.ci, X.Gx.ci, sizeof(int32_t) * X.Gx.nnz, hipMemcpyDeviceToDevice, s) != hipSuccess) { st = fail(GP_ERR_HIP, "copy"); break; }
    k_rowscale2<<<blocks_for(n), 256, 0, s>>>(n, X.Gx.rp, X.Gx.v[0], H->W, H->epsr, H->sig, X.Bx.v[0], X.Bx.v[1]);
    X.L.clear();
    X.L.emplace_back();
    X.L[0].n = naux;
    if ((st = spgemm(A, X.GxT, X.Bx, &X.L[0].N, s)) != GP_OK) break;
    // classes and coordinates
    int32_t *cls = nullptr, *crd = nullptr;
    if ((st = tmp.alloc(&cls, naux)) != GP_OK) break;
    if ((st = tmp.alloc(&crd, 3 * naux)) != GP_OK) break;
    k_aux_class<<<blocks_for(naux), 256, 0, s>>>(naux, V.ndiel, nd, V.dnode_of_col, V.cnode_of_col, lay->mark_node, lay->g, cls, crd);
    const int64_t nsup = nd - V.ndiel;
    if (nsup > 0) {
      unsigned long long *acc = nullptr;
      if ((st = tmp.alloc(&acc, 4 * nsup)) != GP_OK) break;
      if (hipMemsetAsync(acc, 0, sizeof(unsigned long long) * 4 * nsup, s) != hipSuccess) { st = fail(GP_ERR_HIP, "memset"); break; }
      k_super_centroid<<<blocks_for(V.v0d.n), 256, 0, s>>>(V.v0d.n, V.v0d.row, V.v0d.col, V.ndiel, lay->g, acc);
      k_super_crd<<<blocks_for(nsup), 256, 0, s>>>(nsup, V.ndiel, acc, crd);
    }
    // aggregation levels
    while (true) {
      AuxLevel &F = X.L.back();
      const int64_t nf = F.n;
      if (nf <= 256 || X.L.size() >= 16) break;
      uint64_t *key = nullptr, *key2 = nullptr;
      int64_t *idx = nullptr, *idx2 = nullptr, *head = nullptr, *hscan = nullptr, *agg = nullptr;
      if ((st = tmp.alloc(&key, nf)) != GP_OK) break;
      if ((st = tmp.alloc(&key2, nf)) != GP_OK) break;
      if ((st = tmp.alloc(&idx, nf)) != GP_OK) break;
      if ((st = tmp.alloc(&idx2, nf)) != GP_OK) break;
      if ((st = tmp.alloc(&head, nf)) != GP_OK) break;
      if ((st = tmp.alloc(&hscan, nf + 1)) != GP_OK) break;
      if ((st = tmp.alloc(&agg, nf)) != GP_OK) break;
      k_agg_keys<<<blocks_for(nf), 256, 0, s>>>(nf, cls, crd, key, idx);
      size_t bytes = 0;
      if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, key, key2, idx, idx2, (int)nf, 0, 64, s) != hipSuccess) { st = fail(GP_ERR_HIP, "sort"); break; }
      uint8_t *buf = nullptr;
      if ((st = tmp.alloc(&buf, (int64_t)bytes + 16)) != GP_OK) break;
      if (hipcub::DeviceRadixSort::SortPairs(buf, bytes, key, key2, idx, idx2, (int)nf, 0, 64, s) != hipSuccess) { st = fail(GP_ERR_HIP, "sort"); break; }
      k_heads64<<<blocks_for(nf), 256, 0, s>>>(nf, key2, head);
      if ((st = exclusive_scan_i64(head, hscan, nf, s, tmp)) != GP_OK) break;
      int64_t nbox = 0;
      if (hipMemcpyAsync(&nbox, hscan + nf, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "count"); break; }
      int32_t *bcls = nullptr, *bcrd = nullptr;
      if ((st = tmp.alloc(&bcls, nbox)) != GP_OK) break;
      if ((st = tmp.alloc(&bcrd, 3 * nbox)) != GP_OK) break;
      k_agg_assign<<<blocks_for(nf), 256, 0, s>>>(nf, idx2, hscan, head, key2, agg, bcls, bcrd);
      // split every box into its connected components
      int64_t *la = nullptr, *lb = nullptr;
      int *chg = nullptr;
      if ((st = tmp.alloc(&la, nf)) != GP_OK) break;
      if ((st = tmp.alloc(&lb, nf)) != GP_OK) break;
      if ((st = tmp.alloc(&chg, 1)) != GP_OK) break;
      k_cc_init<<<blocks_for(nf), 256, 0, s>>>(nf, la);
      for (int it = 0; it < 64; ++it) {
        int hc = 0;
        if (hipMemsetAsync(chg, 0, sizeof(int), s) != hipSuccess) { st = fail(GP_ERR_HIP, "memset"); break; }
        k_cc_prop<<<blocks_for(nf), 256, 0, s>>>(nf, F.N.rp, F.N.ci, F.N.v[0], F.N.v[1], cls, agg, la, lb, chg);
        std::swap(la, lb);
        if (hipMemcpyAsync(&hc, chg, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "labels"); break; }
        if (hc == 0) break;
      }
      if (st != GP_OK) break;
      k_cc_keys<<<blocks_for(nf), 256, 0, s>>>(nf, la, key, idx);
      if (hipcub::DeviceRadixSort::SortPairs(buf, bytes, key, key2, idx, idx2, (int)nf, 0, 64, s) != hipSuccess) { st = fail(GP_ERR_HIP, "sort"); break; }
      k_heads64<<<blocks_for(nf), 256, 0, s>>>(nf, key2, head);
      if ((st = exclusive_scan_i64(head, hscan, nf, s, tmp)) != GP_OK) break;
      int64_t ncoarse = 0;
      if (hipMemcpyAsync(&ncoarse, hscan + nf, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "count"); break; }
      if (ncoarse > (int64_t)(0.85 * (double)nf)) break;
      int32_t *ccls = nullptr, *ccrd = nullptr;
      if ((st = tmp.alloc(&ccls, ncoarse)) != GP_OK) break;
      if ((st = tmp.alloc(&ccrd, 3 * ncoarse)) != GP_OK) break;
      k_cc_assign<<<blocks_for(nf), 256, 0, s>>>(nf, idx2, hscan, head, key2, cls, crd, agg, ccls, ccrd);
      Csr Pt;
      if ((st = csr_alloc(tmp, nf, ncoarse, nf, 1, &Pt)) != GP_OK) break;
      k_ptent<<<blocks_for(nf + 1), 256, 0, s>>>(nf, agg, Pt.rp, Pt.ci, Pt.v[0]);
      // smoothed prolongation P = Pt - 4/(3 lambda) D_F^-1 A_F Pt
      Csr AF;
      double *dF = nullptr;
      if ((st = csr_alloc(tmp, nf, nf, F.N.nnz, 1, &AF)) != GP_OK) break;
      if ((st = tmp.alloc(&dF, nf)) != GP_OK) break;
      if (hipMemcpyAsync(AF.rp, F.N.rp, sizeof(int64_t) * (nf + 1), hipMemcpyDeviceToDevice, s) != hipSuccess ||
          hipMemcpyAsync(AF.ci, F.N.ci, sizeof(int32_t) * F.N.nnz, hipMemcpyDeviceToDevice, s) != hipSuccess) { st = fail(GP_ERR_HIP, "copy"); break; }
      k_filter<<<blocks_for(nf), 256, 0, s>>>(nf, F.N.rp, F.N.ci, F.N.v[0], F.N.v[1], cls, AF.v[0], dF);
      double *x = nullptr, *y = nullptr;
      if ((st = tmp.alloc(&x, nf)) != GP_OK) break;
      if ((st = tmp.alloc(&y, nf)) != GP_OK) break;
      k_fill_hash<<<blocks_for(nf), 256, 0, s>>>(nf, x);
      double lam = 1.0;
      for (int it = 0; it < 12; ++it) {
        k_dinv_spmv<<<blocks_for(nf * 64), 256, 0, s>>>(nf, AF.rp, AF.ci, AF.v[0], dF, x, y);
        lam = dev_norm(y, nf, s) / std::max(dev_norm(x, nf, s), 1e-300);
        k_scal<<<blocks_for(nf), 256, 0, s>>>(nf, 1.0 / std::max(dev_norm(y, nf, s), 1e-300), y);
        std::swap(x, y);
      }
      lam = std::max(lam, 1e-3);
      k_scale_rows_inv<<<blocks_for(nf), 256, 0, s>>>(nf, AF.rp, dF, 4.0 / (3.0 * lam), AF.v[0]);
      Csr AFc, SP;
      if ((st = csr_drop_zeros(tmp, AF, &AFc, s)) != GP_OK) break;
      if ((st = spgemm(tmp, AFc, Pt, &SP, s)) != GP_OK) break;
      if (kAuxTrunc > 0) {
        Csr Ps;
        if ((st = csr_add(tmp, Pt, -1.0, SP, &Ps, s)) != GP_OK) break;
        if ((st = truncate_prolongation(A, tmp, Ps, Pt, kAuxTrunc, &F.P, s)) != GP_OK) break;
      } else if ((st = csr_add(A, Pt, -1.0, SP, &F.P, s)) != GP_OK) {
        break;
      }
      if ((st = transpose(A, F.P, &F.PT, s)) != GP_OK) break;
      AuxLevel Cl;
      Cl.n = ncoarse;
      Csr NP;
      if ((st = spgemm(tmp, F.N, F.P, &NP, s)) != GP_OK) break;
      if ((st = spgemm(A, F.PT, NP, &Cl.N, s)) != GP_OK) break;
      X.L.push_back(Cl);
      cls = ccls;
      crd = ccrd;
    }
    if (st != GP_OK) break;
    for (AuxLevel &L : X.L) GP_TRY(aux_smoother_data(A, L.N, L.nd, L.nl, s));
    GP_TRY(csr_mark_long(A, X.GxT, s));
    for (AuxLevel &L : X.L) {
      GP_TRY(csr_mark_long(A, L.N, s));
      if (L.P.nr) { GP_TRY(csr_mark_long(A, L.P, s)); GP_TRY(csr_mark_long(A, L.PT, s)); }
    }
    // SELL copies for the cycle kernels (long rows stay with the long-row pass)
    GP_TRY(sell_from_csr(A, X.GxT, &X.GxTs, s, "X.GxT"));
    GP_TRY(sell_from_csr(A, X.Bx, &X.Bxs, s, "X.Bx", H->ord_e));  // rows: level-0 edges
    for (size_t l = 0; l + 1 < X.L.size(); ++l) {
      AuxLevel &L = X.L[l];
      GP_TRY(sell_from_csr(A, L.N, &L.Ns, s, "L.N"));
      GP_TRY(sell_from_csr(A, L.P, &L.Ps, s, "L.P"));
      GP_TRY(sell_from_csr(A, L.PT, &L.PTs, s, "L.PT"));
    }
    const AuxLevel &C = X.L.back();
    X.nL = C.n;
    for (int k = 0; k < 2; ++k) GP_TRY(dense_host_aux(C.N, k, C.n, &X.NL[k]));
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) { st = fail(GP_ERR_HIP, "aux setup kernels"); break; }
  } while (0);
  (void)hipStreamSynchronize(s);
  tmp.free_all();
  return st;
}


// FP32 copies of the auxiliary hierarchy's operators (see hierarchy_add_f32).
gp_status aux_add_f32(Arena &A, Hierarchy *H, hipStream_t s) {
  AuxAMG &X = H->aux;
  GP_TRY(csr_to_float(A, X.Gx, &X.Gxf, s));
  GP_TRY(csr_to_float(A, X.GxT, &X.GxTf, s));
  GP_TRY(csr_to_float(A, X.Bx, &X.Bxf, s));
  GP_TRY(sell_to_float(A, X.GxTs, &X.GxTsf, s));
  GP_TRY(sell_to_float(A, X.Bxs, &X.Bxsf, s));
  for (size_t l = 0; l + 1 < X.L.size(); ++l) {
    AuxLevel &L = X.L[l];
    GP_TRY(sell_to_float(A, L.Ns, &L.Nsf, s));
    GP_TRY(sell_to_float(A, L.Ps, &L.Psf, s));
    GP_TRY(sell_to_float(A, L.PTs, &L.PTsf, s));
  }
  for (AuxLevel &L : X.L) {
    GP_TRY(csr_to_float(A, L.N, &L.Nf, s));
    if (L.P.nr) { GP_TRY(csr_to_float(A, L.P, &L.Pf, s)); GP_TRY(csr_to_float(A, L.PT, &L.PTf, s)); }
    for (int k = 0; k < 2; ++k) {
      GP_TRY(A.alloc(&L.ndf[k], L.n)); GP_TRY(A.alloc(&L.nlf[k], L.n));
    }
    for (int k = 0; k < 2; ++k) { GP_TRY(d2f_copy(L.nd[k], L.ndf[k], L.n, s)); GP_TRY(d2f_copy(L.nl[k], L.nlf[k], L.n, s)); }
  }
  return GP_OK;
}

}  // namespace gp

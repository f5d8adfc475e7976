// assemble.hip -- curl-curl stiffness assembly S = Sh * Se^T / MU on the
// structured Yee mesh, replacing generateStiff + mklMatrixMulti_nt
// (src/generateStiff.cpp:6-565, 871-980).
//
// One thread per kept edge (row).  The thread enumerates the <= 4 faces the
// edge bounds in ascending face-id order (the reference's Se/Sh column order),
// forms Sh(row, f) * Se(col, f) for the 4 edges of each face with the same
// floating-point expressions the reference writes, adds repeated columns (only
// the diagonal repeats) in face order, drops PEC columns, renumbers through
// mapEdge and divides by MU.  Two passes: row lengths -> scan -> fill.
// Built with FP contraction off: every product and sum rounds as written.
#pragma clang fp contract(off)

#include "gp_dev.hpp"

using namespace gp;

namespace {

struct Geo {
  const double *xn, *yn, *zn;
  Grid g;
};

// dual widths dxa/dya/dza (src/generateStiff.cpp:250-267)
__device__ __forceinline__ double dual(const double *c, int64_t i, int64_t N) {
  if (i == 0) return c[1] - c[0];
  if (i == N) return c[N] - c[N - 1];
  return (c[i + 1] - c[i - 1]) / 2;
}

struct Inc { int64_t e; double se, sh; };

// The four incidences of a face in the reference's emission order with their
// Se (primal) and Sh (dual) values.  type 0: xy face at plane a, cell (b, c);
// type 1: yz face at x = a, y-cell b, z-cell c; type 2: xz face at y = a,
// x-cell b, z-cell c.
__device__ void face_incidences(const Geo &G, int type, int64_t a, int64_t b, int64_t c, Inc inc[4]) {
  const Grid &g = G.g;
  if (type == 0) {
    const int64_t pz = a, cx = b, cy = c;
    const double P = dual(G.zn, pz < g.Nz ? pz : g.Nz - 1, g.Nz);  // top face reuses dza[Nz-1]
    const double hx = G.xn[cx + 1] - G.xn[cx], hy = G.yn[cy + 1] - G.yn[cy];
    inc[0] = {g.ey(pz, cx, cy), -1 / hx, -P / (dual(G.xn, cx, g.Nx) * P)};
    inc[1] = {g.ey(pz, cx + 1, cy), 1 / hx, P / (dual(G.xn, cx + 1, g.Nx) * P)};
    inc[2] = {g.ex(pz, cx, cy), 1 / hy, P / (P * dual(G.yn, cy, g.Ny))};
    inc[3] = {g.ex(pz, cx, cy + 1), -1 / hy, -P / (P * dual(G.yn, cy + 1, g.Ny))};
  } else if (type == 1) {
    const int64_t fx = a, cy = b, cz = c;
    const double P = dual(G.xn, fx, g.Nx);
    const double hy = G.yn[cy + 1] - G.yn[cy], hz = G.zn[cz + 1] - G.zn[cz];
    inc[0] = {g.ey(cz, fx, cy), 1 / hz, P / (P * dual(G.zn, cz, g.Nz))};
    inc[1] = {g.ez(cz, fx, cy), -1 / hy, -P / (P * dual(G.yn, cy, g.Ny))};
    inc[2] = {g.ez(cz, fx, cy + 1), 1 / hy, P / (P * dual(G.yn, cy + 1, g.Ny))};
    inc[3] = {g.ey(cz + 1, fx, cy), -1 / hz, -P / (P * dual(G.zn, cz + 1, g.Nz))};
  } else {
    const int64_t fy = a, cx = b, cz = c;
    const double P = dual(G.yn, fy, g.Ny);
    const double hx = G.xn[cx + 1] - G.xn[cx], hz = G.zn[cz + 1] - G.zn[cz];
    inc[0] = {g.ex(cz, cx, fy), -1 / hz, -P / (dual(G.zn, cz, g.Nz) * P)};
    inc[1] = {g.ez(cz, cx, fy), 1 / hx, P / (dual(G.xn, cx, g.Nx) * P)};
    inc[2] = {g.ez(cz, cx + 1, fy), -1 / hx, -P / (P * dual(G.xn, cx + 1, g.Nx))};
    inc[3] = {g.ex(cz + 1, cx, fy), 1 / hz, P / (P * dual(G.zn, cz + 1, g.Nz))};
  }
}

struct FaceRef { int64_t id; int type; int64_t a, b, c; int role; };

// Faces bounded by edge e, sorted by face id; role = e's incidence index.
__device__ int edge_faces(const Grid &g, int64_t e, FaceRef f[4]) {
  int d; int64_t iz, ix, iy;
  g.decode(e, d, iz, ix, iy);
  int n = 0;
  if (d == 0) {
    if (ix >= 1) f[n++] = {g.f_xy(iz, ix - 1, iy), 0, iz, ix - 1, iy, 1};
    if (ix < g.Nx) f[n++] = {g.f_xy(iz, ix, iy), 0, iz, ix, iy, 0};
    if (iz >= 1) f[n++] = {g.f_yz(iz - 1, ix, iy), 1, ix, iy, iz - 1, 3};
    if (iz < g.Nz) f[n++] = {g.f_yz(iz, ix, iy), 1, ix, iy, iz, 0};
  } else if (d == 1) {
    if (iy >= 1) f[n++] = {g.f_xy(iz, ix, iy - 1), 0, iz, ix, iy - 1, 3};
    if (iy < g.Ny) f[n++] = {g.f_xy(iz, ix, iy), 0, iz, ix, iy, 2};
    if (iz >= 1) f[n++] = {g.f_xz(iz - 1, ix, iy), 2, iy, ix, iz - 1, 3};
    if (iz < g.Nz) f[n++] = {g.f_xz(iz, ix, iy), 2, iy, ix, iz, 0};
  } else {
    if (iy >= 1) f[n++] = {g.f_yz(iz, ix, iy - 1), 1, ix, iy - 1, iz, 2};
    if (iy < g.Ny) f[n++] = {g.f_yz(iz, ix, iy), 1, ix, iy, iz, 1};
    if (ix >= 1) f[n++] = {g.f_xz(iz, ix - 1, iy), 2, iy, ix - 1, iz, 2};
    if (ix < g.Nx) f[n++] = {g.f_xz(iz, ix, iy), 2, iy, ix, iz, 1};
  }
  for (int i = 1; i < n; ++i) {  // insertion sort by face id
    FaceRef t = f[i];
    int j = i - 1;
    while (j >= 0 && f[j].id > t.id) { f[j + 1] = f[j]; --j; }
    f[j + 1] = t;
  }
  return n;
}

__global__ void k_row_count(Geo G, const int64_t *__restrict__ mapr, const uint8_t *__restrict__ rm, int64_t n,
                            int64_t *__restrict__ cnt) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t e = mapr[r];
  FaceRef f[4];
  int nf = edge_faces(G.g, e, f);
  int64_t c = 1;
  for (int q = 0; q < nf; ++q) {
    Inc inc[4];
    face_incidences(G, f[q].type, f[q].a, f[q].b, f[q].c, inc);
    for (int k = 0; k < 4; ++k)
      if (k != f[q].role && !rm[inc[k].e]) ++c;
  }
  cnt[r] = c;
}

__global__ void k_row_fill(Geo G, const int64_t *__restrict__ mapr, const int64_t *__restrict__ map,
                           const uint8_t *__restrict__ rm, int64_t n, const int64_t *__restrict__ rp,
                           int32_t *__restrict__ ci, double *__restrict__ v) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t e = mapr[r];
  FaceRef f[4];
  int nf = edge_faces(G.g, e, f);
  int64_t col[16];
  double val[16];
  int m = 0;
  for (int q = 0; q < nf; ++q) {
    Inc inc[4];
    face_incidences(G, f[q].type, f[q].a, f[q].b, f[q].c, inc);
    const double sh = inc[f[q].role].sh;
    for (int k = 0; k < 4; ++k) {
      col[m] = inc[k].e;
      val[m] = sh * inc[k].se;
      ++m;
    }
  }
  // stable insertion sort by column: ties keep ascending face order
  for (int i = 1; i < m; ++i) {
    int64_t tc = col[i];
    double tv = val[i];
    int j = i - 1;
    while (j >= 0 && col[j] > tc) { col[j + 1] = col[j]; val[j + 1] = val[j]; --j; }
    col[j + 1] = tc;
    val[j + 1] = tv;
  }
  int64_t o = rp[r];
  for (int i = 0; i < m;) {
    int j = i;
    double s = 0.;
    while (j < m && col[j] == col[i]) { s += val[j]; ++j; }
    if (!rm[col[i]]) {
      ci[o] = (int32_t)map[col[i]];
      v[o] = s / kMu;
      ++o;
    }
    i = j;
  }
}

}  // namespace

extern "C" gp_status gp_assemble_stiffness(gp_layout *lay) {
  if (!lay) return fail(GP_ERR_ARG, "null layout");
  if (lay->s_ready) return GP_OK;
  const Grid &g = lay->g;
  if (g.Ne > 2147483647LL) return fail(GP_ERR_ARG, "more than 2^31 edges: int32 columns would overflow");
  hipStream_t s = lay->stream;
  Geo G{lay->xn, lay->yn, lay->zn, g};
  const int64_t n = lay->n;
  CsrD &S = lay->S;
  S.n = n;
  GP_TRY(lay->arena.alloc(&S.rp, n + 1));
  Arena tmp;
  int64_t *cnt = nullptr;
  gp_status st = tmp.alloc(&cnt, n);
  if (st == GP_OK) {
    k_row_count<<<blocks_for(n), 256, 0, s>>>(G, lay->map_edge_r, lay->rm_edge, n, cnt);
    st = exclusive_scan_i64(cnt, S.rp, n, s, tmp);
  }
  if (st == GP_OK) {
    if (hipMemcpyAsync(&S.nnz, S.rp + n, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      st = fail(GP_ERR_HIP, "reading nnz");
  }
  if (st == GP_OK) st = lay->arena.alloc(&S.ci, S.nnz);
  if (st == GP_OK) st = lay->arena.alloc(&S.v, S.nnz);
  if (st == GP_OK) {
    k_row_fill<<<blocks_for(n), 256, 0, s>>>(G, lay->map_edge_r, lay->map_edge, lay->rm_edge, n, S.rp, S.ci, S.v);
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) st = fail(GP_ERR_HIP, "assembly kernels");
  }
  tmp.free_all();
  if (st != GP_OK) return st;
  lay->s_ready = true;
  return GP_OK;
}

extern "C" gp_status gp_stiffness_copy_csr(const gp_layout *lay, int64_t *row_ptr, int32_t *col, double *val) {
  if (!lay || !lay->s_ready) return fail(GP_ERR_STATE, "stiffness not assembled");
  const CsrD &S = lay->S;
  if (row_ptr) GP_HIP(hipMemcpy(row_ptr, S.rp, sizeof(int64_t) * (S.n + 1), hipMemcpyDeviceToHost));
  if (col) GP_HIP(hipMemcpy(col, S.ci, sizeof(int32_t) * S.nnz, hipMemcpyDeviceToHost));
  if (val) GP_HIP(hipMemcpy(val, S.v, sizeof(double) * S.nnz, hipMemcpyDeviceToHost));
  return GP_OK;
}

extern "C" gp_status gp_stiffness_copy_coo(const gp_layout *lay, int64_t *row, int64_t *col, double *val) {
  // the reference's COO layout (SRowId/SColId/Sval): rows ascending, mapEdge numbering
  if (!lay || !lay->s_ready) return fail(GP_ERR_STATE, "stiffness not assembled");
  const CsrD &S = lay->S;
  std::vector<int64_t> rp(S.n + 1);
  std::vector<int32_t> ci(S.nnz);
  GP_HIP(hipMemcpy(rp.data(), S.rp, sizeof(int64_t) * (S.n + 1), hipMemcpyDeviceToHost));
  GP_HIP(hipMemcpy(ci.data(), S.ci, sizeof(int32_t) * S.nnz, hipMemcpyDeviceToHost));
  if (val) GP_HIP(hipMemcpy(val, S.v, sizeof(double) * S.nnz, hipMemcpyDeviceToHost));
  for (int64_t r = 0; r < S.n; ++r)
    for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
      if (row) row[k] = r;
      if (col) col[k] = ci[k];
    }
  return GP_OK;
}

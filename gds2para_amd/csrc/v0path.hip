// v0path.hip -- the V0 quasi-static network solve on the GPU: paraGenerator's
// V0 section (src/matrixCon.cpp:366-560), Construct_Z_V0
// (src/fdtd.hpp:4178-4215) and print_z_V0's store (src/fdtd.hpp:4249-4316).
//
// For each source port, with J the port's edge current (src/matrixCon.cpp:367-374),
// D_eps = stackEpsn[e / (Es+Ev)] EPSILON0 per original edge and w0 = 2 pi freqStart:
//   y0d  = Ad^-1 (-V0da^T J) / w0          (:377-403)
//   yd1  = V0d y0d                          (:411, :441)
//   y0c  = Ac^-1 (-V0ca^T J + V0ca^T (-w0 D_eps yd1))   (:434-511)
//   yc   = V0c y0c                          (:531)
//   y0d2 = Ad^-1 (V0da^T (-D_eps yc))       (:536-544)
//   yd   = V0d y0d2 - i yd1 + yc            (:552-587)
//   Z0[resp + np src] = -sum_{side 0 edges of resp} yd[e] len_e dir / sum area(src)
// The real solves (hypreSolve in the reference) run on the real AMG of
// real.hip, set up once per layout for Ad and once for Ac and kept with the
// layout; the V0 matrices are the unscaled values the reference's MKL handles
// hold (v0d1valo / v0d1avalo / v0cvalo / v0cavalo, :104-108, :316-322).
#pragma clang fp contract(off)

#include <chrono>
#include <cmath>
#include <functional>
#include <unordered_map>
#include <vector>

#include "real.hpp"
#include "v0.hpp"

namespace gp {

struct V0Path {
  Arena A;
  Csr Vd, VdT, Vc, VcT;            // v[0] = V0 values, v[1] = the test ("a") values
  RealAMG *ad = nullptr, *ac = nullptr;
  int smoother = GP_SMOOTH_CHEBYSHEV;  // of ad and ac
  int64_t nd = 0, nc = 0, ne = 0;
  double setup_ms = 0.;
  ~V0Path() {
    real_amg_destroy(ad);
    real_amg_destroy(ac);
    A.free_all();
  }
};

void destroy_v0path(V0Path *p) { delete p; }
int64_t v0path_bytes(const V0Path *p);

namespace {

// y = A x (mode 0), y = -(A x) (mode 1), y = y + A x (mode 2); one wavefront
// per row, so the long rows of V0d^T (the conductor aggregate columns of V0d)
// are split over 64 lanes
__global__ void k_mv_wave(int64_t nr, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                          const double *__restrict__ v, const double *__restrict__ x, double *__restrict__ y, int mode) {
  const int64_t row = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (row >= nr) return;
  double s = 0.;
  for (int64_t p = rp[row] + lane; p < rp[row + 1]; p += 64) s += v[p] * x[ci[p]];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) y[row] = mode == 0 ? s : mode == 1 ? -s : y[row] + s;
}

// y0d /= (2 pi freqStart freqUnit)
__global__ void k_div(int64_t n, double w, double *__restrict__ x) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] /= w;
}

// t[e] = yd1[e] * (-1.0 * w0 * stackEpsn[e / El] * EPSILON0)   (src/matrixCon.cpp:442)
__global__ void k_scale_eps_w(int64_t ne, int64_t El, const double *__restrict__ epsn, double w0,
                              const double *__restrict__ yd1, double *__restrict__ t) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < ne) t[e] = yd1[e] * (-1.0 * w0 * epsn[e / El] * kEps0);
}

// t[e] = -yc[e] * stackEpsn[e / El] * EPSILON0   (src/matrixCon.cpp:536-538)
__global__ void k_neg_eps(int64_t ne, int64_t El, const double *__restrict__ epsn, const double *__restrict__ yc,
                          double *__restrict__ t) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < ne) t[e] = -yc[e] * epsn[e / El] * kEps0;
}

__global__ void k_scatter(int64_t n, const int64_t *__restrict__ idx, const double *__restrict__ val,
                          double *__restrict__ J) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) J[idx[i]] = val[i];
}

// yd = (yd2 - i yd1) + yc at the response edges only
__global__ void k_gather_yd(int64_t n, const int64_t *__restrict__ idx, const double *__restrict__ yd2,
                            const double *__restrict__ yd1, const double *__restrict__ yc, double *__restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t e = idx[i];
  out[2 * i] = (yd2[e] - 0. * yd1[e]) + (yc ? yc[e] : 0.);
  out[2 * i + 1] = -yd1[e];
}

unsigned wave_blocks(int64_t nr) { return (unsigned)((nr + 3) / 4); }

gp_status setup(gp_layout *lay, V0Path **out) {
  auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = lay->stream;
  const V0Data &V = *lay->v0;
  V0Path *P = new V0Path();
  P->nd = V.v0d_cols;
  P->nc = V.v0c_cols;
  P->ne = lay->g.Ne;
  P->smoother = lay->v0_smoother;
  gp_status st = GP_OK;
  do {
    const double *dv[2] = {V.v0d.val, V.v0d.aval};
    if ((st = csr_from_coo(P->A, P->ne, P->nd, V.v0d.n, V.v0d.row, V.v0d.col, dv, 2, &P->Vd, s)) != GP_OK) break;
    if ((st = csr_from_coo(P->A, P->nd, P->ne, V.v0d.n, V.v0d.col, V.v0d.row, dv, 2, &P->VdT, s)) != GP_OK) break;
    if (P->nc > 0) {
      const double *cv[2] = {V.v0c.val, V.v0c.aval};
      if ((st = csr_from_coo(P->A, P->ne, P->nc, V.v0c.n, V.v0c.row, V.v0c.col, cv, 2, &P->Vc, s)) != GP_OK) break;
      if ((st = csr_from_coo(P->A, P->nc, P->ne, V.v0c.n, V.v0c.col, V.v0c.row, cv, 2, &P->VcT, s)) != GP_OK) break;
    }
    // Ad: the kept entries only; the reference's trailing (0, 0, 0) slots are
    // never read by hypreSolve's row-by-row insertion (src/hypreSolve.cpp:47-59)
    if ((st = real_amg_create(P->nd, V.ad.n, V.ad.row, V.ad.col, V.ad.val, P->smoother, &P->ad, s)) != GP_OK) break;
    if (P->nc > 0 && (st = real_amg_create(P->nc, V.ac.n, V.ac.row, V.ac.col, V.ac.val, P->smoother, &P->ac, s)) != GP_OK)
      break;
    if (hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "V0 path setup"); break; }
  } while (0);
  if (st != GP_OK) { delete P; return st; }
  P->setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  *out = P;
  return GP_OK;
}

double edge_len(const gp_mesh &m, const Grid &g, int64_t e) {
  int dir; int64_t iz, ix, iy;
  g.decode(e, dir, iz, ix, iy);
  return dir == 2 ? m.zn[iz + 1] - m.zn[iz] : dir == 1 ? m.xn[ix + 1] - m.xn[ix] : m.yn[iy + 1] - m.yn[iy];
}


// Device buffers of one port's V0 solves (N_edge-long unless noted).  The
// last three are optional: the test-value ("a") products that find_Vh's
// per-frequency solve needs (src/matrixCon.cpp:412, :529, :553).
struct PortBufs {
  double *J = nullptr, *t = nullptr, *yd1 = nullptr, *yc = nullptr, *yd2 = nullptr;
  double *bd = nullptr, *xd = nullptr, *bc = nullptr, *xc = nullptr;  // nd / nc long
  double *ydat = nullptr, *yca = nullptr, *yd2a = nullptr;
};

typedef std::function<gp_status(RealAMG *, const double *, double *)> RealRun;

// paraGenerator's V0 section for source port src (src/matrixCon.cpp:367-587):
// J, then
//   y0d  = Ad^-1 (-V0da^T J) / w0,               yd1 = V0d y0d   (ydat = V0da y0d)
//   y0c  = Ac^-1 (-V0ca^T J + V0ca^T (-w0 D_eps yd1)), yc = V0c y0c (yca = V0ca y0c)
//   y0d2 = Ad^-1 (V0da^T (-D_eps yc)),           yd2 = V0d y0d2  (yd2a = V0da y0d2)
gp_status port_solves(gp_layout *lay, V0Path &P, int32_t src, double w0, PortBufs &b, Arena &tmp,
                      const RealRun &run) {
  hipStream_t s = lay->stream;
  const gp_mesh &m = lay->hm.view;
  const int64_t ne = P.ne, nd = P.nd, nc = P.nc;
  const unsigned ge = blocks_for(ne);
  // J: portDirection of each side, assigned side by side (last assignment wins)
  std::unordered_map<int64_t, size_t> jm;
  std::vector<int64_t> ji;
  std::vector<double> jvh;
  for (int32_t sd = m.port_side_ptr[src]; sd < m.port_side_ptr[src + 1]; ++sd)
    for (int64_t q = m.side_edge_ptr[sd]; q < m.side_edge_ptr[sd + 1]; ++q) {
      auto it = jm.find(m.side_edges[q]);
      if (it == jm.end()) { jm.emplace(m.side_edges[q], ji.size()); ji.push_back(m.side_edges[q]); jvh.push_back(m.side_dir[sd]); }
      else jvh[it->second] = m.side_dir[sd];
    }
  const int64_t nj = (int64_t)ji.size();
  GP_HIP(hipMemsetAsync(b.J, 0, sizeof(double) * ne, s));
  if (nj > 0) {
    int64_t *jidx = nullptr;
    double *jv = nullptr;
    GP_TRY(tmp.alloc(&jidx, nj)); GP_TRY(tmp.alloc(&jv, nj));
    GP_HIP(hipMemcpyAsync(jidx, ji.data(), sizeof(int64_t) * nj, hipMemcpyHostToDevice, s));
    GP_HIP(hipMemcpyAsync(jv, jvh.data(), sizeof(double) * nj, hipMemcpyHostToDevice, s));
    k_scatter<<<blocks_for(nj), 256, 0, s>>>(nj, jidx, jv, b.J);
    GP_HIP(hipStreamSynchronize(s));
    tmp.release(jidx, nj);
    tmp.release(jv, nj);
  }
  // y0d = Ad^-1 (-V0da^T J) / w0 ; yd1 = V0d y0d
  k_mv_wave<<<wave_blocks(nd), 256, 0, s>>>(nd, P.VdT.rp, P.VdT.ci, P.VdT.v[1], b.J, b.bd, 1);
  GP_TRY(run(P.ad, b.bd, b.xd));
  k_div<<<blocks_for(nd), 256, 0, s>>>(nd, w0, b.xd);
  k_mv_wave<<<wave_blocks(ne), 256, 0, s>>>(ne, P.Vd.rp, P.Vd.ci, P.Vd.v[0], b.xd, b.yd1, 0);
  if (b.ydat) k_mv_wave<<<wave_blocks(ne), 256, 0, s>>>(ne, P.Vd.rp, P.Vd.ci, P.Vd.v[1], b.xd, b.ydat, 0);
  if (nc > 0) {
    // y0c = Ac^-1 (-V0ca^T J + V0ca^T (yd1 * (-w0 D_eps))) ; yc = V0c y0c
    k_scale_eps_w<<<ge, 256, 0, s>>>(ne, lay->g.El, lay->epsn, w0, b.yd1, b.t);
    k_mv_wave<<<wave_blocks(nc), 256, 0, s>>>(nc, P.VcT.rp, P.VcT.ci, P.VcT.v[1], b.J, b.bc, 1);
    k_mv_wave<<<wave_blocks(nc), 256, 0, s>>>(nc, P.VcT.rp, P.VcT.ci, P.VcT.v[1], b.t, b.bc, 2);
    GP_TRY(run(P.ac, b.bc, b.xc));
    k_mv_wave<<<wave_blocks(ne), 256, 0, s>>>(ne, P.Vc.rp, P.Vc.ci, P.Vc.v[0], b.xc, b.yc, 0);
    if (b.yca) k_mv_wave<<<wave_blocks(ne), 256, 0, s>>>(ne, P.Vc.rp, P.Vc.ci, P.Vc.v[1], b.xc, b.yca, 0);
  } else {
    GP_HIP(hipMemsetAsync(b.yc, 0, sizeof(double) * ne, s));
    if (b.yca) GP_HIP(hipMemsetAsync(b.yca, 0, sizeof(double) * ne, s));
  }
  // y0d2 = Ad^-1 (V0da^T (-yc D_eps)) ; yd2 = V0d y0d2
  k_neg_eps<<<ge, 256, 0, s>>>(ne, lay->g.El, lay->epsn, b.yc, b.t);
  k_mv_wave<<<wave_blocks(nd), 256, 0, s>>>(nd, P.VdT.rp, P.VdT.ci, P.VdT.v[1], b.t, b.bd, 0);
  GP_TRY(run(P.ad, b.bd, b.xd));
  k_mv_wave<<<wave_blocks(ne), 256, 0, s>>>(ne, P.Vd.rp, P.Vd.ci, P.Vd.v[0], b.xd, b.yd2, 0);
  if (b.yd2a) k_mv_wave<<<wave_blocks(ne), 256, 0, s>>>(ne, P.Vd.rp, P.Vd.ci, P.Vd.v[1], b.xd, b.yd2a, 0);
  return GP_OK;
}

gp_status ensure_path(gp_layout *lay) {
  GP_TRY(ensure_device());
  if (!lay->v0) GP_TRY(build_v0(lay));
  if (lay->v0->v0d_cols < 1) return fail(GP_ERR_ARG, "the layout has no V0d columns");
  if (!lay->v0path) GP_TRY(setup(lay, &lay->v0path));
  return GP_OK;
}
}  // namespace

int64_t v0path_bytes(const V0Path *p) {
  if (!p) return 0;
  return p->A.bytes + (p->ad ? real_amg_bytes(p->ad) : 0) + (p->ac ? real_amg_bytes(p->ac) : 0);
}

}  // namespace gp

using namespace gp;

extern "C" gp_status gp_v0_set_smoother(gp_layout *lay, int32_t smoother) {
  if (!lay) return fail(GP_ERR_ARG, "null layout");
  if (smoother != GP_SMOOTH_L1_JACOBI && smoother != GP_SMOOTH_CHEBYSHEV) return fail(GP_ERR_ARG, "unknown real smoother");
  // a hierarchy built with the other smoother is rebuilt by the next V0 call
  if (lay->v0path && lay->v0path->smoother != smoother) {
    GP_HIP(hipStreamSynchronize(lay->stream));
    destroy_v0path(lay->v0path);
    lay->v0path = nullptr;
  }
  lay->v0_smoother = smoother;
  return GP_OK;
}

extern "C" gp_status gp_solve_v0(gp_layout *lay, double rtol, double *Z0, double *X, gp_v0_stats *stats) {
  if (!lay || !Z0) return fail(GP_ERR_ARG, "gp_solve_v0: null layout or Z0");
  GP_TRY(ensure_path(lay));
  V0Path &P = *lay->v0path;
  hipStream_t s = lay->stream;
  const gp_mesh &m = lay->hm.view;
  const int32_t np = m.num_ports;
  const double tol = rtol > 0 ? rtol : 1e-4;  // the reference's HYPRE tolerance (src/hypreSolver.h:8)
  const double w0 = 2 * kPi * m.freq_start * m.freq_unit;
  auto t0 = std::chrono::steady_clock::now();

  // response edges: side 0 of every port (Construct_Z_V0 reads only that side)
  std::vector<int64_t> redge;
  std::vector<int64_t> rptr(np + 1, 0);
  for (int32_t p = 0; p < np; ++p) {
    const int32_t s0 = m.port_side_ptr[p];
    for (int64_t q = m.side_edge_ptr[s0]; q < m.side_edge_ptr[s0 + 1]; ++q) redge.push_back(m.side_edges[q]);
    rptr[p + 1] = (int64_t)redge.size();
  }
  Arena tmp;
  struct TGuard { Arena *a; hipStream_t s; ~TGuard() { (void)hipStreamSynchronize(s); a->free_all(); } } tg{&tmp, s};
  const int64_t ne = P.ne, nd = P.nd, nc = P.nc;
  const int64_t nr = std::max<int64_t>((int64_t)redge.size(), 1);
  double *J, *t, *yd1, *yc, *yd2, *bd, *xd, *bc = nullptr, *xc = nullptr, *gat;
  int64_t *ridx;
  GP_TRY(tmp.alloc(&J, ne)); GP_TRY(tmp.alloc(&t, ne)); GP_TRY(tmp.alloc(&yd1, ne));
  GP_TRY(tmp.alloc(&yc, ne)); GP_TRY(tmp.alloc(&yd2, ne));
  GP_TRY(tmp.alloc(&bd, nd)); GP_TRY(tmp.alloc(&xd, nd));
  if (nc > 0) { GP_TRY(tmp.alloc(&bc, nc)); GP_TRY(tmp.alloc(&xc, nc)); }
  GP_TRY(tmp.alloc(&ridx, nr)); GP_TRY(tmp.alloc(&gat, 2 * nr));
  if (!redge.empty())
    GP_HIP(hipMemcpyAsync(ridx, redge.data(), sizeof(int64_t) * redge.size(), hipMemcpyHostToDevice, s));

  std::vector<double> Zh(2 * (size_t)np * np, 0.), yh(2 * nr);
  int32_t solves = 0, maxit = 0;
  double maxrel = 0.;
  gp_status result = GP_OK;
  RealRun run = [&](RealAMG *G, const double *b, double *x) -> gp_status {
    int32_t it = 0;
    double rel = 0.;
    gp_status st = real_amg_solve(G, b, x, tol, 1000, &it, &rel, s);
    ++solves;
    maxit = std::max(maxit, it);
    maxrel = std::max(maxrel, rel);
    if (st == GP_ERR_NOCONV) { result = GP_ERR_NOCONV; return GP_OK; }  // reported, x kept (as hypreSolve)
    return st;
  };
  PortBufs pb;
  pb.J = J; pb.t = t; pb.yd1 = yd1; pb.yc = yc; pb.yd2 = yd2; pb.bd = bd; pb.xd = xd; pb.bc = bc; pb.xc = xc;
  for (int32_t src = 0; src < np; ++src) {
    GP_TRY(port_solves(lay, P, src, w0, pb, tmp, run));
    // yd at the response edges, then Construct_Z_V0 on the host
    if (!redge.empty()) {
      k_gather_yd<<<blocks_for((int64_t)redge.size()), 256, 0, s>>>((int64_t)redge.size(), ridx, yd2, yd1, yc, gat);
      GP_HIP(hipMemcpyAsync(yh.data(), gat, sizeof(double) * 2 * redge.size(), hipMemcpyDeviceToHost, s));
    }
    GP_HIP(hipStreamSynchronize(s));
    double cur = 0.;
    for (int32_t sd = m.port_side_ptr[src]; sd < m.port_side_ptr[src + 1]; ++sd) cur += m.side_area[sd];
    for (int32_t p = 0; p < np; ++p) {
      const size_t zi = (size_t)p + (size_t)np * src;
      double zr = Zh[2 * zi], zim = Zh[2 * zi + 1];
      const double dir = m.side_dir[m.port_side_ptr[p]] * 1.0;
      for (int64_t q = rptr[p]; q < rptr[p + 1]; ++q) {
        const double leng = edge_len(m, lay->g, redge[q]);
        double tr = yh[2 * q] * leng, ti = yh[2 * q + 1] * leng;
        tr = tr * dir; ti = ti * dir;
        zr -= tr; zim -= ti;
      }
      Zh[2 * zi] = zr / cur;
      Zh[2 * zi + 1] = zim / cur;
    }
  }
  std::copy(Zh.begin(), Zh.end(), Z0);
  if (X) {
    const int32_t nf = std::max(m.nfreq, 1);
    std::fill(X, X + 2 * (size_t)nf * np * np, 0.);
    std::copy(Zh.begin(), Zh.end(), X);
    GP_TRY(gp_z_v0_sweep(&m, X));
  }
  if (stats) {
    stats->solves = solves;
    stats->max_iters = maxit;
    stats->max_rel_residual = maxrel;
    stats->setup_ms = P.setup_ms;
    stats->solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats->ad_levels = real_amg_levels(P.ad);
    stats->ac_levels = P.ac ? real_amg_levels(P.ac) : 0;
    stats->ad_rows = nd;
    stats->ac_rows = nc;
  }
  return result == GP_OK ? GP_OK : fail(GP_ERR_NOCONV, "a V0 real solve stopped above the tolerance");
}

// The V0 vectors of find_Vh's per-frequency solve for one source port
// (src/matrixCon.cpp:404-427, :547-568, :569-587), from the same solves as
// gp_solve_v0; the norms and the final scaling run on the host in the
// reference's order.
extern "C" gp_status gp_v0_vh_inputs(gp_layout *lay, int32_t src, double rtol, double *u0, double *u0a,
                                     double *yd) {
  if (!lay || !u0 || !u0a || !yd) return fail(GP_ERR_ARG, "gp_v0_vh_inputs: null argument");
  if (src < 0 || src >= lay->hm.view.num_ports) return fail(GP_ERR_ARG, "gp_v0_vh_inputs: source port out of range");
  GP_TRY(ensure_path(lay));
  V0Path &P = *lay->v0path;
  hipStream_t s = lay->stream;
  const gp_mesh &m = lay->hm.view;
  const double tol = rtol > 0 ? rtol : 1e-4;
  const double w0 = 2 * kPi * m.freq_start * m.freq_unit;
  const int64_t ne = P.ne, nd = P.nd, nc = P.nc, n = lay->n;
  Arena tmp;
  struct TGuard { Arena *a; hipStream_t s; ~TGuard() { (void)hipStreamSynchronize(s); a->free_all(); } } tg{&tmp, s};
  PortBufs b;
  GP_TRY(tmp.alloc(&b.J, ne)); GP_TRY(tmp.alloc(&b.t, ne)); GP_TRY(tmp.alloc(&b.yd1, ne));
  GP_TRY(tmp.alloc(&b.yc, ne)); GP_TRY(tmp.alloc(&b.yd2, ne));
  GP_TRY(tmp.alloc(&b.ydat, ne)); GP_TRY(tmp.alloc(&b.yca, ne)); GP_TRY(tmp.alloc(&b.yd2a, ne));
  GP_TRY(tmp.alloc(&b.bd, nd)); GP_TRY(tmp.alloc(&b.xd, nd));
  if (nc > 0) { GP_TRY(tmp.alloc(&b.bc, nc)); GP_TRY(tmp.alloc(&b.xc, nc)); }
  bool noconv = false;
  RealRun run = [&](RealAMG *G, const double *rhs, double *x) -> gp_status {
    int32_t it = 0;
    double rel = 0.;
    gp_status st = real_amg_solve(G, rhs, x, tol, 1000, &it, &rel, s);
    if (st == GP_ERR_NOCONV) { noconv = true; return GP_OK; }
    return st;
  };
  GP_TRY(port_solves(lay, P, src, w0, b, tmp, run));
  std::vector<double> yd1(ne), ydat(ne), yd2(ne), yd2a(ne), yc(ne), yca(ne);
  std::vector<int64_t> mr(n);
  GP_HIP(hipMemcpyAsync(yd1.data(), b.yd1, sizeof(double) * ne, hipMemcpyDeviceToHost, s));
  GP_HIP(hipMemcpyAsync(ydat.data(), b.ydat, sizeof(double) * ne, hipMemcpyDeviceToHost, s));
  GP_HIP(hipMemcpyAsync(yd2.data(), b.yd2, sizeof(double) * ne, hipMemcpyDeviceToHost, s));
  GP_HIP(hipMemcpyAsync(yd2a.data(), b.yd2a, sizeof(double) * ne, hipMemcpyDeviceToHost, s));
  GP_HIP(hipMemcpyAsync(yc.data(), b.yc, sizeof(double) * ne, hipMemcpyDeviceToHost, s));
  GP_HIP(hipMemcpyAsync(yca.data(), b.yca, sizeof(double) * ne, hipMemcpyDeviceToHost, s));
  GP_HIP(hipMemcpyAsync(mr.data(), lay->map_edge_r, sizeof(int64_t) * n, hipMemcpyDeviceToHost, s));
  GP_HIP(hipStreamSynchronize(s));
  // u0 = [V0d y0d, V0d y0d2 + V0c y0c], u0a = [V0da y0d, V0da y0d2 + V0ca y0c],
  // each column normalised over all N_edge entries, rows by mapEdgeR
  double nn = 0, nna = 0;
  for (int64_t e = 0; e < ne; ++e) { nn += yd1[e] * yd1[e]; nna += ydat[e] * ydat[e]; }
  nn = std::sqrt(nn); nna = std::sqrt(nna);
  for (int64_t i = 0; i < n; ++i) { u0[i] = yd1[mr[i]] / nn; u0a[i] = ydat[mr[i]] / nna; }
  nn = 0; nna = 0;
  for (int64_t e = 0; e < ne; ++e) {
    nn += (yd2[e] + yc[e]) * (yd2[e] + yc[e]);
    nna += (yd2a[e] + yca[e]) * (yd2a[e] + yca[e]);
  }
  nn = std::sqrt(nn); nna = std::sqrt(nna);
  for (int64_t i = 0; i < n; ++i) {
    u0[n + i] = (yd2[mr[i]] + yc[mr[i]]) / nn;
    u0a[n + i] = (yd2a[mr[i]] + yca[mr[i]]) / nna;
  }
  // yd = yd2 - i yd1, then += yc
  for (int64_t e = 0; e < ne; ++e) { yd[2 * e] = yd2[e] + yc[e]; yd[2 * e + 1] = -yd1[e]; }
  return noconv ? fail(GP_ERR_NOCONV, "a V0 real solve stopped above the tolerance") : GP_OK;
}

// print_z_V0's store (src/fdtd.hpp:4249-4316)
extern "C" gp_status gp_z_v0_sweep(const gp_mesh *m, double *X) {
  if (!m || !X) return fail(GP_ERR_ARG, "gp_z_v0_sweep: null argument");
  const int32_t np = m->num_ports, nf = m->nfreq;
  if (nf <= 1) return GP_OK;
  for (int32_t id = 1; id < nf; ++id) {
    double f;
    if (id == nf - 1) f = m->freq_end * m->freq_unit;
    else if (m->freq_scale == 1) f = (m->freq_start + id * (m->freq_end - m->freq_start) / (nf - 1)) * m->freq_unit;
    else f = m->freq_start * m->freq_unit * std::pow(m->freq_end / m->freq_start, (id * 1.0 / (nf - 1)));
    for (int32_t i = 0; i < np; ++i)
      for (int32_t j = 0; j < np; ++j) {
        const double re = X[2 * (j + (size_t)i * np)], im = X[2 * (j + (size_t)i * np) + 1];
        const size_t o = (size_t)id * np * np + i + (size_t)j * np;  // stored transposed
        X[2 * o] = re;
        X[2 * o + 1] = im * m->freq_start * m->freq_unit / f;
      }
  }
  return GP_OK;
}

// solver.hip -- full-wave solve (the system reference() hands to PARDISO,
// src/generateStiff.cpp:568-699) as batched flexible GMRES in FP64 with a
// multigrid preconditioner, and Z extraction (Construct_Z_V0_Vh,
// src/fdtd.hpp:4133-4170).
//
// Columns are (port, frequency) pairs scheduled onto k device slots.  The k
// slots share every sparse matrix read (SpMM); each keeps its own Arnoldi
// recurrence, Hessenberg matrix and Givens rotations.  At every restart the
// FP64 residual of each slot is formed exactly; a slot whose column has
// converged hands its solution to Z extraction and takes the next pending
// column (continuous batching), so slots never wait for the slowest column.
//
// Iterates are kept in split form x = u + G phi + Gx a (G the nodal gradient,
// Gx the [V0d | V0c] basis) so that A x = S u + (S G) phi + (S Gx) a +
// D(w) x is evaluated without cancellation: S G and S Gx vanish up to the
// rounding of S and are formed exactly, which is what lets a 1e-8 relative
// residual be met at low frequency where the gradient part of x dominates.
#include <algorithm>
#include <type_traits>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>

#include "mg.hpp"

namespace gp {

// Device buffers of the batched solve (Krylov blocks, iterate, V-cycle
// vectors), kept with the layout between calls so that repeated solves with
// the same shape allocate nothing; a call with another shape replaces them.
struct SlotWorkBase {
  virtual ~SlotWorkBase() {}
  int k = 0, m = 0, tsz = 0, tvsz = 0;
  bool want_x = false;
  int64_t bytes = 0;
};

// The autotuner's choice for the mixed-precision path (gp_autotune).
struct Tuning {
  bool done = false;
  int k = 16;                    // device slots (block width)
  int lanes = 16;                // lanes-per-row cap of the V-cycle's SELL kernels
  double us[3] = {-1, -1, -1};   // per column-iteration at 8, 16, 32 slots
  double lanes_us[3] = {-1, -1, -1};  // the chosen width at caps 16, 4, 2
  double ms = 0;
};

struct Solver {
  Arena arena;                   // the hierarchy's buffers
  Hierarchy H;
  bool f32 = false;
  CoarseCache cc;
  SlotWorkBase *ws = nullptr;
  Tuning tune;
  int64_t graph_captures = 0, graph_replays = 0;
  bool graph_failed = false;
};

void release_workspace(Solver *s) {
  if (!s || !s->ws) return;
  delete s->ws;
  s->ws = nullptr;
}

void destroy_solver(Solver *s) {
  if (!s) return;
  release_workspace(s);
  coarse_cache_free(s->cc);
  s->arena.free_all();
  delete s;
}

const Hierarchy *solver_hierarchy(const Solver *s) { return s ? &s->H : nullptr; }

int64_t solver_bytes(const Solver *s) {
  if (!s) return 0;
  return s->arena.bytes + s->cc.bytes + (s->ws ? s->ws->bytes : 0);
}

namespace {

using cd = std::complex<double>;
typedef C2<double> Z2;

// ---------------------------------------------------------------- FP64 outer kernels

// Real part of the system diagonal beyond S_ii.  reference() forms the diagonal
// entry as the FP64 sum S_ii + w * (-w * eps_r * EPSILON0) (src/generateStiff.cpp:
// 618-626); next to S_ii ~ 1e18 that sum drops the low bits of the w^2 eps term,
// a relative change of up to ~1e-4 in it at 100 MHz that moves Z by ~1e-6.  The
// operator applied here is S + D' with D' = fl(S_ii + D) - S_ii (exact when the
// two terms are within a factor of two, else correct to one rounding), i.e. the
// reference's matrix entry for entry.
__device__ __forceinline__ double reference_diag_re(double sii, double w, double epsr) {
#pragma clang fp contract(off)
  const double d = w * (-(w)*epsr * kEps0);
  const double t = sii + d;
  return t - sii;
}

struct OuterOp {
  int64_t n;
  const int64_t *rp; const int32_t *ci; const double *sv;       // S
  const int64_t *grp; const int32_t *gci; const double *gv;     // G
  const int64_t *xrp; const int32_t *xci; const double *xv;     // Gx
  const int64_t *sgrp; const int32_t *sgci; const double *sgv;  // S G (exact)
  const int64_t *sxrp; const int32_t *sxci; const double *sxv;  // S Gx (exact)
  const double *epsr, *sig, *sdiag;
};

// out = S u [+ SG phi + SGx a] + D(w)(u + G phi + Gx a)   (mode 0)   or b - (...)   (mode 1)
// D(w) = w * complex(-w * eps_r * EPSILON0, sigma) with the real part rounded as
// above (src/generateStiff.cpp:620-626).  exact = 1 adds the S G / S Gx terms.
template <class TZ, class TO>
__global__ void k_outer(OuterOp op, int k, const double *__restrict__ omega, const C2<TZ> *__restrict__ u,
                        const C2<TZ> *__restrict__ phi, const C2<TZ> *__restrict__ a, const Z2 *__restrict__ b,
                        C2<TO> *__restrict__ out, int mode, int exact) {
  int64_t t = xcd_block_id() * blockDim.x + threadIdx.x;
  int64_t i = t / k;
  int j = (int)(t - i * k);
  if (i >= op.n) return;
  Z2 acc = {0., 0.};
  for (int64_t p = op.rp[i]; p < op.rp[i + 1]; ++p) {
    const double m = op.sv[p];
    const C2<TZ> x = u[(int64_t)op.ci[p] * k + j];
    acc.x += m * (double)x.x;
    acc.y += m * (double)x.y;
  }
  Z2 xi = {(double)u[i * k + j].x, (double)u[i * k + j].y};
  for (int64_t p = op.grp[i]; p < op.grp[i + 1]; ++p) {
    const double g = op.gv[p];
    const C2<TZ> x = phi[(int64_t)op.gci[p] * k + j];
    xi.x += g * (double)x.x;
    xi.y += g * (double)x.y;
  }
  for (int64_t p = op.xrp[i]; p < op.xrp[i + 1]; ++p) {
    const double g = op.xv[p];
    const C2<TZ> x = a[(int64_t)op.xci[p] * k + j];
    xi.x += g * (double)x.x;
    xi.y += g * (double)x.y;
  }
  if (exact) {  // rounding residue of S on the gradient parts (exact products, tiny)
    for (int64_t p = op.sgrp[i]; p < op.sgrp[i + 1]; ++p) {
      const C2<TZ> x = phi[(int64_t)op.sgci[p] * k + j];
      acc.x += op.sgv[p] * (double)x.x;
      acc.y += op.sgv[p] * (double)x.y;
    }
    for (int64_t p = op.sxrp[i]; p < op.sxrp[i + 1]; ++p) {
      const C2<TZ> x = a[(int64_t)op.sxci[p] * k + j];
      acc.x += op.sxv[p] * (double)x.x;
      acc.y += op.sxv[p] * (double)x.y;
    }
  }
  const double w = omega[j];
  const double dr = reference_diag_re(op.sdiag[i], w, op.epsr[i]), di = w * op.sig[i];
  acc.x += dr * xi.x - di * xi.y;
  acc.y += dr * xi.y + di * xi.x;
  if (mode == 1) { const Z2 bb = b[i * k + j]; acc.x = bb.x - acc.x; acc.y = bb.y - acc.y; }
  out[i * k + j] = {(TO)acc.x, (TO)acc.y};
}

template <class TZ, int KS, int NB>
__device__ __forceinline__ void outer_batch(const SellV<double> &SS, int64_t b0, int t0, int lane,
                                            const C2<TZ> *__restrict__ u, int k, int j0, Z2 (&acc)[KS]) {
  int32_t c[NB];
  double m[NB];
#pragma unroll
  for (int h = 0; h < NB; ++h) {
    const int64_t e = b0 + (int64_t)(t0 + h) * 64 + lane;
    // streamed once per launch: non-temporal, keeping the caches for x
    c[h] = __builtin_nontemporal_load(SS.ci + e);
    m[h] = __builtin_nontemporal_load(SS.v[0] + e);
  }
  C2<TZ> x[NB][KS];
#pragma unroll
  for (int h = 0; h < NB; ++h) load_slots<TZ, KS>(u + (int64_t)c[h] * k + j0, x[h]);
#pragma unroll
  for (int h = 0; h < NB; ++h)
#pragma unroll
    for (int q = 0; q < KS; ++q) { acc[q].x += m[h] * (double)x[h][q].x; acc[q].y += m[h] * (double)x[h][q].y; }
}

// k_outer on the SELL copy of S: LPR adjacent lanes share a row, each with
// SPL slots (one 16-byte load of FP32 u per gathered row), S read once per
// lane group; G, Gx, S G, S Gx rows are short CSR walks.
template <class TZ, class TO, int SPL, int LPR>
__global__ void __launch_bounds__(256) k_outer_sell(OuterOp op, SellV<double> SS, int k, const double *__restrict__ omega,
                                                    const C2<TZ> *__restrict__ u, const C2<TZ> *__restrict__ phi,
                                                    const C2<TZ> *__restrict__ a, const Z2 *__restrict__ b,
                                                    C2<TO> *__restrict__ out, int mode, int exact) {
  constexpr int KS = SPL;
  const int64_t t_ = xcd_block_id() * blockDim.x + threadIdx.x;
  const int64_t pos = t_ / LPR;
  if (pos >= op.n) return;
  const int j0 = ((int)blockIdx.y * LPR + (int)(t_ - pos * LPR)) * SPL;
  const int64_t i = SS.rows ? (int64_t)SS.rows[pos] : pos;
  const int64_t sl = pos >> 6;
  const int lane = (int)(pos & 63);
  Z2 acc[KS], xi[KS];
#pragma unroll
  for (int q = 0; q < KS; ++q) acc[q] = {0., 0.};
  // The G and Gx rows hold two entries or fewer on nearly every edge: their
  // first two indices and values are loaded before the S gathers, so the phi
  // and a gathers after them wait on no index load of their own
  const int64_t g0 = op.grp[i], g1 = op.grp[i + 1], x0 = op.xrp[i], x1 = op.xrp[i + 1];
  int32_t gc[2] = {0, 0}, xc[2] = {0, 0};
  double gw[2] = {0., 0.}, xw[2] = {0., 0.};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (g0 + h < g1) { gc[h] = op.gci[g0 + h]; gw[h] = op.gv[g0 + h]; }
    if (x0 + h < x1) { xc[h] = op.xci[x0 + h]; xw[h] = op.xv[x0 + h]; }
  }
  {
    const int64_t b0 = SS.base[sl];
    const int wd = SS.w[sl];
    // batches of 4, then single entries: indices and values first, then the
    // batch's row gathers in flight together; order unchanged
    int t0 = 0;
    for (; t0 + 4 <= wd; t0 += 4) outer_batch<TZ, KS, 4>(SS, b0, t0, lane, u, k, j0, acc);
    for (; t0 < wd; ++t0) outer_batch<TZ, KS, 1>(SS, b0, t0, lane, u, k, j0, acc);
  }
  {
    C2<TZ> x[KS];
    load_slots<TZ, KS>(u + i * k + j0, x);
#pragma unroll
    for (int q = 0; q < KS; ++q) xi[q] = {(double)x[q].x, (double)x[q].y};
  }
  {
    C2<TZ> xg[2][KS];
#pragma unroll
    for (int h = 0; h < 2; ++h)
      if (g0 + h < g1) load_slots<TZ, KS>(phi + (int64_t)gc[h] * k + j0, xg[h]);
#pragma unroll
    for (int h = 0; h < 2; ++h)
      if (g0 + h < g1)
#pragma unroll
        for (int q = 0; q < KS; ++q) { xi[q].x += gw[h] * (double)xg[h][q].x; xi[q].y += gw[h] * (double)xg[h][q].y; }
    for (int64_t p = g0 + 2; p < g1; ++p) {
      const double g = op.gv[p];
      C2<TZ> x[KS];
      load_slots<TZ, KS>(phi + (int64_t)op.gci[p] * k + j0, x);
#pragma unroll
      for (int q = 0; q < KS; ++q) { xi[q].x += g * (double)x[q].x; xi[q].y += g * (double)x[q].y; }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
      if (x0 + h < x1) load_slots<TZ, KS>(a + (int64_t)xc[h] * k + j0, xg[h]);
#pragma unroll
    for (int h = 0; h < 2; ++h)
      if (x0 + h < x1)
#pragma unroll
        for (int q = 0; q < KS; ++q) { xi[q].x += xw[h] * (double)xg[h][q].x; xi[q].y += xw[h] * (double)xg[h][q].y; }
    for (int64_t p = x0 + 2; p < x1; ++p) {
      const double g = op.xv[p];
      C2<TZ> x[KS];
      load_slots<TZ, KS>(a + (int64_t)op.xci[p] * k + j0, x);
#pragma unroll
      for (int q = 0; q < KS; ++q) { xi[q].x += g * (double)x[q].x; xi[q].y += g * (double)x[q].y; }
    }
  }
  if (exact) {
    // rows of S G and S Gx, four entries at a time: their indices and values
    // first, then the four row gathers in flight together; order unchanged
    auto walk = [&](const int64_t *__restrict__ rp, const int32_t *__restrict__ ci, const double *__restrict__ v,
                    const C2<TZ> *__restrict__ src) {
      const int64_t p1 = rp[i + 1];
      int64_t p = rp[i];
      for (; p + 4 <= p1; p += 4) {
        int32_t cc[4];
        double gg[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) { cc[h] = ci[p + h]; gg[h] = v[p + h]; }
        C2<TZ> x[4][KS];
#pragma unroll
        for (int h = 0; h < 4; ++h) load_slots<TZ, KS>(src + (int64_t)cc[h] * k + j0, x[h]);
#pragma unroll
        for (int h = 0; h < 4; ++h)
#pragma unroll
          for (int q = 0; q < KS; ++q) { acc[q].x += gg[h] * (double)x[h][q].x; acc[q].y += gg[h] * (double)x[h][q].y; }
      }
      for (; p < p1; ++p) {
        const double g = v[p];
        C2<TZ> x[KS];
        load_slots<TZ, KS>(src + (int64_t)ci[p] * k + j0, x);
#pragma unroll
        for (int q = 0; q < KS; ++q) { acc[q].x += g * (double)x[q].x; acc[q].y += g * (double)x[q].y; }
      }
    };
    walk(op.sgrp, op.sgci, op.sgv, phi);
    walk(op.sxrp, op.sxci, op.sxv, a);
  }
  const double sii = op.sdiag[i], er = op.epsr[i], sg = op.sig[i];
  Z2 bb[KS];
  if (mode == 1) load_slots<double, KS>(b + i * k + j0, bb);
#pragma unroll
  for (int q = 0; q < KS; ++q) {
    const double w = omega[j0 + q];
    const double dr = reference_diag_re(sii, w, er), di = w * sg;
    acc[q].x += dr * xi[q].x - di * xi[q].y;
    acc[q].y += dr * xi[q].y + di * xi[q].x;
    if (mode == 1) acc[q] = {bb[q].x - acc[q].x, bb[q].y - acc[q].y};
  }
  C2<TO> o[KS];
#pragma unroll
  for (int q = 0; q < KS; ++q) o[q] = {(TO)acc[q].x, (TO)acc[q].y};
  store_slots<TO, KS>(out + i * k + j0, o);
}

// y = A(w) x: the complex matrix reference() assembles, S + diag(w (-w eps_r
// EPSILON0) + i w sigma) with the diagonal summed in FP64 as the reference
// does (src/generateStiff.cpp:618-626), applied to k columns interleaved
// (row i, column j at i*k + j).  S's SELL copy in level-0 tile order; LPR
// lanes per row, one complex FP64 slot (one 16-byte load) per lane, four row
// gathers in flight per lane.
template <int SPL, int LPR>
__global__ void __launch_bounds__(256) k_amul_sell(SellV<double> SS, int64_t n, int k, const double *__restrict__ omega,
                                                   const double *__restrict__ sdiag, const int16_t *__restrict__ mcode,
                                                   const double *__restrict__ tab_eps, const double *__restrict__ tab_sig,
                                                   const Z2 *__restrict__ x, Z2 *__restrict__ y) {
  const int64_t t_ = xcd_block_id() * blockDim.x + threadIdx.x;
  const int64_t pos = t_ / LPR;
  if (pos >= n) return;
  const int j0 = ((int)blockIdx.y * LPR + (int)(t_ - pos * LPR)) * SPL;
  const int64_t i = SS.rows ? (int64_t)SS.rows[pos] : pos;
  const int64_t sl = pos >> 6;
  const int lane = (int)(pos & 63);
  Z2 acc[SPL];
#pragma unroll
  for (int q = 0; q < SPL; ++q) acc[q] = {0., 0.};
  const int64_t b0 = SS.base[sl];
  const int wd = SS.w[sl];
  int t0 = 0;
  if constexpr (SPL == 1)
    for (; t0 + 8 <= wd; t0 += 8) outer_batch<double, SPL, 8>(SS, b0, t0, lane, x, k, j0, acc);
  for (; t0 + 4 <= wd; t0 += 4) outer_batch<double, SPL, 4>(SS, b0, t0, lane, x, k, j0, acc);
  for (; t0 < wd; ++t0) outer_batch<double, SPL, 1>(SS, b0, t0, lane, x, k, j0, acc);
  const int mc = mcode[i];  // layer * 2 + (conductor edge)
  const double er = tab_eps[mc >> 1], sg = (mc & 1) ? tab_sig[mc >> 1] : 0.;
  const double sii = sdiag[i];
  Z2 xi[SPL];
  load_slots<double, SPL>(x + i * k + j0, xi);
#pragma unroll
  for (int q = 0; q < SPL; ++q) {
    const double w = omega[j0 + q];
    const double dr = reference_diag_re(sii, w, er), di = w * sg;
    acc[q].x += dr * xi[q].x - di * xi[q].y;
    acc[q].y += dr * xi[q].y + di * xi[q].x;
  }
  store_slots<double, SPL>(y + i * k + j0, acc);
}

}  // namespace

void amul(const Hierarchy &H, int64_t n, int k, const double *om, const Z2 *x, Z2 *y, hipStream_t s) {
  // one column: the copy in the reference's row order, whose wavefronts read
  // and write whole lines of the row-indexed vectors
  const Sell &S = (k == 1 && H.Ss1.nr == H.Ss.nr) ? H.Ss1 : H.Ss;
  SellV<double> SS;
  SS.nr = S.nr; SS.nslice = S.nslice; SS.base = S.base; SS.w = S.w; SS.ci = S.ci; SS.nv = 1;
  SS.v[0] = S.v[0];
  SS.rows = S.rows;
  const int sp = (k % 2 == 0) ? 2 : 1;  // two complex FP64 slots = two 16-byte loads per lane
  int lpr = 1;
  while (lpr < 16 && (k / sp) % (lpr * 2) == 0) lpr *= 2;
  const dim3 grid((unsigned)((n * lpr + 255) / 256), (unsigned)(k / (sp * lpr)));
#define GP_AMUL(SP, LP) \
  k_amul_sell<SP, LP><<<grid, 256, 0, s>>>(SS, n, k, om, H.sdiag, H.mcode, H.tab_eps, H.tab_sig, x, y)
  if (sp == 2) {
    if (lpr == 16) GP_AMUL(2, 16);
    else if (lpr == 8) GP_AMUL(2, 8);
    else if (lpr == 4) GP_AMUL(2, 4);
    else if (lpr == 2) GP_AMUL(2, 2);
    else GP_AMUL(2, 1);
  } else {
    if (lpr == 16) GP_AMUL(1, 16);
    else if (lpr == 8) GP_AMUL(1, 8);
    else if (lpr == 4) GP_AMUL(1, 4);
    else if (lpr == 2) GP_AMUL(1, 2);
    else GP_AMUL(1, 1);
  }
#undef GP_AMUL
}

// column-major host layout (column j at j*n) <-> interleaved rows (i*k + j)
__global__ void k_cols_to_rows(int64_t n, int k, const Z2 *__restrict__ a, Z2 *__restrict__ b) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * k) return;
  const int64_t i = t / k;
  const int j = (int)(t - i * k);
  b[t] = a[(int64_t)j * n + i];
}
__global__ void k_rows_to_cols(int64_t n, int k, const Z2 *__restrict__ a, Z2 *__restrict__ b) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * k) return;
  const int64_t i = t / k;
  const int j = (int)(t - i * k);
  b[(int64_t)j * n + i] = a[t];
}

namespace {

// Multi-vector dots.  Block (blk, c) handles vectors t in [4c, 4c+4) of V
// (stride vstride) against w over a grid-stride range of rows and writes
// partial[(blk * nout + t) * k + j] = sum conj(V_t) w; with self, chunk 0 also
// writes t = nvec: sum |w|^2.  nout = nvec + self.  With w in FP32, small
// chunks (more wavefronts per SIMD, w read once per chunk) won: at
// 200x200x24 the solve ran about 1 % faster with 8 than with 16, and about
// 1.5 % faster again with 4.
constexpr int kDotChunk = 4;
template <class TV, class TW>
__global__ void __launch_bounds__(256) k_mdots(int64_t n, int k, const C2<TV> *__restrict__ V, int64_t vstride, int nvec,
                                               const C2<TW> *__restrict__ w, int self, Z2 *__restrict__ partial) {
  __shared__ Z2 sh[256];
  const int c = blockIdx.y;
  const int t0 = c * kDotChunk;
  const int nt = min(kDotChunk, nvec - t0);
  const int do_self = self && c == 0;
  const int rows_per = blockDim.x / k;
  const int j = threadIdx.x % k;
  Z2 acc[kDotChunk + 1];
#pragma unroll
  for (int q = 0; q <= kDotChunk; ++q) acc[q] = {0., 0.};
  const int64_t step = (int64_t)gridDim.x * rows_per;
  // two rows per pass, every load of both issued before the products: one
  // row per pass left each wavefront waiting a full memory latency per row
  auto row = [&](const C2<TW> &wv, const C2<TV> (&v)[kDotChunk]) {
    const Z2 b = {(double)wv.x, (double)wv.y};
#pragma unroll
    for (int q = 0; q < kDotChunk; ++q)
      if (q < nt) {
        const Z2 a = {(double)v[q].x, (double)v[q].y};
        acc[q].x += a.x * b.x + a.y * b.y;
        acc[q].y += a.x * b.y - a.y * b.x;
      }
    if (do_self) acc[kDotChunk].x += b.x * b.x + b.y * b.y;
  };
  int64_t i = (int64_t)blockIdx.x * rows_per + threadIdx.x / k;
  for (; i + step < n; i += 2 * step) {
    const int64_t i2 = i + step;
    const C2<TW> w1 = w[i * k + j], w2 = w[i2 * k + j];
    C2<TV> v1[kDotChunk], v2[kDotChunk];
#pragma unroll
    for (int q = 0; q < kDotChunk; ++q)
      if (q < nt) {
        v1[q] = V[(int64_t)(t0 + q) * vstride + i * k + j];
        v2[q] = V[(int64_t)(t0 + q) * vstride + i2 * k + j];
      }
    row(w1, v1);
    row(w2, v2);
  }
  if (i < n) {
    const C2<TW> w1 = w[i * k + j];
    C2<TV> v1[kDotChunk];
#pragma unroll
    for (int q = 0; q < kDotChunk; ++q)
      if (q < nt) v1[q] = V[(int64_t)(t0 + q) * vstride + i * k + j];
    row(w1, v1);
  }
  const int nout = nvec + self;
  for (int q = 0; q <= kDotChunk; ++q) {
    const bool live = q < nt || (q == kDotChunk && do_self);
    if (!live) continue;  // uniform across the block
    sh[threadIdx.x] = acc[q];
    __syncthreads();
    for (int s = blockDim.x / 2; s >= k; s >>= 1) {
      if ((int)threadIdx.x < s) { sh[threadIdx.x].x += sh[threadIdx.x + s].x; sh[threadIdx.x].y += sh[threadIdx.x + s].y; }
      __syncthreads();
    }
    const int t = q < kDotChunk ? t0 + q : nvec;
    if ((int)threadIdx.x < k) partial[((int64_t)blockIdx.x * nout + t) * k + threadIdx.x] = sh[threadIdx.x];
    __syncthreads();
  }
}

// h[q] = sum over the nb blocks of partial[b * nout * k + q]: one wavefront
// per output, lanes over blocks, then a fixed-order shuffle tree.
__global__ void k_mdots_final(int nout, int nb, int k, const Z2 *__restrict__ partial, Z2 *__restrict__ h) {
  const int q = (int)(blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64);
  const int lane = (int)(threadIdx.x & 63);
  if (q >= nout * k) return;  // uniform across the wavefront
  Z2 s = {0., 0.};
  for (int b = lane; b < nb; b += 64) { const Z2 v = partial[(int64_t)b * nout * k + q]; s.x += v.x; s.y += v.y; }
  for (int off = 32; off > 0; off >>= 1) { s.x += __shfl_down(s.x, off); s.y += __shfl_down(s.y, off); }
  if (lane == 0) h[q] = s;
}

// out = (w - sum_{t<nvec} V_t h_t) * scale_j   (per column; out may be w);
// with rout, also rout = W out (the next preconditioner input, from the
// stored value of out)
template <class TV, class TW, class TO, class TR>
__global__ void k_gs_update(int64_t n, int k, int nvec, const C2<TV> *__restrict__ V, int64_t vstride,
                            const Z2 *__restrict__ h, const double *__restrict__ scale, const C2<TW> *w, C2<TO> *out,
                            const double *__restrict__ W, C2<TR> *__restrict__ rout) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * k) return;
  int j = (int)(q % k);
  const C2<TW> w0 = w[q];
  Z2 acc = {(double)w0.x, (double)w0.y};
#pragma unroll 4
  for (int t = 0; t < nvec; ++t) {
    const C2<TV> a = V[(int64_t)t * vstride + q];
    const Z2 c = h[t * k + j];
    acc.x -= (double)a.x * c.x - (double)a.y * c.y;
    acc.y -= (double)a.x * c.y + (double)a.y * c.x;
  }
  const double sc = scale[j];
  const C2<TO> o = {(TO)(acc.x * sc), (TO)(acc.y * sc)};
  out[q] = o;
  if (rout) {
    const double wt = W[q / k];
    rout[q] = {(TR)(wt * (double)o.x), (TR)(wt * (double)o.y)};
  }
}

// k_gs_update for an even k, two adjacent slots per thread: one 16-byte load
// of FP32 basis data per vector instead of two 8-byte ones; per slot the
// same operations in the same order
template <class TV, class TW, class TO, class TR>
__global__ void k_gs_update2(int64_t n, int k, int nvec, const C2<TV> *__restrict__ V, int64_t vstride,
                             const Z2 *__restrict__ h, const double *__restrict__ scale, const C2<TW> *w, C2<TO> *out,
                             const double *__restrict__ W, C2<TR> *__restrict__ rout) {
  const int64_t q0 = 2 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (q0 >= n * k) return;
  const int j0 = (int)(q0 % k);
  Z2 acc[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) acc[e] = {(double)w[q0 + e].x, (double)w[q0 + e].y};
#pragma unroll 4
  for (int t = 0; t < nvec; ++t) {
    const float4 v4 = *reinterpret_cast<const float4 *>(V + (int64_t)t * vstride + q0);
    const C2<TV> a[2] = {{(TV)v4.x, (TV)v4.y}, {(TV)v4.z, (TV)v4.w}};
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const Z2 c = h[t * k + j0 + e];
      acc[e].x -= (double)a[e].x * c.x - (double)a[e].y * c.y;
      acc[e].y -= (double)a[e].x * c.y + (double)a[e].y * c.x;
    }
  }
  const double wt = rout ? W[q0 / k] : 0.;
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const double sc = scale[j0 + e];
    const C2<TO> o = {(TO)(acc[e].x * sc), (TO)(acc[e].y * sc)};
    out[q0 + e] = o;
    if (rout) rout[q0 + e] = {(TR)(wt * (double)o.x), (TR)(wt * (double)o.y)};
  }
}

// x += sum_t Z_t y_t (per column)
template <class TZ>
__global__ void k_xupdate(int64_t n, int k, int nvec, const C2<TZ> *__restrict__ Zb, int64_t zstride,
                          const Z2 *__restrict__ y, Z2 *__restrict__ x) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * k) return;
  int j = (int)(q % k);
  Z2 acc = x[q];
#pragma unroll 4
  for (int t = 0; t < nvec; ++t) {
    const C2<TZ> a = Zb[(int64_t)t * zstride + q];
    const Z2 c = y[t * k + j];
    acc.x += (double)a.x * c.x - (double)a.y * c.y;
    acc.y += (double)a.x * c.y + (double)a.y * c.x;
  }
  x[q] = acc;
}

// out = w * inv_j per column, and rout = W out (the first preconditioner input)
template <class TO, class TR>
__global__ void k_scale_cols(int64_t n, int k, const Z2 *__restrict__ w, const double *__restrict__ inv,
                             C2<TO> *__restrict__ out, const double *__restrict__ W, C2<TR> *__restrict__ rout) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * k) return;
  const double s = inv[q % k];
  const C2<TO> o = {(TO)(w[q].x * s), (TO)(w[q].y * s)};
  out[q] = o;
  const double wt = W[q / k];
  rout[q] = {(TR)(wt * (double)o.x), (TR)(wt * (double)o.y)};
}



// dst[i*k + j] = src[i*k + j] for every column j with mask[j]
__global__ void k_copy_cols(int64_t n, int k, const int32_t *__restrict__ mask, const Z2 *__restrict__ src,
                            Z2 *__restrict__ dst) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * k) return;
  if (mask[q % k]) dst[q] = src[q];
}

// v[i*k + j] = 0 for every column j with mask[j]
__global__ void k_zero_cols(int64_t n, int k, const int32_t *__restrict__ mask, Z2 *__restrict__ v) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * k) return;
  if (mask[q % k]) v[q] = {0., 0.};
}

// b[row*k + j] = (0, -((dir * f) * 2) * pi) at the port's edges (src/generateStiff.cpp:583-593)
__global__ void k_rhs(int64_t nent, int k, const int64_t *__restrict__ row, const int32_t *__restrict__ col,
                      const double *__restrict__ val, Z2 *__restrict__ b) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nent) return;
  b[row[q] * k + col[q]] = {0., val[q]};
}

// x = u + G phi + Gx a at selected rows (idx) or at every row (idx == null)
__global__ void k_assemble_x(int64_t nidx, const int64_t *__restrict__ idx, int k, const int64_t *__restrict__ grp,
                             const int32_t *__restrict__ gci, const double *__restrict__ gv,
                             const int64_t *__restrict__ xrp, const int32_t *__restrict__ xci,
                             const double *__restrict__ xv, const Z2 *__restrict__ u, const Z2 *__restrict__ phi,
                             const Z2 *__restrict__ a, Z2 *__restrict__ out) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t r = q / k;
  int j = (int)(q - r * k);
  if (r >= nidx) return;
  int64_t i = idx ? idx[r] : r;
  Z2 x = u[i * k + j];
  for (int64_t p = grp[i]; p < grp[i + 1]; ++p) {
    const Z2 f = phi[(int64_t)gci[p] * k + j];
    x.x += gv[p] * f.x;
    x.y += gv[p] * f.y;
  }
  for (int64_t p = xrp[i]; p < xrp[i + 1]; ++p) {
    const Z2 f = a[(int64_t)xci[p] * k + j];
    x.x += xv[p] * f.x;
    x.y += xv[p] * f.y;
  }
  out[r * k + j] = x;
}

struct DevBuf {
  Arena A;
  ~DevBuf() { A.free_all(); }
};

int pow2_at_least(int x) { int p = 1; while (p < x) p <<= 1; return p; }

OuterOp outer_op(gp_layout *lay, const Hierarchy &H) {
  OuterOp op;
  const CsrD &S = lay->S;
  const Csr &G = H.L[0].G, &Gx = H.aux.Gx;
  op.n = lay->n;
  op.rp = S.rp; op.ci = S.ci; op.sv = S.v;
  op.grp = G.rp; op.gci = G.ci; op.gv = G.v[0];
  op.xrp = Gx.rp; op.xci = Gx.ci; op.xv = Gx.v[0];
  op.sgrp = H.SG.rp; op.sgci = H.SG.ci; op.sgv = H.SG.v[0];
  op.sxrp = H.SGx.rp; op.sxci = H.SGx.ci; op.sxv = H.SGx.v[0];
  op.epsr = H.epsr; op.sig = H.sig; op.sdiag = H.sdiag;
  return op;
}

// Outer operator launch: the SELL kernel when S is large (sell_min_rows), else
// the (row, slot) CSR kernel.  out in TO: FP64 for residuals, the basis
// precision for the Arnoldi vector w = A z.
template <class TZ, class TO = double>
void outer(const OuterOp &op, const Hierarchy &H, int64_t smin, int k, const double *om, const C2<TZ> *u,
           const C2<TZ> *phi, const C2<TZ> *a, const Z2 *b, C2<TO> *out, int mode, int exact, hipStream_t s) {
  if (op.n < smin || H.Ss.nr != op.n) {
    k_outer<TZ, TO><<<blocks_for(op.n * k), 256, 0, s>>>(op, k, om, u, phi, a, b, out, mode, exact);
    return;
  }
  SellV<double> SS;
  SS.nr = H.Ss.nr; SS.nslice = H.Ss.nslice; SS.base = H.Ss.base; SS.w = H.Ss.w; SS.ci = H.Ss.ci; SS.nv = 1;
  SS.v[0] = H.Ss.v[0];
  SS.rows = H.Ss.rows;
  const int sp = sizeof(TZ) == 4 ? ((k % 4 == 0) ? 4 : (k % 2 == 0) ? 2 : 1) : ((k % 2 == 0) ? 2 : 1);
  int lpr = 1;
  while (lpr < 16 && (k / sp) % (lpr * 2) == 0) lpr *= 2;
  const dim3 grid((unsigned)((op.n * lpr + 255) / 256), (unsigned)(k / (sp * lpr)));
#define GP_OUTER_SELL(SP, LP) \
  k_outer_sell<TZ, TO, SP, LP><<<grid, 256, 0, s>>>(op, SS, k, om, u, phi, a, b, out, mode, exact)
  if (sp == 4) {
    if (lpr == 16) GP_OUTER_SELL(4, 16);
    else if (lpr == 8) GP_OUTER_SELL(4, 8);
    else if (lpr == 4) GP_OUTER_SELL(4, 4);
    else if (lpr == 2) GP_OUTER_SELL(4, 2);
    else GP_OUTER_SELL(4, 1);
  } else if (sp == 2) {
    if (lpr == 16) GP_OUTER_SELL(2, 16);
    else if (lpr == 8) GP_OUTER_SELL(2, 8);
    else if (lpr == 4) GP_OUTER_SELL(2, 4);
    else if (lpr == 2) GP_OUTER_SELL(2, 2);
    else GP_OUTER_SELL(2, 1);
  } else {
    if (lpr == 16) GP_OUTER_SELL(1, 16);
    else if (lpr == 8) GP_OUTER_SELL(1, 8);
    else if (lpr == 4) GP_OUTER_SELL(1, 4);
    else if (lpr == 2) GP_OUTER_SELL(1, 2);
    else GP_OUTER_SELL(1, 1);
  }
#undef GP_OUTER_SELL
}

// Column-batched dot products on the device, copied to the host:
// h[t*k + j] = V_t^H w for t < nvec, and h[nvec*k + j] = |w_j|^2 if self.
struct Dots {
  int nb = 1024;  // 4096 wavefronts: four per SIMD on 256 CUs
  Z2 *partial = nullptr, *hd = nullptr;
  gp_status init(Arena &A, int k, int maxv) {
    GP_TRY(A.alloc(&partial, (int64_t)nb * (maxv + 1) * k));
    GP_TRY(A.alloc(&hd, (int64_t)(maxv + 1) * k));
    return GP_OK;
  }
  template <class TV, class TW>
  gp_status run(hipStream_t s, int64_t n, int k, const C2<TV> *V, int64_t vstride, int nvec, const C2<TW> *w, int self,
                std::vector<cd> &h) {
    const int nout = nvec + self;
    const int nch = std::max(1, (nvec + kDotChunk - 1) / kDotChunk);
    k_mdots<TV, TW><<<dim3(nb, nch), 256, 0, s>>>(n, k, V, vstride, nvec, w, self, partial);
    k_mdots_final<<<(nout * k + 3) / 4, 256, 0, s>>>(nout, nb, k, partial, hd);
    std::vector<Z2> tmp((size_t)nout * k);
    GP_HIP(hipMemcpyAsync(tmp.data(), hd, sizeof(Z2) * nout * k, hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    h.resize(tmp.size());
    for (size_t q = 0; q < tmp.size(); ++q) h[q] = cd(tmp[q].x, tmp[q].y);
    return GP_OK;
  }
};

struct Column { int32_t port; double freq; int64_t id; };

// Where the slots take their columns from and hand the results back to.
struct ColumnIO {
  virtual ~ColumnIO() {}
  // 1: c filled; 0: nothing waiting right now; -1: no further columns.
  // block: every slot is idle, so the call may wait for work to arrive
  virtual int next(Column &c, bool block) = 0;
  virtual int64_t waiting() = 0;
  virtual bool want_x() const = 0;
  // xport: port-edge values (Construct_Z_V0_Vh order); x: full solution
  // (n values at stride xs) or null
  virtual gp_status done(const Column &c, const std::vector<cd> &xport, const Z2 *x, int64_t xs,
                         const gp_column_report &r) = 0;
};

struct SolveCtx {
  gp_layout *lay;
  Solver *sv;
  const gp_solve_opts *o;
  gp_solve_stats *st;
  std::chrono::steady_clock::time_point t0;
};

// Port-edge rows of every port's first side, in Construct_Z_V0_Vh order.
std::vector<int64_t> port_rows(gp_layout *lay) {
  const gp_mesh &mesh = lay->hm.view;
  std::vector<int64_t> map(lay->g.Ne), idx;
  gp_edge_map(&mesh, map.data(), nullptr);
  for (int32_t p = 0; p < mesh.num_ports; ++p) {
    int32_t s0 = mesh.port_side_ptr[p];
    for (int64_t q = mesh.side_edge_ptr[s0]; q < mesh.side_edge_ptr[s0 + 1]; ++q) idx.push_back(map[mesh.side_edges[q]]);
  }
  return idx;
}

// The batched solve's device buffers for k slots and restart m (see
// SlotWorkBase): the solution in split form, the right-hand sides, the Krylov
// basis V (m+1 blocks) and the preconditioned blocks Zu, Zp, Za (m blocks),
// the FP64 restart residual w and the Arnoldi vector wa, per-slot scalars,
// the Gram-Schmidt reduction buffers and the port-edge gather lists.
template <class T, class TV>
struct SlotWork : SlotWorkBase {
  VWork<T> vw;
  Arena A;
  Z2 *b = nullptr, *xu = nullptr, *xphi = nullptr, *xa = nullptr, *w = nullptr, *yd = nullptr;
  C2<TV> *V = nullptr, *wa = nullptr;
  C2<T> *Zu = nullptr, *Zp = nullptr, *Za = nullptr;
  double *om_d = nullptr, *sc_d = nullptr;
  int32_t *mask_d = nullptr;
  Dots dots;
  std::vector<int64_t> prow, map;
  int64_t *prow_d = nullptr;
  Z2 *xo = nullptr, *xfull = nullptr;
  int64_t max_port_edges = 1;
  int64_t *rhs_row = nullptr; int32_t *rhs_col = nullptr; double *rhs_val = nullptr;
  // the preconditioner of Krylov step i captured in a HIP graph (built on the
  // step's first use); gkey identifies what the captured launches depend on
  std::vector<hipGraphExec_t> gex;
  uint64_t gkey = 0;
  bool graph_off = false;        // capture failed once: launch directly
  void drop_graphs() {
    for (hipGraphExec_t &g : gex)
      if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
  }
  ~SlotWork() override { drop_graphs(); vwork_free(&vw); A.free_all(); }
};

// What a captured preconditioner depends on besides the workspace: the
// smoothing options, the lanes cap and the coarse-level table (re-allocated
// when the cache grows).
uint64_t graph_key(const gp_solve_opts &o, const void *tab, int cap, int lanes) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
  mix((uint64_t)o.smooth_edge); mix((uint64_t)o.smooth_aux); mix((uint64_t)o.aux_post);
  mix((uint64_t)o.sell_min_rows); mix((uint64_t)(uintptr_t)tab); mix((uint64_t)cap); mix((uint64_t)lanes);
  mix((uint64_t)o.edge_lds);
  return h;
}

template <class T, class TV>
gp_status slot_work(SolveCtx &cx, int k, int m, bool want_x, SlotWork<T, TV> **out) {
  Solver *sv = cx.sv;
  if (sv->ws && sv->ws->k == k && sv->ws->m == m && sv->ws->tsz == (int)sizeof(T) &&
      sv->ws->tvsz == (int)sizeof(TV) && sv->ws->want_x == want_x) {
    *out = static_cast<SlotWork<T, TV> *>(sv->ws);
    return GP_OK;
  }
  release_workspace(sv);
  gp_layout *lay = cx.lay;
  const Hierarchy &H = sv->H;
  hipStream_t s = lay->stream;
  const int64_t n = lay->n, nn = H.L[0].nn, na = H.aux.naux;
  const gp_mesh &mesh = lay->hm.view;
  auto *W = new SlotWork<T, TV>();
  W->k = k; W->m = m; W->tsz = (int)sizeof(T); W->tvsz = (int)sizeof(TV); W->want_x = want_x;
  auto st = [&]() -> gp_status {
    GP_TRY(vwork_create<T>(lay, H, k, &W->vw));
    Arena &A = W->A;
    GP_TRY(A.alloc(&W->b, n * k)); GP_TRY(A.alloc(&W->xu, n * k)); GP_TRY(A.alloc(&W->xphi, nn * k));
    GP_TRY(A.alloc(&W->xa, na * k));
    GP_TRY(A.alloc(&W->V, (int64_t)(m + 1) * n * k)); GP_TRY(A.alloc(&W->w, n * k));
    // the Arnoldi vector w = A z is stored in the basis precision (its products
    // and sums are FP64); w itself holds the FP64 restart residuals
    if constexpr (std::is_same<TV, double>::value) W->wa = W->w;
    else GP_TRY(A.alloc(&W->wa, n * k));
    GP_TRY(A.alloc(&W->Zu, (int64_t)m * n * k)); GP_TRY(A.alloc(&W->Zp, (int64_t)m * nn * k));
    GP_TRY(A.alloc(&W->Za, (int64_t)m * na * k));
    GP_TRY(A.alloc(&W->om_d, k)); GP_TRY(A.alloc(&W->sc_d, k)); GP_TRY(A.alloc(&W->mask_d, k));
    GP_TRY(A.alloc(&W->yd, (int64_t)m * k));
    GP_TRY(W->dots.init(A, k, m + 1));
    W->prow = port_rows(lay);
    W->map.resize(lay->g.Ne);
    gp_edge_map(&mesh, W->map.data(), nullptr);
    GP_TRY(A.alloc(&W->prow_d, (int64_t)W->prow.size())); GP_TRY(A.alloc(&W->xo, (int64_t)W->prow.size() * k));
    GP_HIP(hipMemcpyAsync(W->prow_d, W->prow.data(), sizeof(int64_t) * W->prow.size(), hipMemcpyHostToDevice, s));
    if (want_x) GP_TRY(A.alloc(&W->xfull, n * k));
    // right-hand-side entries of newly placed columns: at most k times the
    // largest port's edge count, allocated once (a free per refill would
    // synchronise the device and stall other layouts' streams)
    for (int32_t p = 0; p < mesh.num_ports; ++p) {
      int64_t c = 0;
      for (int32_t sd = mesh.port_side_ptr[p]; sd < mesh.port_side_ptr[p + 1]; ++sd)
        c += mesh.side_edge_ptr[sd + 1] - mesh.side_edge_ptr[sd];
      W->max_port_edges = std::max(W->max_port_edges, c);
    }
    GP_TRY(A.alloc(&W->rhs_row, W->max_port_edges * k)); GP_TRY(A.alloc(&W->rhs_col, W->max_port_edges * k));
    GP_TRY(A.alloc(&W->rhs_val, W->max_port_edges * k));
    GP_HIP(hipStreamSynchronize(s));
    return GP_OK;
  }();
  if (st != GP_OK) { delete W; return st; }
  W->bytes = W->A.bytes + W->vw.arena.bytes;
  sv->ws = W;
  *out = W;
  return GP_OK;
}

// Solve the columns io provides on k slots with preconditioner precision T
// and the Arnoldi basis stored in TV.  Residuals, operator products, dot
// products, the Hessenberg matrix and the iterate are FP64 in every case; an
// FP32 basis (TV = float) halves the bytes the Gram-Schmidt passes read, and
// every restart begins from the explicit FP64 residual, so the basis
// precision bounds only what one cycle can gain (restarted GMRES with a
// low-precision inner solve, Lindquist, Luszczek and Dongarra, 2020).
// Second Gram-Schmidt pass when a slot's first pass left less than kReorth
// of |w|^2 (severe cancellation).  The pass reads every basis block for all
// slots, and the textbook trigger (1/2) fired for some converging slot in
// nearly every step: at 200x200x24 the solve ran 15.1 solves/s with 1/2 and
// 17.2 with 1e-4, at the same iteration counts and residuals.
constexpr double kReorth = 1e-4;

template <class T, class TV>
gp_status solve_slots(SolveCtx &cx, ColumnIO &io, int k) {
  gp_layout *lay = cx.lay;
  const Hierarchy &H = cx.sv->H;
  const gp_solve_opts &o = *cx.o;
  hipStream_t s = lay->stream;
  const int64_t n = lay->n, nn = H.L[0].nn, na = H.aux.naux;
  const int m = std::min(std::max(2, o.restart), 64);
  const int nu_edge = o.smooth_edge > 0 ? std::min(o.smooth_edge, 8) : 1;
  const gp_mesh &mesh = lay->hm.view;
  const OuterOp op = outer_op(lay, H);
  const Csr &G0 = H.L[0].G, &Gx = H.aux.Gx;

  if (std::is_same<T, float>::value && o.edge_lds)
    GP_TRY(build_edge_march(lay, cx.sv->arena, &cx.sv->H, s));
  SlotWork<T, TV> *W = nullptr;
  GP_TRY((slot_work<T, TV>(cx, k, m, io.want_x(), &W)));
  VWork<T> &vw = W->vw;
  vw.om64 = W->om_d;
  // the preconditioner of each Krylov step replays a captured graph, except
  // while the kernel timer records per-launch events (replays would overwrite
  // them)
  const bool use_graphs = o.graphs != 0 && !lay->ktimer.on;
  const int lanes_now = set_lanes_cap(0);
  set_lanes_cap(lanes_now);
  vw.nu_aux = o.smooth_aux > 0 ? std::min(o.smooth_aux, 8) : 1;
  vw.aux_post = o.aux_post > 0;
  vw.kt = &lay->ktimer;
  vw.sell_min_rows = o.sell_min_rows > 0 ? o.sell_min_rows : 1;
  vw.edge_march = o.edge_lds != 0;
  Z2 *b = W->b, *xu = W->xu, *xphi = W->xphi, *xa = W->xa, *w = W->w, *yd = W->yd;
  C2<TV> *V = W->V, *wa = W->wa;
  C2<T> *Zu = W->Zu, *Zp = W->Zp, *Za = W->Za;
  double *om_d = W->om_d, *sc_d = W->sc_d;
  int32_t *mask_d = W->mask_d;
  Dots &dots = W->dots;
  GP_HIP(hipMemsetAsync(b, 0, sizeof(Z2) * n * k, s));
  GP_HIP(hipMemsetAsync(xu, 0, sizeof(Z2) * n * k, s));
  GP_HIP(hipMemsetAsync(xphi, 0, sizeof(Z2) * nn * k, s));
  GP_HIP(hipMemsetAsync(xa, 0, sizeof(Z2) * na * k, s));
  const std::vector<int64_t> &prow = W->prow, &map = W->map;
  int64_t *prow_d = W->prow_d;
  Z2 *xo = W->xo, *xfull = W->xfull;
  const int64_t max_port_edges = W->max_port_edges;
  int64_t *rhs_row = W->rhs_row; int32_t *rhs_col = W->rhs_col; double *rhs_val = W->rhs_val;
  // coarse-cache entries held by the slots (released when a slot takes
  // another column, and all of them when the call ends)
  std::vector<int32_t> pinned(k, -1);
  struct PinGuard {
    CoarseCache &c; std::vector<int32_t> &p;
    ~PinGuard() { for (int32_t &q : p) { coarse_cache_unpin(c, q); q = -1; } }
  } pg{cx.sv->cc, pinned};

  const unsigned gnk = blocks_for(n * k), gnnk = blocks_for(nn * k), gnak = blocks_for(na * k);
  // the paired update reads 16-byte pieces of an FP32 basis: k even
  const bool gs2 = sizeof(TV) == 4 && k % 2 == 0;
  const unsigned gnk2 = blocks_for((n * k + 1) / 2);
  // slot state
  std::vector<int32_t> cid(k, -1), its(k, 0), fsl(k, 0);
  std::vector<Column> ccol(k);
  std::vector<double> t_in(k, 0.);   // when each slot's column entered (ms since the call)
  auto now_ms = [&]() { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - cx.t0).count(); };
  std::vector<double> bnorm(k, 0.), omega(k, 2. * kPi * 1e9), beta(k, 0.);
  bool closed = false;  // the source will deliver no further columns
  std::vector<cd> Hm((size_t)(m + 1) * m * k), cs((size_t)m * k), sn((size_t)m * k), g((size_t)(m + 1) * k), y((size_t)m * k);
  std::vector<double> sch(k);
  std::vector<cd> h, h2;
  std::vector<Z2> yh((size_t)m * k);
  int cycles = 0;

  // place new columns into the slots listed in mask (idle when none comes).
  // With block, only the first column is waited for: the other slots take
  // what is already waiting.  Waiting for every slot left the columns already
  // placed unsolved until further columns arrived, so the last request to an
  // idle layout never finished.
  auto refill = [&](const std::vector<int32_t> &mask, bool block) -> gp_status {
    std::vector<int64_t> row;
    std::vector<int32_t> col;
    std::vector<double> val;
    bool placed = false;
    for (int j = 0; j < k; ++j) {
      if (!mask[j]) continue;
      Column c;
      const int got = closed ? -1 : io.next(c, block && !placed);
      if (got == 1) placed = true;
      if (got < 0) closed = true;
      if (got != 1) { cid[j] = -1; bnorm[j] = 0.; continue; }
      ccol[j] = c;
      cid[j] = 1; its[j] = 0; t_in[j] = now_ms();
      const double f = c.freq;
      omega[j] = 2. * kPi * f;
      {
        std::vector<int32_t> sl;
        GP_TRY(coarse_cache_lookup(H, cx.sv->cc, std::vector<double>{omega[j]}, &sl, s, 1));
        coarse_cache_unpin(cx.sv->cc, pinned[j]);
        pinned[j] = fsl[j] = sl[0];
      }
      double b2 = 0.;
      for (int32_t sd = mesh.port_side_ptr[c.port]; sd < mesh.port_side_ptr[c.port + 1]; ++sd)
        for (int64_t q = mesh.side_edge_ptr[sd]; q < mesh.side_edge_ptr[sd + 1]; ++q) {
          const double v = 0. - (double)mesh.side_dir[sd] * f * 2. * kPi;
          row.push_back(map[mesh.side_edges[q]]);
          col.push_back(j);
          val.push_back(v);
          b2 += v * v;
        }
      bnorm[j] = std::sqrt(b2);
    }
    GP_HIP(hipMemcpyAsync(mask_d, mask.data(), sizeof(int32_t) * k, hipMemcpyHostToDevice, s));
    k_zero_cols<<<gnk, 256, 0, s>>>(n, k, mask_d, b);
    k_zero_cols<<<gnk, 256, 0, s>>>(n, k, mask_d, xu);
    k_zero_cols<<<gnnk, 256, 0, s>>>(nn, k, mask_d, xphi);
    k_zero_cols<<<gnak, 256, 0, s>>>(na, k, mask_d, xa);
    if (!row.empty()) {
      if ((int64_t)row.size() > max_port_edges * k) return fail(GP_ERR_STATE, "right-hand side larger than planned");
      GP_HIP(hipMemcpyAsync(rhs_row, row.data(), sizeof(int64_t) * row.size(), hipMemcpyHostToDevice, s));
      GP_HIP(hipMemcpyAsync(rhs_col, col.data(), sizeof(int32_t) * row.size(), hipMemcpyHostToDevice, s));
      GP_HIP(hipMemcpyAsync(rhs_val, val.data(), sizeof(double) * row.size(), hipMemcpyHostToDevice, s));
      k_rhs<<<blocks_for((int64_t)row.size()), 256, 0, s>>>((int64_t)row.size(), k, rhs_row, rhs_col, rhs_val, b);
      GP_HIP(hipStreamSynchronize(s));
    }
    GP_HIP(hipMemcpyAsync(om_d, omega.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
    // idle slots still run the V-cycle (on zero vectors) and index the coarse
    // table: a slot that never held a column points at an entry another slot
    // pins, so it never reads an evicted entry
    {
      int32_t live = -1;
      for (int j = 0; j < k && live < 0; ++j) live = pinned[j];
      if (live >= 0)
        for (int j = 0; j < k; ++j)
          if (pinned[j] < 0) fsl[j] = live;
    }
    GP_TRY(vwork_set_columns<T>(vw, omega.data(), fsl.data(), cx.sv->cc, s));
    return GP_OK;
  };

  // hand finished slots (mask) to the caller: port-edge values, optional full x, reports
  auto finish = [&](const std::vector<int32_t> &mask, const std::vector<double> &relres) -> gp_status {
    k_assemble_x<<<blocks_for((int64_t)prow.size() * k), 256, 0, s>>>((int64_t)prow.size(), prow_d, k, G0.rp, G0.ci,
                                                                      G0.v[0], Gx.rp, Gx.ci, Gx.v[0], xu, xphi, xa, xo);
    std::vector<Z2> xh(prow.size() * k);
    GP_HIP(hipMemcpyAsync(xh.data(), xo, sizeof(Z2) * xh.size(), hipMemcpyDeviceToHost, s));
    std::vector<Z2> xf;
    if (io.want_x()) {
      k_assemble_x<<<gnk, 256, 0, s>>>(n, nullptr, k, G0.rp, G0.ci, G0.v[0], Gx.rp, Gx.ci, Gx.v[0], xu, xphi, xa, xfull);
      xf.resize(n * k);
      GP_HIP(hipMemcpyAsync(xf.data(), xfull, sizeof(Z2) * n * k, hipMemcpyDeviceToHost, s));
    }
    GP_HIP(hipStreamSynchronize(s));
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - cx.t0).count();
    std::vector<cd> xp(prow.size());
    for (int j = 0; j < k; ++j) {
      if (!mask[j] || cid[j] < 0) continue;
      for (size_t a = 0; a < prow.size(); ++a) xp[a] = cd(xh[a * k + j].x, xh[a * k + j].y);
      const bool ok = relres[j] <= o.rtol;
      GP_TRY(io.done(ccol[j], xp, io.want_x() ? xf.data() + j : nullptr, k,
                     gp_column_report{its[j], ok ? 1 : 0, relres[j], t_in[j], ms}));
      cx.st->columns += 1;
      cx.st->converged += ok ? 1 : 0;
      cx.st->max_iters_used = std::max(cx.st->max_iters_used, its[j]);
      cx.st->max_rel_residual = std::max(cx.st->max_rel_residual, relres[j]);
    }
    return GP_OK;
  };

  {  // initial fill
    std::vector<int32_t> all(k, 1);
    GP_TRY(refill(all, false));
  }
  for (;;) {
    {  // with every slot idle, wait for the source (its next() may block)
      bool any = false;
      for (int j = 0; j < k; ++j) any = any || cid[j] >= 0;
      while (!any && !closed) {
        std::vector<int32_t> all(k, 1);
        GP_TRY(refill(all, true));
        for (int j = 0; j < k; ++j) any = any || cid[j] >= 0;
      }
      if (!any) break;
    }
    // explicit FP64 residual of the split iterate (S G, S Gx included exactly)
    outer<double>(op, H, vw.sell_min_rows, k, om_d, xu, xphi, xa, b, w, 1, 1, s);
    GP_TRY(dots.run(s, n, k, V, 0, 0, w, 1, h));
    std::vector<double> relres(k, 0.);
    std::vector<int32_t> done(k, 0);
    int ndone = 0, nactive = 0;
    for (int j = 0; j < k; ++j) {
      beta[j] = std::sqrt(std::max(h[j].real(), 0.0));
      if (cid[j] < 0) continue;
      relres[j] = bnorm[j] > 0 ? beta[j] / bnorm[j] : 0.;
      if (relres[j] <= o.rtol || its[j] >= o.max_iters) { done[j] = 1; ++ndone; }
      else ++nactive;
    }
    if (ndone) GP_TRY(finish(done, relres));
    // refill the finished slots and, while columns may still come, the idle
    // ones (slots left without a column carry zero vectors)
    std::vector<int32_t> fill(k, 0);
    bool anyfill = false;
    for (int j = 0; j < k; ++j) {
      fill[j] = done[j] || (cid[j] < 0 && !closed && io.waiting() > 0);
      anyfill = anyfill || fill[j];
    }
    if (anyfill) {
      GP_TRY(refill(fill, false));
      // the refilled columns start from x = 0, so their residual b - A x is b
      // exactly; the other columns keep the residual formed above (their x
      // and b are unchanged), so no second FP64 operator pass is needed
      k_copy_cols<<<gnk, 256, 0, s>>>(n, k, mask_d, b, w);
      GP_TRY(dots.run(s, n, k, V, 0, 0, w, 1, h));
      nactive = 0;
      for (int j = 0; j < k; ++j) {
        beta[j] = cid[j] >= 0 ? std::sqrt(std::max(h[j].real(), 0.0)) : 0.;
        if (cid[j] >= 0) ++nactive;
      }
    }
    if (nactive == 0) continue;  // every slot idle: the wait at the top decides
    ++cycles;
    for (int j = 0; j < k; ++j) sch[j] = (cid[j] >= 0 && beta[j] > 0) ? 1.0 / beta[j] : 0.0;
    GP_HIP(hipMemcpyAsync(sc_d, sch.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
    k_scale_cols<TV, T><<<gnk, 256, 0, s>>>(n, k, w, sc_d, V, H.W, vw.r[0]);
    std::fill(g.begin(), g.end(), cd(0));
    for (int j = 0; j < k; ++j) g[j] = (cid[j] >= 0) ? beta[j] : 0.;
    int used = 0;
    for (int i = 0; i < m; ++i) {
      C2<TV> *Vn = V + (int64_t)(i + 1) * n * k;
      // preconditioner: (u, phi, a) = M(W v_i).  vw.r[0] already holds W v_i,
      // written with v_i by k_scale_cols (i = 0) or by the previous step's
      // k_gs_update
      // the preconditioner writes its level-0 outputs straight into this
      // step's Z blocks (same n*k interleaved layout), no copies
      C2<T> *zu = Zu + (int64_t)i * n * k, *zp = Zp + (int64_t)i * nn * k, *za = Za + (int64_t)i * na * k;
      C2<T> *own_u = vw.u[0], *own_phi = vw.phi[0], *own_a = vw.a;
      vw.u[0] = zu; vw.phi[0] = zp; vw.a = za;
      gp_status pst = GP_OK;
      if (use_graphs && !W->graph_off) {
        const uint64_t key = graph_key(o, vw.tab, vw.cap, lanes_now);
        if (key != W->gkey || (int)W->gex.size() != m) {
          W->drop_graphs();
          W->gex.assign(m, nullptr);
          W->gkey = key;
        }
        if (!W->gex[i]) {
          hipGraph_t g = nullptr;
          hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
          if (e == hipSuccess) {
            pst = precond<T>(H, vw, s, nu_edge, nu_edge);
            e = hipStreamEndCapture(s, &g);
          }
          if (e == hipSuccess && pst == GP_OK) e = hipGraphInstantiate(&W->gex[i], g, nullptr, nullptr, 0);
          if (g) (void)hipGraphDestroy(g);
          if (e != hipSuccess || pst != GP_OK) {   // no graph for this workspace: launch directly
            (void)hipGetLastError();
            W->gex[i] = nullptr;
            W->graph_off = true;
            lay->solver->graph_failed = true;
            pst = GP_OK;
          } else {
            ++lay->solver->graph_captures;
          }
        }
        if (W->gex[i]) {
          const hipError_t e = hipGraphLaunch(W->gex[i], s);
          if (e != hipSuccess) pst = fail(GP_ERR_HIP, "hipGraphLaunch of the preconditioner failed");
          ++lay->solver->graph_replays;
        } else {
          pst = precond<T>(H, vw, s, nu_edge, nu_edge);
        }
      } else {
        pst = precond<T>(H, vw, s, nu_edge, nu_edge);
      }
      vw.u[0] = own_u; vw.phi[0] = own_phi; vw.a = own_a;
      GP_TRY(pst);
      // w = A z (split form, FP64 products, stored in the basis precision)
      outer<T, TV>(op, H, vw.sell_min_rows, k, om_d, zu, zp, za, nullptr, wa, 0, 0, s);
      // classical Gram-Schmidt with the norm from the same pass; a second pass
      // only when the first one cancelled all but kReorth of |w|^2 in a slot
      GP_TRY(dots.run(s, n, k, V, n * k, i + 1, wa, 1, h));
      bool reorth = false;
      std::vector<double> rn2(k);
      for (int j = 0; j < k; ++j) {
        double hh = 0.;
        for (int t = 0; t <= i; ++t) hh += std::norm(h[t * k + j]);
        const double wn2 = h[(i + 1) * k + j].real();
        rn2[j] = wn2 - hh;
        if (cid[j] >= 0 && wn2 > 0 && rn2[j] < kReorth * wn2) reorth = true;
      }
      h2.assign((size_t)(i + 1) * k, cd(0));
      if (reorth) {
        std::fill(sch.begin(), sch.end(), 1.0);
        GP_HIP(hipMemcpyAsync(sc_d, sch.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
        if (gs2) k_gs_update2<TV, TV, TV, T><<<gnk2, 256, 0, s>>>(n, k, i + 1, V, n * k, dots.hd, sc_d, wa, wa, nullptr, nullptr);
        else k_gs_update<TV, TV, TV, T><<<gnk, 256, 0, s>>>(n, k, i + 1, V, n * k, dots.hd, sc_d, wa, wa, nullptr, nullptr);
        GP_TRY(dots.run(s, n, k, V, n * k, i + 1, wa, 1, h2));
        for (int j = 0; j < k; ++j) {
          double hh = 0.;
          for (int t = 0; t <= i; ++t) hh += std::norm(h2[t * k + j]);
          rn2[j] = h2[(i + 1) * k + j].real() - hh;
        }
      }
      bool all_in = true;
      for (int j = 0; j < k; ++j) {
        cd *Hc = &Hm[(size_t)j * (m + 1) * m];
        const double hn = std::sqrt(std::max(rn2[j], 0.0));
        sch[j] = (cid[j] >= 0 && hn > 0) ? 1.0 / hn : 0.0;
        if (cid[j] < 0) {
          for (int t = 0; t <= i + 1; ++t) Hc[t * m + i] = 0.;
          cs[i * k + j] = 1.; sn[i * k + j] = 0.;
          continue;
        }
        for (int t = 0; t <= i; ++t) Hc[t * m + i] = h[t * k + j] + (reorth ? h2[t * k + j] : cd(0));
        Hc[(i + 1) * m + i] = hn;
        // apply previous rotations, then form the new one
        for (int t = 0; t < i; ++t) {
          cd c = cs[t * k + j], sn_ = sn[t * k + j];
          cd a = Hc[t * m + i], bq = Hc[(t + 1) * m + i];
          Hc[t * m + i] = std::conj(c) * a + std::conj(sn_) * bq;
          Hc[(t + 1) * m + i] = -sn_ * a + c * bq;
        }
        cd a = Hc[i * m + i], bq = Hc[(i + 1) * m + i];
        double den = std::sqrt(std::norm(a) + std::norm(bq));
        cd c = den > 0 ? a / den : cd(1), sn_ = den > 0 ? bq / den : cd(0);
        cs[i * k + j] = c;
        sn[i * k + j] = sn_;
        Hc[i * m + i] = std::conj(c) * a + std::conj(sn_) * bq;
        Hc[(i + 1) * m + i] = 0;
        cd gi = g[i * k + j];
        g[i * k + j] = std::conj(c) * gi;
        g[(i + 1) * k + j] = -sn_ * gi;
        ++its[j];
        if (bnorm[j] > 0 && std::abs(g[(i + 1) * k + j]) > o.rtol * bnorm[j] && its[j] < o.max_iters) all_in = false;
      }
      // V_{i+1} = (w - V h) / |w - V h|, with the coefficients of the last pass
      // (dots.hd holds them: the second pass overwrote the first when it ran)
      GP_HIP(hipMemcpyAsync(sc_d, sch.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
      if (gs2) k_gs_update2<TV, TV, TV, T><<<gnk2, 256, 0, s>>>(n, k, i + 1, V, n * k, dots.hd, sc_d, wa, Vn, H.W, vw.r[0]);
      else k_gs_update<TV, TV, TV, T><<<gnk, 256, 0, s>>>(n, k, i + 1, V, n * k, dots.hd, sc_d, wa, Vn, H.W, vw.r[0]);
      used = i + 1;
      if (all_in) break;
      // with columns waiting, restart once half of the slots are done
      if (io.waiting() > 0) {
        int conv = 0;
        for (int j = 0; j < k; ++j)
          if (cid[j] < 0 || std::abs(g[(i + 1) * k + j]) <= o.rtol * bnorm[j] || its[j] >= o.max_iters) ++conv;
        if (2 * conv >= k) break;
      }
    }
    // y = R^-1 g per column, then x += Z y
    for (int j = 0; j < k; ++j) {
      cd *Hc = &Hm[(size_t)j * (m + 1) * m];
      for (int t = used - 1; t >= 0; --t) {
        cd acc = g[t * k + j];
        for (int q = t + 1; q < used; ++q) acc -= Hc[t * m + q] * y[q * k + j];
        y[t * k + j] = (cid[j] >= 0 && Hc[t * m + t] != cd(0)) ? acc / Hc[t * m + t] : cd(0);
      }
    }
    for (int t = 0; t < used; ++t)
      for (int j = 0; j < k; ++j) yh[t * k + j] = {y[t * k + j].real(), y[t * k + j].imag()};
    GP_HIP(hipMemcpyAsync(yd, yh.data(), sizeof(Z2) * used * k, hipMemcpyHostToDevice, s));
    k_xupdate<T><<<gnk, 256, 0, s>>>(n, k, used, Zu, n * k, yd, xu);
    k_xupdate<T><<<gnnk, 256, 0, s>>>(nn, k, used, Zp, nn * k, yd, xphi);
    k_xupdate<T><<<gnak, 256, 0, s>>>(na, k, used, Za, na * k, yd, xa);
    if (o.verbose) {
      double est = 0.;
      for (int j = 0; j < k; ++j)
        if (cid[j] >= 0 && bnorm[j] > 0) est = std::max(est, std::abs(g[used * k + j]) / bnorm[j]);
      std::fprintf(stderr, "[gp] k=%d cycle %d: %d steps, active %d, waiting %lld, est %.3e\n", k, cycles, used,
                   nactive, (long long)io.waiting(), est);
    }
  }
  GP_HIP(hipGetLastError());
  cx.st->batches += cycles;
  return GP_OK;
}

// Z column of one finished column (Construct_Z_V0_Vh on the first side of
// every port, src/fdtd.hpp:4133-4170): z[resp] = -sum x*leng*dir / area(src)
void z_column(gp_layout *lay, int32_t src, const std::vector<cd> &xp, cd *z) {
  const gp_mesh &mesh = lay->hm.view;
  const int32_t npt = mesh.num_ports;
  double cur = 0.;
  for (int32_t sd = mesh.port_side_ptr[src]; sd < mesh.port_side_ptr[src + 1]; ++sd) cur += mesh.side_area[sd];
  size_t a = 0;
  for (int32_t p = 0; p < npt; ++p) {
    const int32_t s0 = mesh.port_side_ptr[p];
    cd zz(0., 0.);
    for (int64_t q = mesh.side_edge_ptr[s0]; q < mesh.side_edge_ptr[s0 + 1]; ++q, ++a) {
      int dir; int64_t iz, ix, iy;
      lay->g.decode(mesh.side_edges[q], dir, iz, ix, iy);
      double leng = dir == 2 ? mesh.zn[iz + 1] - mesh.zn[iz] : dir == 1 ? mesh.xn[ix + 1] - mesh.xn[ix] : mesh.yn[iy + 1] - mesh.yn[iy];
      zz -= xp[a] * leng * (mesh.side_dir[s0] * 1.0);
    }
    z[p] = zz / cur;
  }
}

// Columns of one gp_solve_fullwave call: a queue (highest frequency first),
// results into the caller's Z / X arrays and per-column reports.
struct QueueIO : ColumnIO {
  gp_layout *lay;
  std::deque<Column> q;
  const int32_t *ports;
  int32_t np, npt;
  double *Z, *X;
  int64_t n;
  gp_column_report *rep;
  int next(Column &c, bool) override {
    if (q.empty()) return -1;
    c = q.front();
    q.pop_front();
    return 1;
  }
  int64_t waiting() override { return (int64_t)q.size(); }
  bool want_x() const override { return X != nullptr; }
  gp_status done(const Column &c, const std::vector<cd> &xp, const Z2 *x, int64_t xs,
                 const gp_column_report &r) override {
    const int64_t id = c.id, f = id / np, a = id % np;
    rep[id] = r;
    if (Z) {
      std::vector<cd> z(npt);
      z_column(lay, ports[a], xp, z.data());
      for (int32_t p = 0; p < npt; ++p) {
        const size_t zi = (size_t)f * npt * npt + p + (size_t)npt * ports[a];
        Z[2 * zi] = z[p].real();
        Z[2 * zi + 1] = z[p].imag();
      }
    }
    if (X && x)
      for (int64_t i = 0; i < n; ++i) {
        X[2 * (id * n + i)] = x[i * xs].x;
        X[2 * (id * n + i) + 1] = x[i * xs].y;
      }
    return GP_OK;
  }
};

// Columns from the caller's callbacks (gp_solve_stream).
struct CallbackIO : ColumnIO {
  gp_layout *lay;
  gp_pull_fn pull;
  gp_waiting_fn wait_fn;
  gp_push_fn push;
  void *user;
  int32_t npt;
  std::vector<gp_stream_column> inflight;  // by column id
  int next(Column &c, bool block) override {
    gp_stream_column sc{};
    for (;;) {
      const int r = pull(user, &sc, block ? 1 : 0);
      if (r != 1) return r < 0 ? -1 : 0;
      if (sc.port >= 0 && sc.port < npt && sc.freq > 0) break;
      // a malformed column comes straight back as a failed result
      gp_column_report bad{0, 0, -1.0, 0., 0.};
      push(user, &sc, nullptr, &bad);
    }
    c.port = sc.port;
    c.freq = sc.freq;
    c.id = (int64_t)inflight.size();
    inflight.push_back(sc);
    return 1;
  }
  int64_t waiting() override { return wait_fn ? wait_fn(user) : 0; }
  bool want_x() const override { return false; }
  gp_status done(const Column &c, const std::vector<cd> &xp, const Z2 *, int64_t, const gp_column_report &r) override {
    std::vector<cd> z(npt);
    z_column(lay, c.port, xp, z.data());
    std::vector<double> zz(2 * (size_t)npt);
    for (int32_t p = 0; p < npt; ++p) { zz[2 * p] = z[p].real(); zz[2 * p + 1] = z[p].imag(); }
    push(user, &inflight[c.id], zz.data(), &r);
    return GP_OK;
  }
};

// Device bytes one slot of the batched solve needs (basis, preconditioned
// basis, iterate, V-cycle vectors), for the choice of the slot count.
int64_t bytes_per_slot(const Hierarchy &H, int m, bool f32, bool v32) {
  const int64_t n = H.L[0].ne, nn = H.L[0].nn, na = H.aux.naux;
  const int64_t tz = f32 ? 8 : 16, tv = v32 ? 8 : 16;
  int64_t b = (int64_t)(m + 1) * n * tv + (int64_t)m * (n + nn + na) * tz + (n * 3 + nn + na) * 16;
  if (v32) b += n * tv;  // the Arnoldi vector in the basis precision
  for (const MgLevel &L : H.L) b += (5 * L.ne + 3 * L.nn) * tz;
  for (const AuxLevel &L : H.aux.L) b += 3 * L.n * tz;
  return b + (n + na) * tz;
}

// Preconditioner and basis precision from the options: FP32 preconditioner
// with an FP32 basis unless opts.basis asks for FP64; plain FP64 otherwise.
gp_status run_slots(SolveCtx &cx, ColumnIO &io, int k, bool f32) {
  if (!f32) return solve_slots<double, double>(cx, io, k);
  return cx.o->basis == 1 ? solve_slots<float, double>(cx, io, k) : solve_slots<float, float>(cx, io, k);
}

// Block width and lanes-per-row cap for the mixed path, measured: one
// FGMRES cycle of the real solve loop (restart m, the preconditioner, the
// outer operator and Gram-Schmidt) on k synthetic columns at the middle of the
// layout's frequency grid, for k = 8, 16 and 32 while the workspace fits half
// of the device memory, then the lanes-per-row caps 16, 4 and 2 at the best
// width.  The choice with the least time per column-iteration is kept with the
// solver.  The workspace and the coarse entry it leaves are reused by the
// solves that follow.
gp_status autotune(gp_layout *lay, const gp_solve_opts &o) {
  Solver *sv = lay->solver;
  Tuning &T = sv->tune;
  const auto t0 = std::chrono::steady_clock::now();
  const gp_mesh &mesh = lay->hm.view;
  const int m = std::min(std::max(2, o.restart), 64);
  const double fmid = gp_freq_of(&mesh, std::max(0, mesh.nfreq / 2));
  gp_solve_opts to = o;
  to.rtol = 1e-300;              // run the cycle to its end
  to.max_iters = m;
  to.verbose = 0;
  to.basis = 0;
  // trial solves launch directly: each trial sizes a new workspace, so a
  // captured graph would be timed once and never replayed
  to.graphs = 0;
  std::vector<int32_t> ports(1, 0);
  auto run = [&](int k, int lanes, double *us) -> gp_status {
    QueueIO io;
    io.lay = lay; io.ports = ports.data(); io.np = 1; io.npt = mesh.num_ports; io.Z = nullptr; io.X = nullptr;
    io.n = lay->n;
    std::vector<gp_column_report> rep(k);
    io.rep = rep.data();
    for (int j = 0; j < k; ++j) io.q.push_back({(int32_t)(j % std::max(1, mesh.num_ports)), fmid, (int64_t)0});
    gp_solve_stats st{};
    SolveCtx cx{lay, sv, &to, &st, std::chrono::steady_clock::now()};
    const int prev = set_lanes_cap(lanes);
    GP_HIP(hipStreamSynchronize(lay->stream));
    const auto a = std::chrono::steady_clock::now();
    const gp_status r = solve_slots<float, float>(cx, io, k);
    const gp_status q = hipStreamSynchronize(lay->stream) == hipSuccess ? GP_OK : fail(GP_ERR_HIP, "autotune sync");
    set_lanes_cap(prev);
    GP_TRY(r);
    GP_TRY(q);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
    if (us) *us = 1e3 * ms / ((double)k * m);
    return GP_OK;
  };
  size_t fre = 0, tot = 0;
  GP_HIP(hipMemGetInfo(&fre, &tot));
  const int64_t per = bytes_per_slot(sv->H, m, true, true);
  auto fits = [&](int k) { return (int64_t)k * per <= ((int64_t)fre + (sv->ws ? sv->ws->bytes : 0)) / 2; };
  // first launches of every kernel instance load code; keep them out of the timing
  GP_TRY(run(fits(8) ? 8 : 1, 16, nullptr));
  const int ks[3] = {8, 16, 32};
  double best = 1e300;
  T.k = 1;
  for (int q = 0; q < 3; ++q) {
    if (!fits(ks[q])) continue;
    GP_TRY(run(ks[q], 16, &T.us[q]));           // the workspace is allocated in a first pass
    GP_TRY(run(ks[q], 16, &T.us[q]));
    if (T.us[q] < best) { best = T.us[q]; T.k = ks[q]; }
  }
  const int caps[3] = {16, 4, 2};
  T.lanes = 16;
  T.lanes_us[0] = best;
  for (int q = 1; q < 3; ++q) {
    GP_TRY(run(T.k, caps[q], &T.lanes_us[q]));
    if (T.lanes_us[q] < best) { best = T.lanes_us[q]; T.lanes = caps[q]; }
  }
  T.done = true;
  T.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return GP_OK;
}

}  // namespace

}  // namespace gp

using namespace gp;

extern "C" gp_status gp_autotune(gp_layout *lay, int32_t restart, gp_autotune_info *out) {
  if (!lay) return fail(GP_ERR_ARG, "null layout");
  GP_TRY(gp_prepare_solver(lay, GP_PREC_MIXED));
  gp_solve_opts o;
  gp_solve_opts_default(&o);
  if (restart > 0) o.restart = restart;
  lay->solver->tune = Tuning{};
  GP_TRY(autotune(lay, o));
  if (out) {
    const Tuning &T = lay->solver->tune;
    out->slots = T.k;
    out->lanes_cap = T.lanes;
    for (int q = 0; q < 3; ++q) { out->slot_us[q] = T.us[q]; out->lanes_us[q] = T.lanes_us[q]; }
    out->tune_ms = T.ms;
  }
  return GP_OK;
}

extern "C" void gp_solve_opts_default(gp_solve_opts *o) {
  o->rtol = 1e-8;
  o->max_iters = 500;
  o->restart = 16;
  o->precision = GP_PREC_MIXED;
  o->batch = 0;
  o->verbose = 0;
  o->smooth_edge = 0;
  o->smooth_aux = 0;
  o->sell_min_rows = 0;
  o->aux_post = 0;
  o->basis = 0;
  o->graphs = 1;
  o->edge_lds = 0;
}

extern "C" gp_status gp_prepare_solver(gp_layout *lay, int32_t precision) {
  if (!lay) return fail(GP_ERR_ARG, "null layout");
  const bool want32 = precision == GP_PREC_MIXED;
  if (lay->solver && (!want32 || lay->solver->f32)) return GP_OK;
  if (lay->solver) {
    // an FP64 hierarchy gains the FP32 copies in place
    Solver *sv = lay->solver;
    GP_TRY(hierarchy_add_f32(lay, sv->arena, &sv->H, lay->stream));
    sv->f32 = true;
    return GP_OK;
  }
  auto *sv = new Solver();
  gp_status st = build_hierarchy(lay, sv->arena, &sv->H, want32, 12);
  if (st != GP_OK) { destroy_solver(sv); return st; }
  sv->f32 = want32;
  lay->solver = sv;
  return GP_OK;
}

extern "C" gp_status gp_coarse_aux_inverse(gp_layout *lay, double freq, int64_t *n, double *mat, double *inv) {
  if (!lay || !n) return fail(GP_ERR_ARG, "null argument");
  if (!lay->solver) return fail(GP_ERR_STATE, "solver not prepared");
  const Hierarchy &H = lay->solver->H;
  *n = H.aux.nL;
  if (!mat && !inv) return GP_OK;
  if (!(freq > 0)) return fail(GP_ERR_ARG, "frequency must be positive");
  std::vector<cd> m, v;
  GP_TRY(aux_coarse_dense(H, 2 * kPi * freq, &m, &v, lay->stream));
  if (mat) std::memcpy(mat, m.data(), sizeof(cd) * m.size());
  if (inv) std::memcpy(inv, v.data(), sizeof(cd) * v.size());
  return GP_OK;
}

extern "C" gp_status gp_coarse_aux_apply(gp_layout *lay, double freq, int32_t k, int32_t precision, const double *r,
                                         double *x) {
  if (!lay || !r || !x || k < 1 || k > 64 || !(freq > 0)) return fail(GP_ERR_ARG, "bad arguments (1 <= k <= 64)");
  if (!lay->solver) return fail(GP_ERR_STATE, "solver not prepared");
  const AuxAMG &X = lay->solver->H.aux;
  if (!X.eig) return fail(GP_ERR_STATE, "the auxiliary coarsest level is not in eigenvector form");
  const int64_t n = X.nL;
  hipStream_t s = lay->stream;
  DevBuf D;
  double *om = nullptr;
  Z2 *rd = nullptr, *xd = nullptr, *yp = nullptr;
  GP_TRY(D.A.alloc(&om, k)); GP_TRY(D.A.alloc(&rd, n * k)); GP_TRY(D.A.alloc(&xd, n * k)); GP_TRY(D.A.alloc(&yp, n * k));
  std::vector<double> omh(k, 2 * kPi * freq);
  GP_HIP(hipMemcpyAsync(om, omh.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(rd, r, sizeof(Z2) * n * k, hipMemcpyHostToDevice, s));
  if (precision == GP_PREC_MIXED) {
    std::vector<C2<float>> rf(n * k), xf(n * k);
    for (int64_t q = 0; q < n * k; ++q) rf[q] = {(float)r[2 * q], (float)r[2 * q + 1]};
    C2<float> *rfd = (C2<float> *)rd, *xfd = (C2<float> *)xd;
    GP_HIP(hipMemcpyAsync(rfd, rf.data(), sizeof(C2<float>) * n * k, hipMemcpyHostToDevice, s));
    GP_TRY(aux_eig_apply<float>(X, k, om, rfd, xfd, yp, s));
    GP_HIP(hipMemcpyAsync(xf.data(), xfd, sizeof(C2<float>) * n * k, hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    for (int64_t q = 0; q < n * k; ++q) { x[2 * q] = xf[q].x; x[2 * q + 1] = xf[q].y; }
  } else {
    GP_TRY(aux_eig_apply<double>(X, k, om, rd, xd, yp, s));
    GP_HIP(hipMemcpyAsync(x, xd, sizeof(Z2) * n * k, hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
  }
  GP_HIP(hipGetLastError());
  return GP_OK;
}

extern "C" gp_status gp_layout_trim(gp_layout *lay) {
  if (!lay) return fail(GP_ERR_ARG, "null layout");
  if (!lay->solver) return GP_OK;
  GP_HIP(hipStreamSynchronize(lay->stream));
  release_workspace(lay->solver);
  for (const CoarseEntry &x : lay->solver->cc.e)
    if (x.pins) return fail(GP_ERR_STATE, "gp_layout_trim: a solve on this layout is running");
  coarse_cache_free(lay->solver->cc);
  return GP_OK;
}

extern "C" gp_status gp_hierarchy_info_get(gp_layout *lay, gp_hierarchy_info *out) {
  if (!lay || !out) return fail(GP_ERR_ARG, "null argument");
  if (!lay->solver) return fail(GP_ERR_STATE, "solver not prepared");
  const Hierarchy &H = lay->solver->H;
  *out = gp_hierarchy_info{};
  out->edge_levels = (int32_t)std::min<size_t>(H.L.size(), GP_MAX_LEVELS);
  for (int l = 0; l < out->edge_levels; ++l) {
    out->ne[l] = H.L[l].ne; out->nn[l] = H.L[l].nn; out->nnz_k[l] = H.L[l].K.nnz;
  }
  if (!H.L.empty() && H.L.size() <= GP_MAX_LEVELS) out->nnz_k[H.L.size() - 1] = H.neL * H.neL;
  out->aux_levels = (int32_t)std::min<size_t>(H.aux.L.size(), GP_MAX_LEVELS);
  for (int l = 0; l < out->aux_levels; ++l) {
    out->na[l] = H.aux.L[l].n; out->nnz_n[l] = H.aux.L[l].N.nnz; out->nlong[l] = H.aux.L[l].N.nlong;
  }
  out->naux = H.aux.naux;
  out->setup_ms = H.setup_ms;
  const CoarseCache &cc = lay->solver->cc;
  out->coarse_entry_bytes = coarse_entry_bytes(H);
  out->coarse_entries = 0;
  for (const CoarseEntry &x : cc.e) out->coarse_entries += x.mem ? 1 : 0;
  out->coarse_evictions = cc.evictions;
  out->workspace_bytes = lay->solver->ws ? lay->solver->ws->bytes : 0;
  out->aux_coarse_blocks = H.aux.nblk;
  out->aux_coarse_max_block = H.aux.max_blk;
  out->aux_coarse_eigen = H.aux.eig ? 1 : 0;
  out->coarse_form_ms = cc.form_ms;
  out->graph_captures = lay->solver->graph_captures;
  out->graph_replays = lay->solver->graph_replays;
  out->graph_failed = lay->solver->graph_failed ? 1 : 0;
  out->march_block_rows = H.march.code ? H.march.maxb : 0;
  out->march_entries = H.march.entries;
  out->march_outside = H.march.outside;
  return GP_OK;
}

extern "C" gp_status gp_solve_fullwave_ex(gp_layout *lay, const double *freqs, int32_t nf, const int32_t *ports,
                                          int32_t np, const gp_solve_opts *opts, double *X, double *Z,
                                          gp_solve_stats *stats, gp_column_report *reports) {
  if (!lay || !freqs || nf < 1 || !ports || np < 1) return fail(GP_ERR_ARG, "bad solve arguments");
  gp_solve_opts o;
  if (opts) o = *opts; else gp_solve_opts_default(&o);
  if (!(o.rtol > 0) || o.max_iters < 1) return fail(GP_ERR_ARG, "rtol and max_iters must be positive");
  const gp_mesh &mesh = lay->hm.view;
  for (int32_t a = 0; a < np; ++a)
    if (ports[a] < 0 || ports[a] >= mesh.num_ports) return fail(GP_ERR_ARG, "port index out of range");
  for (int32_t f = 0; f < nf; ++f)
    if (!(freqs[f] > 0)) return fail(GP_ERR_ARG, "frequencies must be positive");
  auto t0 = std::chrono::steady_clock::now();
  GP_TRY(gp_prepare_solver(lay, o.precision));
  Solver *sv = lay->solver;
  const bool f32 = sv->f32 && o.precision == GP_PREC_MIXED;
  // coarsest-level inverses for every requested frequency
  std::vector<double> fr(freqs, freqs + nf), om(nf);
  for (int32_t f = 0; f < nf; ++f) om[f] = 2. * kPi * fr[f];
  std::vector<int32_t> cslot;
  if (nf <= sv->cc.max_entries / 2) GP_TRY(coarse_cache_lookup(sv->H, sv->cc, om, &cslot, lay->stream));
  auto t1 = std::chrono::steady_clock::now();
  gp_solve_stats st{};
  // column ids in request order, frequency-major: id = f*np + p.  They enter
  // the slots highest frequency first: those need the most iterations, and
  // started early they overlap the quick ones instead of forming the tail.
  const int ncol = nf * np;
  QueueIO io;
  io.lay = lay; io.ports = ports; io.np = np; io.npt = mesh.num_ports; io.Z = Z; io.X = X; io.n = lay->n;
  std::vector<gp_column_report> rep(ncol, gp_column_report{0, 0, 0., 0., 0.});
  io.rep = rep.data();
  std::vector<int32_t> forder(nf);
  for (int32_t f = 0; f < nf; ++f) forder[f] = f;
  std::stable_sort(forder.begin(), forder.end(), [&](int32_t a, int32_t b) { return fr[a] > fr[b]; });
  for (int32_t f : forder)
    for (int32_t a = 0; a < np; ++a) io.q.push_back({ports[a], fr[f], (int64_t)f * np + a});
  // slot count: requested, else 16 for the mixed path and 1 for plain FP64,
  // bounded by the columns and by half of the free device memory
  const int m = std::min(std::max(2, o.restart), 64);
  if (o.batch <= 0 && f32 && !sv->tune.done) GP_TRY(autotune(lay, o));
  int kb = o.batch > 0 ? o.batch : (f32 ? sv->tune.k : 1);
  kb = std::min(kb, 64);
  kb = std::min(kb, pow2_at_least(ncol));
  {
    // a workspace kept from an earlier call is reused or released, so its
    // bytes count as available
    size_t fre = 0, tot = 0;
    if (hipMemGetInfo(&fre, &tot) == hipSuccess) {
      const int64_t avail = (int64_t)fre + (sv->ws ? sv->ws->bytes : 0);
      const int64_t per = bytes_per_slot(sv->H, m, f32, f32 && o.basis != 1);
      while (kb > 1 && (int64_t)kb * per > avail / 2) kb /= 2;
    }
  }
  const int k = pow2_at_least(kb);
  const int32_t npt = mesh.num_ports;
  if (Z) std::fill(Z, Z + 2 * (size_t)nf * npt * npt, 0.0);
  SolveCtx cx{lay, sv, &o, &st, t0};
  const int cap_prev = set_lanes_cap(f32 && o.batch <= 0 ? sv->tune.lanes : 16);
  gp_status s2 = run_slots(cx, io, k, f32);
  set_lanes_cap(cap_prev);
  if (s2 != GP_OK) return s2;
  if (reports)
    for (int id = 0; id < ncol; ++id) reports[id] = rep[id];
  auto t2 = std::chrono::steady_clock::now();
  st.setup_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  st.solve_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
  if (stats) *stats = st;
  return st.converged == st.columns ? GP_OK : fail(GP_ERR_NOCONV, "some columns stopped above the tolerance");
}

extern "C" gp_status gp_solve_workspace_bytes(gp_layout *lay, int32_t precision, int32_t slots, int32_t restart,
                                              int64_t *bytes) {
  if (!lay || !bytes || slots < 1) return fail(GP_ERR_ARG, "bad arguments");
  GP_TRY(gp_prepare_solver(lay, precision));
  const int m = std::min(std::max(2, restart), 64);
  const int k = pow2_at_least(std::min<int32_t>(slots, 64));
  const bool f32 = lay->solver->f32 && precision == GP_PREC_MIXED;
  const CoarseCache &cc = lay->solver->cc;
  const int64_t coarse = std::min<int64_t>(cc.max_bytes, (int64_t)std::max(cc.max_entries, k) *
                                                         coarse_entry_bytes(lay->solver->H));
  *bytes = (int64_t)k * bytes_per_slot(lay->solver->H, m, f32, f32) + coarse;
  return GP_OK;
}

extern "C" gp_status gp_solve_stream(gp_layout *lay, const gp_solve_opts *opts, int32_t slots, gp_pull_fn pull,
                                     gp_waiting_fn waiting, gp_push_fn push, void *user, gp_solve_stats *stats) {
  if (!lay || !pull || !push || slots < 1) return fail(GP_ERR_ARG, "bad stream arguments");
  gp_solve_opts o;
  if (opts) o = *opts; else gp_solve_opts_default(&o);
  if (!(o.rtol > 0) || o.max_iters < 1) return fail(GP_ERR_ARG, "rtol and max_iters must be positive");
  auto t0 = std::chrono::steady_clock::now();
  GP_TRY(gp_prepare_solver(lay, o.precision));
  Solver *sv = lay->solver;
  const bool f32 = sv->f32 && o.precision == GP_PREC_MIXED;
  CallbackIO io;
  io.lay = lay; io.pull = pull; io.wait_fn = waiting; io.push = push; io.user = user; io.npt = lay->hm.view.num_ports;
  gp_solve_stats st{};
  SolveCtx cx{lay, sv, &o, &st, t0};
  const int k = pow2_at_least(std::min<int32_t>(slots, 64));
  gp_status s2 = run_slots(cx, io, k, f32);
  st.solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = st;
  return s2;
}

extern "C" gp_status gp_solve_fullwave(gp_layout *lay, const double *freqs, int32_t nf, const int32_t *ports, int32_t np,
                                       const gp_solve_opts *opts, double *X, double *Z, gp_solve_stats *stats) {
  return gp_solve_fullwave_ex(lay, freqs, nf, ports, np, opts, X, Z, stats, nullptr);
}

extern "C" gp_status gp_layout_sync(gp_layout *lay) {
  if (!lay) return fail(GP_ERR_ARG, "null layout");
  GP_HIP(hipStreamSynchronize(lay->stream));
  return GP_OK;
}

extern "C" gp_status gp_kernel_timer(gp_layout *lay, int32_t enable) {
  if (!lay) return fail(GP_ERR_ARG, "null layout");
  GP_HIP(hipStreamSynchronize(lay->stream));
  lay->ktimer.on = enable != 0;
  if (enable) lay->ktimer.reset();
  return GP_OK;
}

extern "C" gp_status gp_kernel_timer_read(gp_layout *lay, int64_t *launches, double *total_ms, double *total_bytes) {
  if (!lay || !launches || !total_ms || !total_bytes) return fail(GP_ERR_ARG, "null argument");
  GP_HIP(hipStreamSynchronize(lay->stream));
  const KTimer &t = lay->ktimer;
  double ms = 0, b = 0;
  for (size_t q = 0; q < t.used; ++q) {
    float e = 0;
    GP_HIP(hipEventElapsedTime(&e, t.ev[2 * q], t.ev[2 * q + 1]));
    ms += e;
    b += t.bytes[q];
  }
  *launches = (int64_t)t.used;
  *total_ms = ms;
  *total_bytes = b;
  return GP_OK;
}

extern "C" gp_status gp_apply_fullwave(gp_layout *lay, const double *freqs, int32_t k, const double *x, double *y) {
  if (!lay || !freqs || !x || !y || k < 1 || k > 64) return fail(GP_ERR_ARG, "gp_apply_fullwave: bad arguments (1 <= k <= 64)");
  for (int32_t j = 0; j < k; ++j)
    if (!(freqs[j] > 0.)) return fail(GP_ERR_ARG, "gp_apply_fullwave: frequencies must be positive");
  if (!lay->s_ready) return fail(GP_ERR_STATE, "gp_apply_fullwave: stiffness not assembled");
  GP_TRY(gp_prepare_solver(lay, GP_PREC_FP64));
  const Hierarchy &H = lay->solver->H;
  const int64_t n = lay->n;
  hipStream_t s = lay->stream;
  DevBuf D;
  Z2 *xc, *xr, *yr;
  double *om;
  GP_TRY(D.A.alloc(&xc, n * k)); GP_TRY(D.A.alloc(&xr, n * k)); GP_TRY(D.A.alloc(&yr, n * k)); GP_TRY(D.A.alloc(&om, k));
  std::vector<double> omh(k);
  for (int32_t j = 0; j < k; ++j) omh[j] = 2 * kPi * freqs[j];
  GP_HIP(hipMemcpyAsync(om, omh.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(xc, x, sizeof(Z2) * n * k, hipMemcpyHostToDevice, s));
  k_cols_to_rows<<<blocks_for(n * k), 256, 0, s>>>(n, k, xc, xr);
  amul(H, n, k, om, xr, yr, s);
  k_rows_to_cols<<<blocks_for(n * k), 256, 0, s>>>(n, k, yr, xc);
  GP_HIP(hipMemcpyAsync(y, xc, sizeof(Z2) * n * k, hipMemcpyDeviceToHost, s));
  GP_HIP(hipStreamSynchronize(s));
  return GP_OK;
}

extern "C" gp_status gp_bench_spmv(gp_layout *lay, int32_t k, int32_t precision, int32_t iters, double *avg_ms) {
  if (!lay || k < 1 || k > 64 || iters < 1 || !avg_ms) return fail(GP_ERR_ARG, "bad arguments");
  (void)precision;
  GP_TRY(gp_prepare_solver(lay, GP_PREC_FP64));
  const Hierarchy &H = lay->solver->H;
  const int64_t n = lay->n;
  hipStream_t s = lay->stream;
  DevBuf D;
  Z2 *x, *y;
  double *om;
  GP_TRY(D.A.alloc(&x, n * k)); GP_TRY(D.A.alloc(&y, n * k)); GP_TRY(D.A.alloc(&om, k));
  std::vector<double> omh(k, 2 * kPi * 1e9);
  GP_HIP(hipMemcpyAsync(om, omh.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemsetAsync(x, 0, sizeof(Z2) * n * k, s));
  hipEvent_t e0, e1;
  GP_HIP(hipEventCreate(&e0)); GP_HIP(hipEventCreate(&e1));
  amul(H, n, k, om, x, y, s);
  GP_HIP(hipEventRecord(e0, s));
  for (int it = 0; it < iters; ++it) amul(H, n, k, om, x, y, s);
  GP_HIP(hipEventRecord(e1, s));
  GP_HIP(hipEventSynchronize(e1));
  float ms = 0;
  GP_HIP(hipEventElapsedTime(&ms, e0, e1));
  *avg_ms = ms / iters;
  (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
  return GP_OK;
}

namespace gp {
// plain vector copy, 16 bytes per thread
__global__ void gp_copy_probe(int64_t n, const double2 *__restrict__ a, double2 *__restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}
}  // namespace gp

extern "C" gp_status gp_bench_copy(int64_t bytes, int32_t iters, double *avg_ms) {
  if (bytes < 16 || iters < 1 || !avg_ms) return fail(GP_ERR_ARG, "bad arguments");
  GP_TRY(ensure_device());
  const int64_t n = bytes / 16;
  DevBuf D;
  double2 *a, *b;
  GP_TRY(D.A.alloc(&a, n)); GP_TRY(D.A.alloc(&b, n));
  GP_HIP(hipMemset(a, 0, sizeof(double2) * n));
  hipEvent_t e0, e1;
  GP_HIP(hipEventCreate(&e0)); GP_HIP(hipEventCreate(&e1));
  gp_copy_probe<<<blocks_for(n), 256>>>(n, a, b);
  GP_HIP(hipEventRecord(e0, nullptr));
  for (int it = 0; it < iters; ++it) gp_copy_probe<<<blocks_for(n), 256>>>(n, a, b);
  GP_HIP(hipEventRecord(e1, nullptr));
  GP_HIP(hipEventSynchronize(e1));
  float ms = 0;
  GP_HIP(hipEventElapsedTime(&ms, e0, e1));
  *avg_ms = ms / iters;
  (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
  return GP_OK;
}

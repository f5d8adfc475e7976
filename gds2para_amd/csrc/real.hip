// real.hip -- real sparse solve replacing hypreSolve (src/hypreSolve.cpp:9-283,
// interface src/hypreSolver.h:14): A x = b for a real matrix given as
// row-sorted COO triplets, by flexible GMRES(30) in FP64 preconditioned with
// one V-cycle of smoothed-aggregation AMG built from the matrix alone.
//
// The reference runs HYPRE FlexGMRES (KDim 10, max 100 iterations, relative
// tolerance 1e-4) with BoomerAMG.  Here the tolerance is the caller's (1e-4
// when 0 is passed, the reference's value) and the AMG is:
//   * aggregation: two passes of strength-weighted pairwise matching (each
//     node pairs with its strongest unmatched strong neighbour when the
//     choice is mutual; j is a strong neighbour of i when
//     |a_ij| / sqrt|a_ii a_jj| is at least kTheta of the row's largest such
//     score), so aggregates hold up to four nodes;
//   * prolongation: the aggregates' indicator smoothed once with
//     4/(3 lambda) D^-1 A (lambda from power iteration), truncated (entries
//     below kTrunc of the row's largest dropped, the row's own aggregate
//     kept, the row sum restored), Galerkin R A P;
//   * smoothing, two matrix products each side: l1-Jacobi sweeps (the
//     default) or two steps of the Chebyshev recurrence on D^-1 A over
//     [0.3 b, b], b = 1.1 lambda (gp_solve_real_ex; the reference picks
//     BoomerAMG's relaxation with HYPRE_BoomerAMGSetRelaxType,
//     src/hypreSolve.cpp:230); dense FP64 inverse on the coarsest level
//     (<= 512 unknowns).
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "dense.hpp"
#include "real.hpp"
#include "sparse.hpp"

namespace gp {

namespace {

constexpr double kTheta = 0.25;
constexpr double kTrunc = 0.2;
constexpr int64_t kCoarsest = 512;
// Chebyshev smoothing: steps per side (the matrix products of two Jacobi
// sweeps) and its interval [kChebLower b, b] on the spectrum of D^-1 A, with
// b = kChebUpper times the power-iteration estimate of its largest eigenvalue
// (the estimate approaches it from below)
constexpr int kChebSteps = 2;
constexpr double kChebUpper = 1.1;
constexpr double kChebLower = 0.3;

// The setup kernels below give each row one wavefront (kWave lanes striding
// over its entries): Ad's conductor-aggregate rows couple to every dielectric
// node around a conductor, and one thread per row left each launch waiting
// on those few rows.  The wavefronts stride over the rows, so any n fits the
// grid of wave_blocks(n).
constexpr int kWave = 64;

unsigned wave_blocks(int64_t n) {
  return (unsigned)std::min<int64_t>(std::max<int64_t>((n * kWave + 255) / 256, 1), 1 << 20);
}

__device__ __forceinline__ int64_t wave_first() { return ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; }
__device__ __forceinline__ int64_t wave_count() { return ((int64_t)gridDim.x * blockDim.x) / kWave; }

__device__ __forceinline__ double wave_sum(double x) {
  for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

__device__ __forceinline__ double wave_max(double x) {
  for (int o = kWave / 2; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o));
  return x;
}

__global__ void k_r_diag(int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                         const double *__restrict__ v, double *__restrict__ d, double *__restrict__ dl1) {
  const int lane = threadIdx.x & (kWave - 1);
  for (int64_t i = wave_first(); i < n; i += wave_count()) {  // i is the same on every lane
    double a = 0., off = 0.;
    for (int64_t p = rp[i] + lane; p < rp[i + 1]; p += kWave) {
      if (ci[p] == (int32_t)i) a += v[p];
      else off += fabs(v[p]);
    }
    a = wave_sum(a);
    off = wave_sum(off);
    if (lane == 0) {
      d[i] = a;
      const double m = fabs(a);
      dl1[i] = m > 0 ? a * (1.0 + off / m) : (off > 0 ? off : 1.0);
    }
  }
}

// symmetric perturbation in [1, 1 + 2^-10) of the strength of edge {i, j}:
// with equal strengths (a uniform grid) the heaviest edge at a node is then
// also the heaviest at its partner, so picks are mutual (locally dominant
// edge matching) instead of chaining along the index order
__device__ __forceinline__ double edge_jitter(int64_t i, int64_t j) {
  const uint64_t a = (uint64_t)(i < j ? i : j), b = (uint64_t)(i < j ? j : i);
  uint64_t h = a * 0x9E3779B97F4A7C15ULL ^ (b + 0x632BE59BD9B4E019ULL + (a << 6) + (a >> 2));
  h ^= h >> 31; h *= 0xBF58476D1CE4E5B9ULL; h ^= h >> 27;
  return 1.0 + (double)(h >> 11) * (1.0 / 9007199254740992.0) * (1.0 / 1024.0);
}

// each unmatched node picks its strongest unmatched strong neighbour (the
// largest jittered score, the smallest index among equal ones: max and the
// comparisons are exact, so the lanes' partial picks combine to the pick of
// one pass over the row)
__global__ void k_pick(int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                       const double *__restrict__ v, const double *__restrict__ d, const int32_t *__restrict__ matched,
                       int32_t *__restrict__ pick) {
  const int lane = threadIdx.x & (kWave - 1);
  for (int64_t i = wave_first(); i < n; i += wave_count()) {  // i is the same on every lane
    int32_t best = -1;
    double bs = 0.;
    if (!matched[i]) {
      const int64_t b = rp[i], e = rp[i + 1];
      double rmax = 0.;
      for (int64_t p = b + lane; p < e; p += kWave) {
        const int32_t j = ci[p];
        if (j == (int32_t)i) continue;
        const double den = sqrt(fabs(d[i] * d[j]));
        rmax = fmax(rmax, den > 0 ? fabs(v[p]) / den : 0.);
      }
      const double thr = kTheta * wave_max(rmax);
      for (int64_t p = b + lane; p < e; p += kWave) {
        const int32_t j = ci[p];
        if (j == (int32_t)i || matched[j]) continue;
        const double den = sqrt(fabs(d[i] * d[j]));
        const double sc0 = den > 0 ? fabs(v[p]) / den : 0.;
        if (!(sc0 > 0.) || sc0 < thr) continue;
        const double sc = sc0 * edge_jitter(i, j);
        if (sc > bs || (sc == bs && j < best)) { bs = sc; best = j; }
      }
      for (int o = kWave / 2; o > 0; o >>= 1) {
        const double s2 = __shfl_xor(bs, o);
        const int32_t j2 = __shfl_xor(best, o);
        if (s2 > bs || (s2 == bs && j2 < best)) { bs = s2; best = j2; }
      }
    }
    if (lane == 0) pick[i] = best;
  }
}

// mutual picks become pairs
__global__ void k_pair(int64_t n, const int32_t *__restrict__ pick, int32_t *__restrict__ partner,
                       int32_t *__restrict__ matched_next) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t j = pick[i];
  if (j >= 0 && pick[j] == (int32_t)i) { partner[i] = j; matched_next[i] = 1; }
}

__global__ void k_copy_i32(int64_t n, const int32_t *__restrict__ a, int32_t *__restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_roots(int64_t n, const int32_t *__restrict__ partner, int64_t *__restrict__ root) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  root[i] = (partner[i] < 0 || (int32_t)i < partner[i]) ? 1 : 0;
}

// tentative prolongation: row i has one entry 1 in its aggregate's column
__global__ void k_ptent_pairs(int64_t n, const int32_t *__restrict__ partner, const int64_t *__restrict__ rootscan,
                              int64_t *__restrict__ rp, int32_t *__restrict__ ci, double *__restrict__ v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  rp[i] = i;
  if (i == n) return;
  const int32_t q = partner[i];
  const int64_t r = (q < 0 || (int32_t)i < q) ? i : q;
  ci[i] = (int32_t)rootscan[r];
  v[i] = 1.0;
}

__global__ void k_r_dinv_spmv(int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                              const double *__restrict__ v, const double *__restrict__ d, const double *__restrict__ x,
                              double *__restrict__ y) {
  const int lane = threadIdx.x & (kWave - 1);
  for (int64_t i = wave_first(); i < n; i += wave_count()) {  // i is the same on every lane
    double s = 0.;
    for (int64_t p = rp[i] + lane; p < rp[i + 1]; p += kWave) s += v[p] * x[ci[p]];
    s = wave_sum(s);
    if (lane == 0) y[i] = d[i] != 0. ? s / d[i] : 0.;
  }
}

// rows of A scaled by f / d_i (the smoothing matrix of the prolongation)
__global__ void k_r_scale_rows(int64_t n, const int64_t *__restrict__ rp, const double *__restrict__ v,
                               const double *__restrict__ d, double f, double *__restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double s = d[i] != 0. ? f / d[i] : 0.;
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) out[p] = s * v[p];
}

// Row products split two ways: rows of at most kRLong entries take 8 lanes
// each (k_r_spmv, k_r_jacobi), longer rows -- the conductor aggregate
// columns of V0d couple to every dielectric node around a conductor -- take a
// 256-thread block each (the *_long kernels), listed per matrix at setup.
constexpr int kRLong = 64;
constexpr int kRLanes = 8;

__device__ __forceinline__ double row_dot8(int64_t b, int64_t e, int lane, const int32_t *__restrict__ ci,
                                           const double *__restrict__ v, const double *__restrict__ x) {
  double s = 0.;
  for (int64_t p = b + lane; p < e; p += kRLanes) s += v[p] * x[ci[p]];
  for (int off = kRLanes / 2; off > 0; off >>= 1) s += __shfl_down(s, off, kRLanes);
  return s;
}

__device__ __forceinline__ double row_dot_block(int64_t b, int64_t e, const int32_t *__restrict__ ci,
                                                const double *__restrict__ v, const double *__restrict__ x) {
  __shared__ double sh[256];
  double s = 0.;
  for (int64_t p = b + threadIdx.x; p < e; p += blockDim.x) s += v[p] * x[ci[p]];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  return sh[0];
}

// y = A x (mode 0), y = r - A x (mode 1), y = y + A x (mode 2)
__global__ void k_r_spmv(int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                         const double *__restrict__ v, const double *__restrict__ x, const double *__restrict__ r,
                         double *__restrict__ y, int mode) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kRLanes;
  const int lane = threadIdx.x & (kRLanes - 1);
  if (i >= n) return;
  const int64_t b = rp[i], e = rp[i + 1];
  if (e - b > kRLong) return;
  const double s = row_dot8(b, e, lane, ci, v, x);
  if (lane == 0) y[i] = mode == 0 ? s : mode == 1 ? r[i] - s : y[i] + s;
}

__global__ void k_r_spmv_long(const int32_t *__restrict__ lrows, const int64_t *__restrict__ rp,
                              const int32_t *__restrict__ ci, const double *__restrict__ v,
                              const double *__restrict__ x, const double *__restrict__ r, double *__restrict__ y,
                              int mode) {
  const int64_t i = lrows[blockIdx.x];
  const double s = row_dot_block(rp[i], rp[i + 1], ci, v, x);
  if (threadIdx.x == 0) y[i] = mode == 0 ? s : mode == 1 ? r[i] - s : y[i] + s;
}

// l1-Jacobi, first sweep: xo = r / d
__global__ void k_r_jacobi0(int64_t n, const double *__restrict__ dl1, const double *__restrict__ r,
                            double *__restrict__ xo) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) xo[i] = 0. + r[i] / dl1[i];
}

// l1-Jacobi, later sweeps: xo = x + (r - A x) / d
__global__ void k_r_jacobi(int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                           const double *__restrict__ v, const double *__restrict__ dl1, const double *__restrict__ r,
                           const double *__restrict__ x, double *__restrict__ xo) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kRLanes;
  const int lane = threadIdx.x & (kRLanes - 1);
  if (i >= n) return;
  const int64_t b = rp[i], e = rp[i + 1];
  if (e - b > kRLong) return;
  const double s = row_dot8(b, e, lane, ci, v, x);
  if (lane == 0) xo[i] = x[i] + (r[i] - s) / dl1[i];
}

__global__ void k_r_jacobi_long(const int32_t *__restrict__ lrows, const int64_t *__restrict__ rp,
                                const int32_t *__restrict__ ci, const double *__restrict__ v,
                                const double *__restrict__ dl1, const double *__restrict__ r,
                                const double *__restrict__ x, double *__restrict__ xo) {
  const int64_t i = lrows[blockIdx.x];
  const double s = row_dot_block(rp[i], rp[i + 1], ci, v, x);
  if (threadIdx.x == 0) xo[i] = x[i] + (r[i] - s) / dl1[i];
}

// Chebyshev smoothing (the three-term recurrence written with a direction
// vector dd, Saad, Iterative Methods for Sparse Linear Systems, 2nd ed.,
// Alg. 12.1, preconditioned with D^-1), first step from x = 0:
// dd = c2 D^-1 r, xo = dd
__global__ void k_r_cheb0(int64_t n, const double *__restrict__ dg, const double *__restrict__ r, double c2,
                          double *__restrict__ dd, double *__restrict__ xo) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double q = dg[i] != 0. ? c2 * r[i] / dg[i] : 0.;
  dd[i] = q;
  xo[i] = q;
}

// later steps: dd = c1 dd + c2 D^-1 (r - A x), xo = x + dd (c1 = 0 starts the
// recurrence from a nonzero x; dd is then not read)
__device__ __forceinline__ void cheb_row(int64_t i, double s, const double *__restrict__ dg,
                                         const double *__restrict__ r, const double *__restrict__ x, double c1,
                                         double c2, double *__restrict__ dd, double *__restrict__ xo) {
  double q = dg[i] != 0. ? c2 * (r[i] - s) / dg[i] : 0.;
  if (c1 != 0.) q += c1 * dd[i];
  dd[i] = q;
  xo[i] = x[i] + q;
}

__global__ void k_r_cheb(int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                         const double *__restrict__ v, const double *__restrict__ dg, const double *__restrict__ r,
                         const double *__restrict__ x, double c1, double c2, double *__restrict__ dd,
                         double *__restrict__ xo) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kRLanes;
  const int lane = threadIdx.x & (kRLanes - 1);
  if (i >= n) return;
  const int64_t b = rp[i], e = rp[i + 1];
  if (e - b > kRLong) return;
  const double s = row_dot8(b, e, lane, ci, v, x);
  if (lane == 0) cheb_row(i, s, dg, r, x, c1, c2, dd, xo);
}

__global__ void k_r_cheb_long(const int32_t *__restrict__ lrows, const int64_t *__restrict__ rp,
                              const int32_t *__restrict__ ci, const double *__restrict__ v,
                              const double *__restrict__ dg, const double *__restrict__ r,
                              const double *__restrict__ x, double c1, double c2, double *__restrict__ dd,
                              double *__restrict__ xo) {
  const int64_t i = lrows[blockIdx.x];
  const double s = row_dot_block(rp[i], rp[i + 1], ci, v, x);
  if (threadIdx.x == 0) cheb_row(i, s, dg, r, x, c1, c2, dd, xo);
}

// x = inv r on the coarsest level, one wavefront per row (coalesced rows)
__global__ void k_r_dense(int64_t n, const double *__restrict__ inv, const double *__restrict__ r, double *__restrict__ x) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  double s = 0.;
  for (int64_t j = lane; j < n; j += 64) s += inv[i * n + j] * r[j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) x[i] = s;
}

// partial[t * nb + blk] = sum over the block's rows of V_t . w
__global__ void k_r_dots(int64_t n, const double *__restrict__ V, int64_t vstride, int nvec, const double *__restrict__ w,
                         double *__restrict__ partial) {
  __shared__ double sh[256];
  const int t = blockIdx.y;
  double acc = 0.;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    acc += V[(int64_t)t * vstride + i] * w[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(int64_t)t * gridDim.x + blockIdx.x] = sh[0];
}

// w = (w - sum_t V_t h_t) * sc
__global__ void k_r_gs(int64_t n, const double *__restrict__ V, int64_t vstride, int nvec, const double *__restrict__ h,
                       double sc, double *__restrict__ w) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a = w[i];
  for (int t = 0; t < nvec; ++t) a -= V[(int64_t)t * vstride + i] * h[t];
  w[i] = a * sc;
}

// x += sum_t Z_t y_t
__global__ void k_r_xup(int64_t n, const double *__restrict__ Z, int64_t zstride, int nvec, const double *__restrict__ y,
                        double *__restrict__ x) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a = x[i];
  for (int t = 0; t < nvec; ++t) a += Z[(int64_t)t * zstride + i] * y[t];
  x[i] = a;
}

__global__ void k_r_scal(int64_t n, double f, double *__restrict__ x) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= f;
}

__global__ void k_r_to_c(int64_t n, const double *__restrict__ a, C2<double> *__restrict__ c) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = {a[i], 0.};
}

__global__ void k_c_to_r(int64_t n, const C2<double> *__restrict__ c, double *__restrict__ a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = c[i].x;
}

__global__ void k_csr_dense(int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                            const double *__restrict__ v, double *__restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) out[i * n + ci[p]] += v[p];
}

}  // namespace

struct RLevel {
  Csr A, P, R;
  double *d = nullptr, *dl1 = nullptr;
  double *r = nullptr, *x = nullptr, *t = nullptr;
  // Chebyshev smoothing: direction vector and the coefficients of each step
  double *dd = nullptr;
  double c1[kChebSteps] = {}, c2[kChebSteps] = {};
  // rows longer than kRLong of A, P and R (device lists) and their counts
  int32_t *la = nullptr, *lp = nullptr, *lr = nullptr;
  int64_t nla = 0, nlp = 0, nlr = 0;
};

struct RealAMG {
  int smoother = GP_SMOOTH_L1_JACOBI;
  std::vector<RLevel> L;
  double *inv = nullptr;  // coarsest, dense n x n
  // FGMRES workspace: basis V (m+1 vectors), preconditioned basis Z (m),
  // block partial sums of the dot products, Hessenberg coefficients
  double *V = nullptr, *Zb = nullptr, *partial = nullptr, *hd = nullptr;
  Arena A;
};

namespace {

// rows of M longer than kRLong, as a device list
gp_status long_rows(Arena &A, const Csr &M, int32_t **out, int64_t *nl, hipStream_t s) {
  *out = nullptr;
  *nl = 0;
  if (M.nr == 0) return GP_OK;
  std::vector<int64_t> rp((size_t)M.nr + 1);
  GP_HIP(hipMemcpyAsync(rp.data(), M.rp, sizeof(int64_t) * rp.size(), hipMemcpyDeviceToHost, s));
  GP_HIP(hipStreamSynchronize(s));
  std::vector<int32_t> l;
  for (int64_t i = 0; i < M.nr; ++i)
    if (rp[i + 1] - rp[i] > kRLong) l.push_back((int32_t)i);
  if (l.empty()) return GP_OK;
  GP_TRY(A.alloc(out, (int64_t)l.size()));
  GP_HIP(hipMemcpyAsync(*out, l.data(), sizeof(int32_t) * l.size(), hipMemcpyHostToDevice, s));
  GP_HIP(hipStreamSynchronize(s));
  *nl = (int64_t)l.size();
  return GP_OK;
}

void spmv_r(const Csr &M, const int32_t *lr, int64_t nl, const double *x, const double *r, double *y, int mode,
            hipStream_t s) {
  if (M.nr == 0) return;
  k_r_spmv<<<blocks_for(M.nr * kRLanes), 256, 0, s>>>(M.nr, M.rp, M.ci, M.v[0], x, r, y, mode);
  if (nl) k_r_spmv_long<<<(unsigned)nl, 256, 0, s>>>(lr, M.rp, M.ci, M.v[0], x, r, y, mode);
}

void jacobi_r(const RLevel &L, const double *x, double *xo, hipStream_t s) {
  const int64_t n = L.A.nr;
  if (!x) { k_r_jacobi0<<<blocks_for(n), 256, 0, s>>>(n, L.dl1, L.r, xo); return; }
  k_r_jacobi<<<blocks_for(n * kRLanes), 256, 0, s>>>(n, L.A.rp, L.A.ci, L.A.v[0], L.dl1, L.r, x, xo);
  if (L.nla) k_r_jacobi_long<<<(unsigned)L.nla, 256, 0, s>>>(L.la, L.A.rp, L.A.ci, L.A.v[0], L.dl1, L.r, x, xo);
}

// the Chebyshev coefficients for lambda, the estimate of D^-1 A's largest
// eigenvalue: interval [alpha, beta], theta its centre, delta its half width,
// sigma = theta / delta; step 0: c1 = 0, c2 = 1 / theta; step k: rho_k =
// 1 / (2 sigma - rho_(k-1)) from rho_0 = 1 / sigma, c1 = rho_k rho_(k-1),
// c2 = 2 rho_k / delta
void cheb_coeffs(double lam, double *c1, double *c2) {
  const double beta = kChebUpper * lam, alpha = kChebLower * beta;
  const double theta = 0.5 * (beta + alpha), delta = 0.5 * (beta - alpha), sigma = theta / delta;
  double rho = 1.0 / sigma;
  c1[0] = 0.;
  c2[0] = 1.0 / theta;
  for (int k = 1; k < kChebSteps; ++k) {
    const double rn = 1.0 / (2.0 * sigma - rho);
    c1[k] = rn * rho;
    c2[k] = 2.0 * rn / delta;
    rho = rn;
  }
}

// step k of the Chebyshev smoother from x (null: x = 0, step 0 only) into xo
void cheb_r(const RLevel &L, int k, const double *x, double *xo, hipStream_t s) {
  const int64_t n = L.A.nr;
  if (!x) { k_r_cheb0<<<blocks_for(n), 256, 0, s>>>(n, L.d, L.r, L.c2[0], L.dd, xo); return; }
  k_r_cheb<<<blocks_for(n * kRLanes), 256, 0, s>>>(n, L.A.rp, L.A.ci, L.A.v[0], L.d, L.r, x, L.c1[k], L.c2[k], L.dd, xo);
  if (L.nla)
    k_r_cheb_long<<<(unsigned)L.nla, 256, 0, s>>>(L.la, L.A.rp, L.A.ci, L.A.v[0], L.d, L.r, x, L.c1[k], L.c2[k], L.dd, xo);
}

// one pairwise matching pass on M: tentative P (n x ncoarse)
gp_status match_pass(Arena &A, Arena &tmp, const Csr &M, const double *d, Csr *P, hipStream_t s) {
  const int64_t n = M.nr;
  int32_t *matched, *mnext, *pick, *partner;
  int64_t *root, *rscan;
  GP_TRY(tmp.alloc(&matched, n)); GP_TRY(tmp.alloc(&mnext, n)); GP_TRY(tmp.alloc(&pick, n));
  GP_TRY(tmp.alloc(&partner, n)); GP_TRY(tmp.alloc(&root, n)); GP_TRY(tmp.alloc(&rscan, n + 1));
  GP_HIP(hipMemsetAsync(matched, 0, sizeof(int32_t) * n, s));
  GP_HIP(hipMemsetAsync(mnext, 0, sizeof(int32_t) * n, s));
  GP_HIP(hipMemsetAsync(partner, 0xff, sizeof(int32_t) * n, s));  // -1
  for (int round = 0; round < 6; ++round) {
    k_pick<<<wave_blocks(n), 256, 0, s>>>(n, M.rp, M.ci, M.v[0], d, matched, pick);
    k_pair<<<blocks_for(n), 256, 0, s>>>(n, pick, partner, mnext);
    k_copy_i32<<<blocks_for(n), 256, 0, s>>>(n, mnext, matched);
  }
  k_roots<<<blocks_for(n), 256, 0, s>>>(n, partner, root);
  GP_TRY(exclusive_scan_i64(root, rscan, n, s, tmp));
  int64_t nc = 0;
  GP_HIP(hipMemcpyAsync(&nc, rscan + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  GP_HIP(hipStreamSynchronize(s));
  GP_TRY(csr_alloc(A, n, nc, n, 1, P));
  k_ptent_pairs<<<blocks_for(n + 1), 256, 0, s>>>(n, partner, rscan, P->rp, P->ci, P->v[0]);
  GP_HIP(hipStreamSynchronize(s));
  return GP_OK;
}

double power_lambda(Arena &tmp, const Csr &M, const double *d, hipStream_t s) {
  const int64_t n = M.nr;
  double *x = nullptr, *y = nullptr;
  if (tmp.alloc(&x, n) != GP_OK || tmp.alloc(&y, n) != GP_OK) return 2.0;
  std::vector<double> h(n);
  for (int64_t i = 0; i < n; ++i) h[i] = 0.5 + 0.5 * (double)((i * 2654435761ULL) % 1000) / 1000.0;
  (void)hipMemcpyAsync(x, h.data(), sizeof(double) * n, hipMemcpyHostToDevice, s);
  double lam = 1.0;
  for (int it = 0; it < 12; ++it) {
    k_r_dinv_spmv<<<wave_blocks(n), 256, 0, s>>>(n, M.rp, M.ci, M.v[0], d, x, y);
    (void)hipMemcpyAsync(h.data(), y, sizeof(double) * n, hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    double nr = 0;
    for (double v : h) nr += v * v;
    nr = std::sqrt(nr);
    lam = nr;  // |D^-1 A x| with |x| = 1 after the first step
    if (!(nr > 0)) break;
    for (double &v : h) v /= nr;
    (void)hipMemcpyAsync(x, h.data(), sizeof(double) * n, hipMemcpyHostToDevice, s);
  }
  return std::max(lam, 1e-3);
}

gp_status build_real_amg(RealAMG &G, Csr A0, hipStream_t s) {
  Arena &A = G.A;
  G.L.clear();
  RLevel l0;
  l0.A = A0;
  G.L.push_back(l0);
  for (;;) {
    RLevel &F = G.L.back();
    const int64_t n = F.A.nr;
    GP_TRY(A.alloc(&F.d, n)); GP_TRY(A.alloc(&F.dl1, n));
    k_r_diag<<<wave_blocks(n), 256, 0, s>>>(n, F.A.rp, F.A.ci, F.A.v[0], F.d, F.dl1);
    GP_TRY(A.alloc(&F.r, n)); GP_TRY(A.alloc(&F.x, n)); GP_TRY(A.alloc(&F.t, n));
    GP_TRY(long_rows(A, F.A, &F.la, &F.nla, s));
    if (n <= kCoarsest || G.L.size() >= 24) break;
    Arena tmp;
    gp_status st = GP_OK;
    Csr P1, M2, AP1, P2, Pt, DA, DAP, Ps, P, AP, Ac, R;
    do {
      if ((st = match_pass(tmp, tmp, F.A, F.d, &P1, s)) != GP_OK) break;
      Csr R1;
      if ((st = transpose(tmp, P1, &R1, s)) != GP_OK) break;
      if ((st = spgemm(tmp, F.A, P1, &AP1, s)) != GP_OK) break;
      if ((st = spgemm(tmp, R1, AP1, &M2, s)) != GP_OK) break;
      double *d2 = nullptr, *l2 = nullptr;
      if ((st = tmp.alloc(&d2, M2.nr)) != GP_OK || (st = tmp.alloc(&l2, M2.nr)) != GP_OK) break;
      k_r_diag<<<wave_blocks(M2.nr), 256, 0, s>>>(M2.nr, M2.rp, M2.ci, M2.v[0], d2, l2);
      if ((st = match_pass(tmp, tmp, M2, d2, &P2, s)) != GP_OK) break;
      if ((st = spgemm(tmp, P1, P2, &Pt, s)) != GP_OK) break;
      if (Pt.nc > (int64_t)(0.8 * (double)n)) { st = GP_ERR_NOCONV; break; }  // stalled: stop coarsening
      // P = Pt - 4/(3 lambda) D^-1 A Pt
      const double lam = power_lambda(tmp, F.A, F.d, s);
      if (G.smoother == GP_SMOOTH_CHEBYSHEV) {
        cheb_coeffs(lam, F.c1, F.c2);
        if ((st = A.alloc(&F.dd, n)) != GP_OK) break;
      }
      if ((st = csr_alloc(tmp, n, n, F.A.nnz, 1, &DA)) != GP_OK) break;
      if (hipMemcpyAsync(DA.rp, F.A.rp, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToDevice, s) != hipSuccess ||
          hipMemcpyAsync(DA.ci, F.A.ci, sizeof(int32_t) * F.A.nnz, hipMemcpyDeviceToDevice, s) != hipSuccess) { st = fail(GP_ERR_HIP, "copy"); break; }
      k_r_scale_rows<<<blocks_for(n), 256, 0, s>>>(n, F.A.rp, F.A.v[0], F.d, 4.0 / (3.0 * lam), DA.v[0]);
      if ((st = spgemm(tmp, DA, Pt, &DAP, s)) != GP_OK) break;
      if ((st = csr_add(tmp, Pt, -1.0, DAP, &Ps, s)) != GP_OK) break;
      if ((st = truncate_prolongation(A, tmp, Ps, Pt, kTrunc, &P, s)) != GP_OK) break;
      if ((st = transpose(A, P, &R, s)) != GP_OK) break;
      if ((st = spgemm(tmp, F.A, P, &AP, s)) != GP_OK) break;
      if ((st = spgemm(A, R, AP, &Ac, s)) != GP_OK) break;
    } while (0);
    tmp.free_all();
    if (st == GP_ERR_NOCONV) { (void)hipGetLastError(); break; }
    if (st != GP_OK) return st;
    F.P = P;
    F.R = R;
    GP_TRY(long_rows(A, F.P, &F.lp, &F.nlp, s));
    GP_TRY(long_rows(A, F.R, &F.lr, &F.nlr, s));
    RLevel C;
    C.A = Ac;
    G.L.push_back(C);
  }
  // coarsest: dense inverse on the device
  RLevel &C = G.L.back();
  const int64_t n = C.A.nr;
  if (n > 4096) return fail(GP_ERR_NOCONV, "real AMG coarsening stalled above 4096 unknowns");
  double *dn = nullptr;
  C2<double> *cm = nullptr;
  Arena tmp;
  gp_status st = GP_OK;
  do {
    if ((st = tmp.alloc(&dn, std::max<int64_t>(n * n, 1))) != GP_OK) break;
    if ((st = tmp.alloc(&cm, std::max<int64_t>(n * n, 1))) != GP_OK) break;
    if (hipMemsetAsync(dn, 0, sizeof(double) * n * n, s) != hipSuccess) { st = fail(GP_ERR_HIP, "memset"); break; }
    k_csr_dense<<<blocks_for(n), 256, 0, s>>>(n, C.A.rp, C.A.ci, C.A.v[0], dn);
    k_r_to_c<<<blocks_for(n * n), 256, 0, s>>>(n * n, dn, cm);
    if ((st = dense_inverse_batched(cm, (int)n, 1, s)) != GP_OK) break;
    if ((st = A.alloc(&G.inv, std::max<int64_t>(n * n, 1))) != GP_OK) break;
    k_c_to_r<<<blocks_for(n * n), 256, 0, s>>>(n * n, cm, G.inv);
    if (hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "coarse inverse"); break; }
  } while (0);
  tmp.free_all();
  return st;
}

void vcycle_real(const RealAMG &G, int l, hipStream_t s) {
  const RLevel &L = G.L[l];
  const int64_t n = L.A.nr;
  if (l + 1 == (int)G.L.size()) {
    k_r_dense<<<blocks_for(n * 64), 256, 0, s>>>(n, G.inv, L.r, L.x);
    return;
  }
  static_assert(kChebSteps == 2, "the smoothing steps alternate between x and t twice");
  const bool cheb = G.smoother == GP_SMOOTH_CHEBYSHEV;
  if (cheb) { cheb_r(L, 0, nullptr, L.x, s); cheb_r(L, 1, L.x, L.t, s); }
  else { jacobi_r(L, nullptr, L.x, s); jacobi_r(L, L.x, L.t, s); }
  spmv_r(L.A, L.la, L.nla, L.t, L.r, L.x, 1, s);      // x <- r - A t
  const RLevel &C = G.L[l + 1];
  spmv_r(L.R, L.lr, L.nlr, L.x, nullptr, C.r, 0, s);
  vcycle_real(G, l + 1, s);
  spmv_r(L.P, L.lp, L.nlp, C.x, nullptr, L.t, 2, s);  // t += P xc
  if (cheb) { cheb_r(L, 0, L.t, L.x, s); cheb_r(L, 1, L.x, L.t, s); }
  else { jacobi_r(L, L.t, L.x, s); jacobi_r(L, L.x, L.t, s); }
  (void)hipMemcpyAsync(L.x, L.t, sizeof(double) * n, hipMemcpyDeviceToDevice, s);
}

gp_status rdots(hipStream_t s, int64_t n, const double *V, int64_t vstride, int nvec, const double *w, double *partial,
                std::vector<double> &h) {
  const int nb = 256;
  k_r_dots<<<dim3(nb, nvec), 256, 0, s>>>(n, V, vstride, nvec, w, partial);
  std::vector<double> p((size_t)nb * nvec);
  GP_HIP(hipMemcpyAsync(p.data(), partial, sizeof(double) * p.size(), hipMemcpyDeviceToHost, s));
  GP_HIP(hipStreamSynchronize(s));
  h.assign(nvec, 0.);
  for (int t = 0; t < nvec; ++t)
    for (int b = 0; b < nb; ++b) h[t] += p[(size_t)t * nb + b];
  return GP_OK;
}

constexpr int kRestart = 30;

}  // namespace

gp_status real_amg_create(int64_t n, int64_t nnz, const int64_t *rows, const int64_t *cols, const double *vals,
                          int smoother, RealAMG **out, hipStream_t s) {
  *out = nullptr;
  if (n < 1 || nnz < 1) return fail(GP_ERR_ARG, "empty real matrix");
  if (n > 2147483647LL) return fail(GP_ERR_ARG, "more than 2^31 unknowns");
  if (smoother != GP_SMOOTH_L1_JACOBI && smoother != GP_SMOOTH_CHEBYSHEV) return fail(GP_ERR_ARG, "unknown real smoother");
  RealAMG *G = new RealAMG();
  G->smoother = smoother;
  Csr A0;
  const double *vv[1] = {vals};
  gp_status st = csr_from_coo(G->A, n, n, nnz, rows, cols, vv, 1, &A0, s);
  if (st == GP_OK) st = build_real_amg(*G, A0, s);
  if (st == GP_OK) st = G->A.alloc(&G->V, (int64_t)(kRestart + 1) * n);
  if (st == GP_OK) st = G->A.alloc(&G->Zb, (int64_t)kRestart * n);
  if (st == GP_OK) st = G->A.alloc(&G->partial, (int64_t)256 * (kRestart + 1));
  if (st == GP_OK) st = G->A.alloc(&G->hd, kRestart + 1);
  if (st != GP_OK) { real_amg_destroy(G); return st; }
  *out = G;
  return GP_OK;
}

int64_t real_amg_rows(const RealAMG *G) { return G->L[0].A.nr; }
int32_t real_amg_levels(const RealAMG *G) { return (int32_t)G->L.size(); }
int64_t real_amg_bytes(const RealAMG *G) { return G->A.bytes; }

void real_amg_destroy(RealAMG *G) {
  if (!G) return;
  G->A.free_all();
  delete G;
}

// flexible GMRES(m), FP64, classical Gram-Schmidt with one
// reorthogonalization pass, Givens rotations on the host
gp_status real_amg_solve(RealAMG *G, const double *bd, double *xd, double tol, int maxit, int32_t *iters_out,
                         double *relres_out, hipStream_t s) {
  const int m = kRestart;
  const RLevel &L0 = G->L[0];
  const int64_t n = L0.A.nr;
  double *V = G->V, *Zb = G->Zb, *partial = G->partial, *hd = G->hd;
  GP_HIP(hipMemsetAsync(xd, 0, sizeof(double) * n, s));
  const unsigned gb = blocks_for(n);
  std::vector<double> h, h2;
  GP_TRY(rdots(s, n, bd, 0, 1, bd, partial, h));
  const double bn = std::sqrt(h[0]);
  int iters = 0;
  double rel = 0.;
  if (bn > 0.) {
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
    for (;;) {
      spmv_r(L0.A, L0.la, L0.nla, xd, bd, V, 1, s);
      GP_TRY(rdots(s, n, V, 0, 1, V, partial, h));
      const double beta = std::sqrt(h[0]);
      rel = beta / bn;
      if (rel <= tol || iters >= maxit) break;
      k_r_scal<<<gb, 256, 0, s>>>(n, 1.0 / beta, V);
      std::fill(g.begin(), g.end(), 0.);
      g[0] = beta;
      int used = 0;
      for (int i = 0; i < m && iters < maxit; ++i) {
        double *Vi = V + (int64_t)i * n, *Vn = V + (int64_t)(i + 1) * n, *Zi = Zb + (int64_t)i * n;
        GP_HIP(hipMemcpyAsync(L0.r, Vi, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        vcycle_real(*G, 0, s);
        GP_HIP(hipMemcpyAsync(Zi, L0.x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        spmv_r(L0.A, L0.la, L0.nla, Zi, nullptr, Vn, 0, s);
        for (int pass = 0; pass < 2; ++pass) {
          GP_TRY(rdots(s, n, V, n, i + 1, Vn, partial, h2));
          GP_HIP(hipMemcpyAsync(hd, h2.data(), sizeof(double) * (i + 1), hipMemcpyHostToDevice, s));
          k_r_gs<<<gb, 256, 0, s>>>(n, V, n, i + 1, hd, 1.0, Vn);
          for (int t = 0; t <= i; ++t) H[t * m + i] = (pass ? H[t * m + i] : 0.) + h2[t];
        }
        GP_TRY(rdots(s, n, Vn, 0, 1, Vn, partial, h));
        const double hn = std::sqrt(std::max(h[0], 0.));
        H[(i + 1) * m + i] = hn;
        if (hn > 0) k_r_scal<<<gb, 256, 0, s>>>(n, 1.0 / hn, Vn);
        for (int t = 0; t < i; ++t) {
          const double a = H[t * m + i], bq = H[(t + 1) * m + i];
          H[t * m + i] = cs[t] * a + sn[t] * bq;
          H[(t + 1) * m + i] = -sn[t] * a + cs[t] * bq;
        }
        const double a = H[i * m + i], bq = H[(i + 1) * m + i];
        const double den = std::sqrt(a * a + bq * bq);
        cs[i] = den > 0 ? a / den : 1.;
        sn[i] = den > 0 ? bq / den : 0.;
        H[i * m + i] = cs[i] * a + sn[i] * bq;
        H[(i + 1) * m + i] = 0.;
        g[i + 1] = -sn[i] * g[i];
        g[i] = cs[i] * g[i];
        ++iters;
        used = i + 1;
        if (std::fabs(g[i + 1]) <= tol * bn) break;
      }
      for (int t = used - 1; t >= 0; --t) {
        double acc = g[t];
        for (int q = t + 1; q < used; ++q) acc -= H[t * m + q] * y[q];
        y[t] = H[t * m + t] != 0. ? acc / H[t * m + t] : 0.;
      }
      GP_HIP(hipMemcpyAsync(hd, y.data(), sizeof(double) * used, hipMemcpyHostToDevice, s));
      k_r_xup<<<gb, 256, 0, s>>>(n, Zb, n, used, hd, xd);
    }
  }
  GP_HIP(hipStreamSynchronize(s));
  if (iters_out) *iters_out = iters;
  if (relres_out) *relres_out = rel;
  return rel <= tol ? GP_OK : fail(GP_ERR_NOCONV, "real solve stopped above the tolerance");
}

}  // namespace gp

using namespace gp;

extern "C" gp_status gp_solve_real(const int64_t *row, const int64_t *col, const double *val, int64_t nnz, int64_t n,
                                   const double *b, double *x, double rtol, int32_t *iters_out, double *relres_out) {
  return gp_solve_real_ex(row, col, val, nnz, n, b, x, rtol, GP_SMOOTH_L1_JACOBI, iters_out, relres_out);
}

extern "C" gp_status gp_solve_real_ex(const int64_t *row, const int64_t *col, const double *val, int64_t nnz,
                                      int64_t n, const double *b, double *x, double rtol, int32_t smoother,
                                      int32_t *iters_out, double *relres_out) {
  if (!row || !col || !val || !b || !x || n < 1 || nnz < 1) return fail(GP_ERR_ARG, "bad real solve arguments");
  if (smoother != GP_SMOOTH_L1_JACOBI && smoother != GP_SMOOTH_CHEBYSHEV) return fail(GP_ERR_ARG, "unknown real smoother");
  if (n > 2147483647LL) return fail(GP_ERR_ARG, "more than 2^31 unknowns");
  for (int64_t q = 0; q < nnz; ++q)
    if (row[q] < 0 || row[q] >= n || col[q] < 0 || col[q] >= n) return fail(GP_ERR_ARG, "COO index out of range");
  GP_TRY(ensure_device());
  const double tol = rtol > 0 ? rtol : 1e-4;  // the reference's HYPRE tolerance (src/hypreSolver.h:8)
  hipStream_t s;
  GP_HIP(hipStreamCreate(&s));
  struct StreamGuard { hipStream_t s; ~StreamGuard() { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); } } sg{s};
  Arena tmp;
  struct TGuard { Arena *a; hipStream_t s; ~TGuard() { (void)hipStreamSynchronize(s); a->free_all(); } } tg{&tmp, s};
  int64_t *dr, *dc;
  double *dv, *bd, *xd;
  GP_TRY(tmp.alloc(&dr, nnz)); GP_TRY(tmp.alloc(&dc, nnz)); GP_TRY(tmp.alloc(&dv, nnz));
  GP_TRY(tmp.alloc(&bd, n)); GP_TRY(tmp.alloc(&xd, n));
  GP_HIP(hipMemcpyAsync(dr, row, sizeof(int64_t) * nnz, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(dc, col, sizeof(int64_t) * nnz, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(dv, val, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(bd, b, sizeof(double) * n, hipMemcpyHostToDevice, s));
  RealAMG *G = nullptr;
  GP_TRY(real_amg_create(n, nnz, dr, dc, dv, smoother, &G, s));
  struct GGuard { RealAMG *g; ~GGuard() { real_amg_destroy(g); } } gg{G};
  gp_status st = real_amg_solve(G, bd, xd, tol, 1000, iters_out, relres_out, s);
  if (st != GP_OK && st != GP_ERR_NOCONV) return st;
  GP_HIP(hipMemcpyAsync(x, xd, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  GP_HIP(hipStreamSynchronize(s));
  return st;
}

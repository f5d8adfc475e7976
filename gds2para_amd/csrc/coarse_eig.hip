// coarse_eig.hip -- the auxiliary coarsest level, diagonalised once per layout.
//
// The auxiliary operator N(w) = Gx^T K(w) Gx = -w^2 N_eps + i w N_sig with
// N_eps = Gx^T D_eps Gx symmetric positive definite and N_sig = Gx^T D_sig Gx
// symmetric positive semi-definite (S Gx = 0: the columns of Gx are
// gradients).  Its coarsest Galerkin level keeps that form, so the symmetric
// definite pencil N_sig v = lambda N_eps v, V^T N_eps V = I, solved once,
// gives every frequency's inverse
//     N(w)^-1 = V diag(1 / (-w^2 + i w lambda)) V^T
// with no per-frequency factorization.  The coarse solve x = N(w)^-1 r of k
// columns (each at its own w) is two real FP64 products, V^T R and V Y,
// with the diagonal applied between them; both run on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64).  The pencil is solved by rocSOLVER's dsygvd at
// setup; where it fails (N_eps not numerically definite) the level keeps the
// per-frequency Gauss-Jordan inverse.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "mg.hpp"

namespace gp {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

__global__ void k_transpose_sq(int64_t n, const double *__restrict__ a, double *__restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * n) return;
  const int64_t i = t / n, j = t - i * n;
  b[j * n + i] = a[t];
}

// Real column q of an n x k complex block (interleaved slots): slot q / 2,
// real part for even q, imaginary for odd.
template <class T>
__device__ __forceinline__ double real_col(const C2<T> *__restrict__ r, int64_t i, int k, int q) {
  const C2<T> v = r[i * k + (q >> 1)];
  return (q & 1) ? (double)v.y : (double)v.x;
}

// out[16 rows m0.., 16 real columns c0..] = sum over i of M[m][i] B[i][c]:
// M row-major n x n (real), B the n x k complex block seen as n x 2k real.
// One workgroup of four wavefronts per 16 x 16 output tile; the wavefronts
// split the sum over i four ways, each running v_mfma_f64_16x16x4f64 over
// its quarter (A lane l: M[m0 + l % 16][i0 + l / 16]; B lane l:
// B[i0 + l / 16][c0 + l % 16]; D lane l, element q: row l / 16 + 4 q,
// column l % 16, as measured on gfx950), and the four partial tiles are
// summed in LDS in a fixed
// order.  Returns the summed tile in red[0] (row-major 16 x 16).
template <class T>
__device__ __forceinline__ void mfma_tile(int64_t n, int k, const double *__restrict__ M, const C2<T> *__restrict__ B,
                                          int64_t m0, int c0, double (*red)[16][16]) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t quarter = ((n + 15) / 16) * 4;  // a multiple of 4 covering n / 4
  const int64_t kb = wave * quarter, ke = kb + quarter < n ? kb + quarter : n;
  const int64_t am = m0 + (lane & 15);
  const int bc = c0 + (lane & 15);
  dbl4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i0 = kb; i0 < ke; i0 += 4) {
    const int64_t ii = i0 + (lane >> 4);
    const double a = (am < n && ii < ke) ? M[am * n + ii] : 0.0;
    const double b = (ii < ke && bc < 2 * k) ? real_col<T>(B, ii, k, bc) : 0.0;
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) red[wave][(lane >> 4) + 4 * q][lane & 15] = acc[q];
  __syncthreads();
  const int row = threadIdx.x >> 4, col = threadIdx.x & 15;
  const double s = ((red[0][row][col] + red[1][row][col]) + red[2][row][col]) + red[3][row][col];
  __syncthreads();
  red[0][row][col] = s;
  __syncthreads();
}

// Yp[m][j] = d_m(w_j) * (V^T R)[m][j], d_m(w) = 1 / (-w^2 + i w lambda_m).
template <class T>
__global__ void __launch_bounds__(256) k_eig_project(int64_t n, int k, const double *__restrict__ VT,
                                                     const double *__restrict__ lam, const double *__restrict__ om,
                                                     const C2<T> *__restrict__ r, C2<double> *__restrict__ yp) {
  __shared__ double red[4][16][16];
  const int64_t m0 = (int64_t)blockIdx.x * 16;
  const int c0 = blockIdx.y * 16;
  mfma_tile<T>(n, k, VT, r, m0, c0, red);
  const int row = threadIdx.x >> 4, col = threadIdx.x & 15;
  const int64_t m = m0 + row;
  const int c = c0 + col;
  if (m >= n || c >= 2 * k || (col & 1)) return;
  const int j = c >> 1;
  const double w = om[j], l = lam[m];
  const double dr = -w * w, di = w * l;          // -w^2 + i w lambda
  const double den = dr * dr + di * di;
  const double ir = dr / den, ii = -di / den;    // its reciprocal
  const double yr = red[0][row][col], yi = red[0][row][col + 1];
  yp[m * k + j] = {ir * yr - ii * yi, ir * yi + ii * yr};
}

// x = V Yp (the coarse solution, written in the V-cycle's precision).
template <class T>
__global__ void __launch_bounds__(256) k_eig_expand(int64_t n, int k, const double *__restrict__ V,
                                                    const C2<double> *__restrict__ yp, C2<T> *__restrict__ x) {
  __shared__ double red[4][16][16];
  const int64_t i0 = (int64_t)blockIdx.x * 16;
  const int c0 = blockIdx.y * 16;
  mfma_tile<double>(n, k, V, yp, i0, c0, red);
  const int row = threadIdx.x >> 4, col = threadIdx.x & 15;
  const int64_t i = i0 + row;
  const int c = c0 + col;
  if (i >= n || c >= 2 * k) return;
  T *xs = (T *)(x + i * k + (c >> 1));
  xs[c & 1] = (T)red[0][row][col];
}

}  // namespace

gp_status aux_coarse_eigen(Arena &A, AuxAMG &X, hipStream_t s) {
  X.eig = false;
  const int64_t n = X.nL;
  if (n <= 0) return GP_OK;
  if (n > 46340) return GP_OK;  // n * n must fit rocSOLVER's 32-bit sizes
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) return fail(GP_ERR_HIP, "rocblas_create_handle failed");
  struct HandleGuard { rocblas_handle h; ~HandleGuard() { rocblas_destroy_handle(h); } } hg{h};
  if (rocblas_set_stream(h, s) != rocblas_status_success) return fail(GP_ERR_HIP, "rocblas_set_stream failed");
  Arena T;
  struct Free { Arena &a; ~Free() { a.free_all(); } } fr{T};
  double *a = nullptr, *b = nullptr, *e = nullptr, *lam = nullptr;
  rocblas_int *info = nullptr;
  GP_TRY(T.alloc(&a, n * n)); GP_TRY(T.alloc(&b, n * n)); GP_TRY(T.alloc(&e, n)); GP_TRY(T.alloc(&info, 1));
  GP_TRY(T.alloc(&lam, n));
  // symmetric inputs: row- and column-major agree
  GP_HIP(hipMemcpyAsync(a, X.NLd[1], sizeof(double) * n * n, hipMemcpyDeviceToDevice, s));
  GP_HIP(hipMemcpyAsync(b, X.NLd[0], sizeof(double) * n * n, hipMemcpyDeviceToDevice, s));
  const rocblas_status rs = rocsolver_dsygvd(h, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper,
                                             (rocblas_int)n, a, (rocblas_int)n, b, (rocblas_int)n, lam, e, info);
  rocblas_int hinfo = -1;
  GP_HIP(hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost, s));
  GP_HIP(hipStreamSynchronize(s));
  if (rs != rocblas_status_success || hinfo != 0) return GP_OK;  // keep the per-frequency inverse
  // a holds V column-major, i.e. V^T row-major
  GP_TRY(A.alloc(&X.VT, n * n));
  GP_TRY(A.alloc(&X.V, n * n));
  GP_TRY(A.alloc(&X.lam, n));
  GP_HIP(hipMemcpyAsync(X.VT, a, sizeof(double) * n * n, hipMemcpyDeviceToDevice, s));
  GP_HIP(hipMemcpyAsync(X.lam, lam, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  k_transpose_sq<<<blocks_for(n * n), 256, 0, s>>>(n, a, X.V);
  GP_HIP(hipStreamSynchronize(s));
  GP_HIP(hipGetLastError());
  X.eig = true;
  return GP_OK;
}

template <class T>
gp_status aux_eig_apply(const AuxAMG &X, int k, const double *om, const C2<T> *r, C2<T> *x, C2<double> *yp,
                        hipStream_t s) {
  const int64_t n = X.nL;
  const dim3 grid((unsigned)((n + 15) / 16), (unsigned)((2 * k + 15) / 16));
  k_eig_project<T><<<grid, 256, 0, s>>>(n, k, X.VT, X.lam, om, r, yp);
  k_eig_expand<T><<<grid, 256, 0, s>>>(n, k, X.V, yp, x);
  return GP_OK;
}

template gp_status aux_eig_apply<float>(const AuxAMG &, int, const double *, const C2<float> *, C2<float> *,
                                        C2<double> *, hipStream_t);
template gp_status aux_eig_apply<double>(const AuxAMG &, int, const double *, const C2<double> *, C2<double> *,
                                         C2<double> *, hipStream_t);

}  // namespace gp

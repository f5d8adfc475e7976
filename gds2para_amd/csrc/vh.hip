// vh.hip -- fdtdMesh::find_Vh(k) (src/fdtd.hpp:3577-3859): the Vh basis of
// the quadratic eigenproblem S u + lambda D_sig u + lambda^2 D_eps u = 0 from
// k Arnoldi steps on its companion operator.
//
// The reference runs Arnoldi on the 2n x 2n operator
//     M [v1; v2] = [v2; -D_eps^-1 S v1 / s^2 - D_eps^-1 D_sig v2 / s]
// (s = scale = 1e14 balances the two halves), from the constant start vector
// 1 / sqrt(2n), with modified Gram-Schmidt; takes the eigenvalues of the k x k
// Hessenberg matrix (LAPACK dhseqr), the eigenvectors (dhsein) of those with
// |Im| > eps, and Vh = (top half of V) * eigenvectors.  A Ritz pair (mu, y)
// of M gives lambda = s mu and u = V_top y with S u + lambda D_sig u +
// lambda^2 D_eps u ~ 0.
//
// Here the Arnoldi steps run on the device: the operator is one pass over
// the FP64 CSR of S with the per-edge eps_r and sigma of the system diagonal;
// the orthogonalization is classical Gram-Schmidt with one full
// re-orthogonalization pass (the Krylov subspace and Hessenberg matrix of
// the reference's modified Gram-Schmidt, to rounding).  The k x k eigenvalue
// problem runs on the host: Francis' implicit double-shift QR (Francis,
// 1961-62; the step of Golub and Van Loan, Matrix Computations, Alg. 7.5.1)
// for the eigenvalues and inverse iteration on H - mu I (complex Hessenberg
// elimination with partial pivoting) for the selected eigenvectors.  Vh is
// formed on the device.
#include <chrono>
#include <cmath>
#include <complex>

#include "mg.hpp"

namespace gp {

namespace {

typedef std::complex<double> cplx;

// w = M v (companion operator), rows i < n: w_i = v_{n+i};
// rows n + i: -(S v_top)_i / (eps_i s^2) - sigma_i v_{n+i} / (eps_i s), with
// e = mapEdgeR[i], L = e / (N_edge_s + N_edge_v), eps_i = stackEpsn[L] EPSILON0,
// sigma_i = stackSign[L] on conductor edges (markEdge[e] != 0), else 0.
__global__ void k_vh_apply(int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                           const double *__restrict__ sv, const int64_t *__restrict__ map_r,
                           const int32_t *__restrict__ mark_edge, const double *__restrict__ epsn,
                           const double *__restrict__ sign, int64_t El, double scale, const double *__restrict__ v,
                           double *__restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0;
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) acc += sv[p] * v[ci[p]];
  const int64_t e = map_r[i], L = e / El;
  const double eps = epsn[L] * kEps0;
  const double sig = mark_edge[e] != 0 ? sign[L] : 0.0;
  w[i] = v[n + i];
  w[n + i] = -acc / (eps * scale * scale) - v[n + i] * sig / (eps * scale);
}

// h[t] = V_t . w for t < nv (one block per basis vector, fixed-order block sum)
__global__ void __launch_bounds__(256) k_vh_dots(int64_t len, const double *__restrict__ V, const double *__restrict__ w,
                                                 double *__restrict__ h) {
  __shared__ double red[256];
  const int t = blockIdx.x;
  const double *vt = V + (int64_t)t * len;
  double s = 0;
  for (int64_t i = threadIdx.x; i < len; i += blockDim.x) s += vt[i] * w[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h2 = 128; h2 > 0; h2 >>= 1) {
    if ((int)threadIdx.x < h2) red[threadIdx.x] += red[threadIdx.x + h2];
    __syncthreads();
  }
  if (threadIdx.x == 0) h[t] = red[0];
}

// w -= sum_t h[t] V_t
__global__ void k_vh_update(int64_t len, int nv, const double *__restrict__ V, const double *__restrict__ h,
                            double *__restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  double s = w[i];
  for (int t = 0; t < nv; ++t) s -= h[t] * V[(int64_t)t * len + i];
  w[i] = s;
}

__global__ void k_vh_scale(int64_t len, const double *__restrict__ w, double inv, double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) out[i] = w[i] * inv;
}

// Vh[c][i] = sum_t V_t[i] Y[t][c]  (i < n: the top half; Y complex k x cols)
__global__ void k_vh_form(int64_t n, int64_t len, int k, int cols, const double *__restrict__ V,
                          const double2 *__restrict__ Y, double2 *__restrict__ Vh) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n) return;
  double re = 0, im = 0;
  for (int t = 0; t < k; ++t) {
    const double v = V[(int64_t)t * len + i];
    const double2 y = Y[(int64_t)t * cols + c];
    re += v * y.x;
    im += v * y.y;
  }
  Vh[(int64_t)c * n + i] = {re, im};
}

}  // namespace

// ---------------------------------------------------------------- host: k x k eigenproblem

// Eigenvalues of an upper Hessenberg matrix (row-major n x n, overwritten) by
// Francis' implicit double-shift QR.  A subdiagonal entry is set to zero when
// it is below eps times its diagonal neighbours; an exceptional shift every
// tenth step without a deflation; after 30 such steps the block is split at
// its relatively smallest subdiagonal entry if that is below 1e-7 (see below).
bool hessenberg_eigenvalues(std::vector<double> &H, int n, std::vector<double> &wr, std::vector<double> &wi) {
  auto A = [&](int r, int c) -> double & { return H[(size_t)r * n + c]; };
  wr.assign(n, 0.0);
  wi.assign(n, 0.0);
  double anorm = 0;
  for (int r = 0; r < n; ++r)
    for (int c = std::max(r - 1, 0); c < n; ++c) anorm += std::abs(A(r, c));
  const double eps = 2.220446049250313e-16, tiny = 2.2250738585072014e-308 * (n / eps);
  const int itmax = 30 * std::max(10, n);
  int hi = n - 1, its = 0;
  while (hi >= 0) {
    int l = hi;
    for (; l > 0; --l) {
      const double sub = std::abs(A(l, l - 1));
      if (sub <= tiny) break;
      double s = std::abs(A(l - 1, l - 1)) + std::abs(A(l, l));
      if (s == 0.0) s = anorm;
      if (sub <= eps * s) break;
    }
    if (l > 0) A(l, l - 1) = 0.0;
    if (l < hi - 1 && its >= 30) {
      // stalled: a nearly defective cluster, whose eigenvalues the data fix
      // only to about eps^(1/m), keeps its subdiagonal above eps |h|; split the
      // block at its relatively smallest subdiagonal entry once that is below
      // 1e-7 of its diagonal neighbours
      int best = -1;
      double rbest = 1e-7;
      for (int m = l + 1; m <= hi; ++m) {
        const double d = std::abs(A(m - 1, m - 1)) + std::abs(A(m, m));
        const double r = d > 0 ? std::abs(A(m, m - 1)) / d : 1.0;
        if (r <= rbest) { rbest = r; best = m; }
      }
      if (best > 0) {
        A(best, best - 1) = 0.0;
        its = 0;
        continue;
      }
    }
    if (l == hi) {                            // a 1 x 1 block
      wr[hi] = A(hi, hi); wi[hi] = 0.0;
      --hi; its = 0;
      continue;
    }
    if (l == hi - 1) {                        // a 2 x 2 block
      const double a = A(hi - 1, hi - 1), b = A(hi - 1, hi), c = A(hi, hi - 1), d = A(hi, hi);
      const double p = 0.5 * (a - d), q = p * p + b * c;
      if (q >= 0) {
        const double z = p >= 0 ? p + std::sqrt(q) : p - std::sqrt(q);
        wr[hi - 1] = d + z;
        wr[hi] = z != 0.0 ? d - b * c / z : d + z;
        wi[hi - 1] = wi[hi] = 0.0;
      } else {
        wr[hi - 1] = wr[hi] = d + p;
        wi[hi - 1] = std::sqrt(-q);
        wi[hi] = -std::sqrt(-q);
      }
      hi -= 2; its = 0;
      continue;
    }
    if (++its > itmax) return false;
    // shifts: the eigenvalues of the trailing 2 x 2 (sum s, product t), or an
    // exceptional pair every tenth step without a deflation
    double s, t;
    if (its % 10 == 0) {
      const double e = std::abs(A(hi, hi - 1)) + std::abs(A(hi - 1, hi - 2));
      s = 1.5 * e;
      t = e * e;
    } else {
      s = A(hi - 1, hi - 1) + A(hi, hi);
      t = A(hi - 1, hi - 1) * A(hi, hi) - A(hi - 1, hi) * A(hi, hi - 1);
    }
    // first column of (H - s1 I)(H - s2 I) restricted to the active block
    double x = A(l, l) * A(l, l) + A(l, l + 1) * A(l + 1, l) - s * A(l, l) + t;
    double y = A(l + 1, l) * (A(l, l) + A(l + 1, l + 1) - s);
    double z = A(l + 1, l) * A(l + 2, l + 1);
    for (int k = l; k <= hi - 2; ++k) {
      // Householder reflector P = I - 2 v v^T / v^T v with P (x, y, z)^T = (*, 0, 0)
      const double alpha = std::sqrt(x * x + y * y + z * z);
      if (alpha == 0.0) break;
      const double beta = x >= 0 ? -alpha : alpha;
      double v0 = x - beta, v1 = y, v2 = z;
      const double vn = v0 * v0 + v1 * v1 + v2 * v2;
      if (vn > 0) {
        const double f = 2.0 / vn;
        const int c0 = std::max(l, k - 1);
        for (int c = c0; c <= hi; ++c) {          // rows k..k+2, from the left
          const double d = f * (v0 * A(k, c) + v1 * A(k + 1, c) + v2 * A(k + 2, c));
          A(k, c) -= d * v0; A(k + 1, c) -= d * v1; A(k + 2, c) -= d * v2;
        }
        const int r1 = std::min(k + 3, hi);
        for (int r = l; r <= r1; ++r) {           // columns k..k+2, from the right
          const double d = f * (A(r, k) * v0 + A(r, k + 1) * v1 + A(r, k + 2) * v2);
          A(r, k) -= d * v0; A(r, k + 1) -= d * v1; A(r, k + 2) -= d * v2;
        }
      }
      x = A(k + 1, k);
      y = A(k + 2, k);
      if (k < hi - 2) z = A(k + 3, k);
    }
    {  // final 2 x 2 Givens rotation zeroing y below x at rows hi-1, hi
      const double r = std::hypot(x, y);
      if (r > 0) {
        const double c = x / r, sn = y / r;
        const int k = hi - 1;
        for (int cc = std::max(l, k - 1); cc <= hi; ++cc) {
          const double a = A(k, cc), b = A(k + 1, cc);
          A(k, cc) = c * a + sn * b;
          A(k + 1, cc) = -sn * a + c * b;
        }
        for (int rr = l; rr <= hi; ++rr) {
          const double a = A(rr, k), b = A(rr, k + 1);
          A(rr, k) = c * a + sn * b;
          A(rr, k + 1) = -sn * a + c * b;
        }
      }
    }
  }
  return true;
}

// An eigenvector of the Hessenberg matrix H (row-major n x n, unchanged) for
// the eigenvalue mu: inverse iteration on (H - mu I), elimination with
// partial pivoting on the Hessenberg structure (a row swap only with the next
// row), a pivot that vanishes replaced by eps |H|; three steps from a vector of
// ones, normalised to largest component 1.
std::vector<cplx> hessenberg_eigenvector(const std::vector<double> &H, int n, cplx mu, double hnorm) {
  std::vector<cplx> x(n, cplx(1.0, 0.0));
  const double tiny = 2.220446049250313e-16 * std::max(hnorm, 1e-300);
  for (int it = 0; it < 3; ++it) {
    std::vector<cplx> U((size_t)n * n);
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n; ++c) U[(size_t)r * n + c] = cplx(H[(size_t)r * n + c], 0.0) - (r == c ? mu : cplx(0.0));
    std::vector<cplx> b = x;
    // forward elimination of the subdiagonal
    for (int c = 0; c + 1 < n; ++c) {
      cplx &p = U[(size_t)c * n + c], &q = U[(size_t)(c + 1) * n + c];
      if (std::abs(q) > std::abs(p)) {
        for (int cc = c; cc < n; ++cc) std::swap(U[(size_t)c * n + cc], U[(size_t)(c + 1) * n + cc]);
        std::swap(b[c], b[c + 1]);
      }
      if (std::abs(U[(size_t)c * n + c]) < tiny) U[(size_t)c * n + c] = tiny;
      const cplx f = U[(size_t)(c + 1) * n + c] / U[(size_t)c * n + c];
      if (f != cplx(0.0))
        for (int cc = c; cc < n; ++cc) U[(size_t)(c + 1) * n + cc] -= f * U[(size_t)c * n + cc];
      b[c + 1] -= f * b[c];
    }
    if (std::abs(U[(size_t)(n - 1) * n + (n - 1)]) < tiny) U[(size_t)(n - 1) * n + (n - 1)] = tiny;
    for (int r = n - 1; r >= 0; --r) {
      cplx s = b[r];
      for (int c = r + 1; c < n; ++c) s -= U[(size_t)r * n + c] * x[c];
      x[r] = s / U[(size_t)r * n + r];
    }
    double mx = 0;
    int im = 0;
    for (int r = 0; r < n; ++r)
      if (std::abs(x[r]) > mx) { mx = std::abs(x[r]); im = r; }
    const cplx sc = x[im];
    for (cplx &v : x) v /= sc;
  }
  return x;
}

}  // namespace gp

using namespace gp;

extern "C" gp_status gp_find_vh(gp_layout *lay, int32_t k, double scale, double eps, double *hess, double *ritz,
                                int64_t max_cols, double *vh, gp_vh_info *info) {
  if (!lay || k < 2 || !(scale > 0) || !(eps >= 0) || max_cols < 0 || (max_cols > 0 && !vh))
    return fail(GP_ERR_ARG, "gp_find_vh: bad arguments");
  if (!lay->s_ready) GP_TRY(gp_assemble_stiffness(lay));
  const int64_t n = lay->n, len = 2 * n;
  if (k > len) return fail(GP_ERR_ARG, "gp_find_vh: more Arnoldi steps than the operator's dimension 2n");
  hipStream_t s = lay->stream;
  const auto t0 = std::chrono::steady_clock::now();
  Arena T;
  struct Free { Arena &a; ~Free() { a.free_all(); } } fr{T};
  double *V = nullptr, *w = nullptr, *h = nullptr, *h2 = nullptr;
  GP_TRY(T.alloc(&V, (int64_t)(k + 1) * len));
  GP_TRY(T.alloc(&w, len)); GP_TRY(T.alloc(&h, k + 1)); GP_TRY(T.alloc(&h2, k + 1));
  {
    std::vector<double> v0(len, 1.0 / std::sqrt((double)len));   // the reference's start vector
    GP_HIP(hipMemcpyAsync(V, v0.data(), sizeof(double) * len, hipMemcpyHostToDevice, s));
    GP_HIP(hipStreamSynchronize(s));
  }
  std::vector<double> Hm((size_t)k * k, 0.0), hv(k + 1), hv2(k + 1);
  double hnext = 0;
  const unsigned gl = blocks_for(len), gn = blocks_for(n);
  for (int j = 1; j <= k; ++j) {
    k_vh_apply<<<gn, 256, 0, s>>>(n, lay->S.rp, lay->S.ci, lay->S.v, lay->map_edge_r, lay->mark_edge, lay->epsn,
                                  lay->sign, lay->g.El, scale, V + (int64_t)(j - 1) * len, w);
    k_vh_dots<<<j, 256, 0, s>>>(len, V, w, h);
    k_vh_update<<<gl, 256, 0, s>>>(len, j, V, h, w);
    k_vh_dots<<<j, 256, 0, s>>>(len, V, w, h2);
    k_vh_update<<<gl, 256, 0, s>>>(len, j, V, h2, w);
    k_vh_dots<<<1, 256, 0, s>>>(len, w, w, h2 + j);   // |w|^2 after both passes
    GP_HIP(hipMemcpyAsync(hv.data(), h, sizeof(double) * j, hipMemcpyDeviceToHost, s));
    GP_HIP(hipMemcpyAsync(hv2.data(), h2, sizeof(double) * (j + 1), hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < j; ++i) Hm[(size_t)i * k + (j - 1)] = hv[i] + hv2[i];
    const double nrm = std::sqrt(std::max(hv2[j], 0.0));
    if (j < k) Hm[(size_t)j * k + (j - 1)] = nrm;
    else hnext = nrm;
    if (!(nrm > 0)) return fail(GP_ERR_STATE, "gp_find_vh: Arnoldi breakdown (invariant subspace) at step " +
                                                  std::to_string(j));
    k_vh_scale<<<gl, 256, 0, s>>>(len, w, 1.0 / nrm, V + (int64_t)j * len);
  }
  GP_HIP(hipStreamSynchronize(s));
  GP_HIP(hipGetLastError());
  const auto t1 = std::chrono::steady_clock::now();
  if (hess)
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) hess[(size_t)j * k + i] = Hm[(size_t)i * k + j];
  // eigenvalues of H (dhseqr) and the eigenvectors of |Im| > eps (dhsein)
  std::vector<double> Hc = Hm, wr, wi;
  if (!hessenberg_eigenvalues(Hc, k, wr, wi)) return fail(GP_ERR_NOCONV, "gp_find_vh: Hessenberg QR did not converge");
  if (ritz)
    for (int i = 0; i < k; ++i) { ritz[2 * i] = wr[i] * scale; ritz[2 * i + 1] = wi[i] * scale; }
  double hnorm = 0;
  for (double v : Hm) hnorm = std::max(hnorm, std::abs(v));
  std::vector<int> sel;
  for (int i = 0; i < k; ++i)
    if (std::abs(wi[i]) > eps) sel.push_back(i);
  const int64_t cols = (int64_t)sel.size();
  const auto t2 = std::chrono::steady_clock::now();
  if (info) info->leng_vh = (int32_t)cols;
  if (max_cols > 0 && cols > 0) {
    if (cols > max_cols) return fail(GP_ERR_ARG, "gp_find_vh: Vh has " + std::to_string(cols) + " columns, above max_cols");
    if (cols > 65535) return fail(GP_ERR_ARG, "gp_find_vh: more than 65535 Vh columns");
    std::vector<double2> Y((size_t)k * cols);
    for (int64_t c = 0; c < cols; ++c) {
      const int i = sel[c];
      if (c > 0 && sel[c - 1] == i - 1 && wi[i] < 0 && wi[i - 1] == -wi[i] && wr[i - 1] == wr[i]) {
        // the second of a conjugate pair: conj of the first column (dhsein's pair storage)
        for (int t = 0; t < k; ++t) Y[(size_t)t * cols + c] = {Y[(size_t)t * cols + c - 1].x, -Y[(size_t)t * cols + c - 1].y};
        continue;
      }
      const std::vector<cplx> y = hessenberg_eigenvector(Hm, k, cplx(wr[i], wi[i]), hnorm);
      for (int t = 0; t < k; ++t) Y[(size_t)t * cols + c] = {y[t].real(), y[t].imag()};
    }
    double2 *Yd = nullptr, *Vhd = nullptr;
    GP_TRY(T.alloc(&Yd, (int64_t)k * cols));
    GP_TRY(T.alloc(&Vhd, n * cols));
    GP_HIP(hipMemcpyAsync(Yd, Y.data(), sizeof(double2) * Y.size(), hipMemcpyHostToDevice, s));
    k_vh_form<<<dim3(gn, (unsigned)cols), 256, 0, s>>>(n, len, k, (int)cols, V, Yd, Vhd);
    GP_HIP(hipMemcpyAsync(vh, Vhd, sizeof(double2) * n * cols, hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    GP_HIP(hipGetLastError());
  }
  const auto t3 = std::chrono::steady_clock::now();
  if (info) {
    info->k = k;
    info->leng_vh = (int32_t)cols;
    info->h_next = hnext;
    info->arnoldi_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    info->eig_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    info->vh_ms = std::chrono::duration<double, std::milli>(t3 - t2).count();
  }
  return GP_OK;
}

extern "C" gp_status gp_hessenberg_eigenvalues(int32_t n, const double *h, double *w) {
  if (n < 1 || !h || !w) return fail(GP_ERR_ARG, "gp_hessenberg_eigenvalues: bad arguments");
  std::vector<double> H((size_t)n * n), wr, wi;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) H[(size_t)i * n + j] = i <= j + 1 ? h[(size_t)j * n + i] : 0.0;
  if (!hessenberg_eigenvalues(H, n, wr, wi)) return fail(GP_ERR_NOCONV, "gp_hessenberg_eigenvalues: QR did not converge");
  for (int i = 0; i < n; ++i) { w[2 * i] = wr[i]; w[2 * i + 1] = wi[i]; }
  return GP_OK;
}

extern "C" gp_status gp_hessenberg_eigenvector(int32_t n, const double *h, double re, double im, double *v) {
  if (n < 1 || !h || !v) return fail(GP_ERR_ARG, "gp_hessenberg_eigenvector: bad arguments");
  std::vector<double> H((size_t)n * n);
  double hnorm = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      H[(size_t)i * n + j] = i <= j + 1 ? h[(size_t)j * n + i] : 0.0;
      hnorm = std::max(hnorm, std::abs(H[(size_t)i * n + j]));
    }
  const std::vector<cplx> x = hessenberg_eigenvector(H, n, cplx(re, im), hnorm);
  for (int i = 0; i < n; ++i) { v[2 * i] = x[i].real(); v[2 * i + 1] = x[i].imag(); }
  return GP_OK;
}

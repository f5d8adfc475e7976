// dense.hip -- batched dense complex FP64 inverse (Gauss-Jordan, partial
// pivoting) for the coarsest multigrid levels: one matrix per frequency.
//
// The augmented matrices [A | I] (n x 2n per matrix) live in device memory.
// Each elimination step is two launches covering every matrix of the batch:
// k_gj_pivot (one block per matrix: pivot search, row swap, pivot-row scaling,
// and a copy of the pivot column) and k_gj_elim (all rows other than the pivot
// row, columns from the pivot column on).  Matrices here are a few hundred
// rows, so a step is L2-resident work; the whole inverse is O(n) launches.
#include <cmath>

#include "dense.hpp"

namespace gp {

namespace {

typedef C2<double> Z2;

__global__ void k_gj_init(int n, int nb, const Z2 *__restrict__ a, Z2 *__restrict__ aug) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t w = 2 * (int64_t)n;
  if (t >= (int64_t)nb * n * w) return;
  const int64_t b = t / (n * w), rem = t - b * n * w;
  const int64_t r = rem / w, c = rem - r * w;
  Z2 v;
  if (c < n) v = a[(b * n + r) * n + c];
  else v = {c - n == r ? 1.0 : 0.0, 0.0};
  aug[t] = v;
}

// One block per matrix.  bad[b] is set when the pivot column is exactly zero.
__global__ void __launch_bounds__(256) k_gj_pivot(int n, int c, Z2 *__restrict__ aug, Z2 *__restrict__ fcol,
                                                  int *__restrict__ bad) {
  __shared__ double sv[256];
  __shared__ int si[256];
  const int b = blockIdx.x;
  const int64_t w = 2 * (int64_t)n;
  Z2 *A = aug + (int64_t)b * n * w;
  double best = -1.0;
  int bi = c;
  for (int r = c + threadIdx.x; r < n; r += blockDim.x) {
    const Z2 v = A[r * w + c];
    const double m = v.x * v.x + v.y * v.y;
    if (m > best) { best = m; bi = r; }
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double o = sv[threadIdx.x + s];
      const int oi = si[threadIdx.x + s];
      if (o > sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) { sv[threadIdx.x] = o; si[threadIdx.x] = oi; }
    }
    __syncthreads();
  }
  const int piv = si[0];
  if (!(sv[0] > 0.0)) {
    if (threadIdx.x == 0) bad[b] = 1;
    return;
  }
  if (piv != c)
    for (int64_t j = c + threadIdx.x; j < w; j += blockDim.x) {
      const Z2 t = A[c * w + j];
      A[c * w + j] = A[(int64_t)piv * w + j];
      A[(int64_t)piv * w + j] = t;
    }
  __syncthreads();
  const Z2 p = A[c * w + c];
  const double d = p.x * p.x + p.y * p.y;
  const Z2 inv = {p.x / d, -p.y / d};
  __syncthreads();
  for (int64_t j = c + threadIdx.x; j < w; j += blockDim.x) {
    const Z2 v = A[c * w + j];
    A[c * w + j] = {v.x * inv.x - v.y * inv.y, v.x * inv.y + v.y * inv.x};
  }
  __syncthreads();
  for (int r = threadIdx.x; r < n; r += blockDim.x) fcol[(int64_t)b * n + r] = A[(int64_t)r * w + c];
}

// rows r != c: A[r][j] -= fcol[r] * A[c][j] for j >= c
__global__ void k_gj_elim(int n, int nb, int c, Z2 *__restrict__ aug, const Z2 *__restrict__ fcol,
                          const int *__restrict__ bad) {
  const int64_t w = 2 * (int64_t)n, wc = w - c;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nb * n * wc) return;
  const int64_t b = t / (n * wc), rem = t - b * n * wc;
  const int64_t r = rem / wc, j = c + (rem - r * wc);
  if (r == c || bad[b]) return;
  Z2 *A = aug + b * n * w;
  const Z2 f = fcol[b * n + r], v = A[(int64_t)c * w + j];
  Z2 x = A[r * w + j];
  x.x -= f.x * v.x - f.y * v.y;
  x.y -= f.x * v.y + f.y * v.x;
  A[r * w + j] = x;
}

__global__ void k_gj_extract(int n, int nb, const Z2 *__restrict__ aug, Z2 *__restrict__ a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nb * n * n) return;
  const int64_t b = t / ((int64_t)n * n), rem = t - b * n * n;
  const int64_t r = rem / n, c = rem - r * n;
  a[t] = aug[(b * n + r) * 2 * n + n + c];
}

}  // namespace

gp_status dense_inverse_batched(C2<double> *a, int n, int nb, hipStream_t s) {
  if (n <= 0 || nb <= 0) return GP_OK;
  Arena T;
  Z2 *aug, *fcol;
  int *bad;
  gp_status st = GP_OK;
  do {
    if ((st = T.alloc(&aug, (int64_t)nb * n * 2 * n)) != GP_OK) break;
    if ((st = T.alloc(&fcol, (int64_t)nb * n)) != GP_OK) break;
    if ((st = T.alloc(&bad, (int64_t)nb)) != GP_OK) break;
    if (hipMemsetAsync(bad, 0, sizeof(int) * nb, s) != hipSuccess) { st = fail(GP_ERR_HIP, "memset"); break; }
    k_gj_init<<<blocks_for((int64_t)nb * n * 2 * n), 256, 0, s>>>(n, nb, a, aug);
    for (int c = 0; c < n; ++c) {
      k_gj_pivot<<<nb, 256, 0, s>>>(n, c, aug, fcol, bad);
      k_gj_elim<<<blocks_for((int64_t)nb * n * (2 * (int64_t)n - c)), 256, 0, s>>>(n, nb, c, aug, fcol, bad);
    }
    k_gj_extract<<<blocks_for((int64_t)nb * n * n), 256, 0, s>>>(n, nb, aug, a);
    std::vector<int> hb(nb);
    if (hipMemcpyAsync(hb.data(), bad, sizeof(int) * nb, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "dense inverse"); break; }
    for (int b = 0; b < nb; ++b)
      if (hb[b]) { st = fail(GP_ERR_ARG, "singular coarsest-level matrix"); break; }
  } while (0);
  (void)hipStreamSynchronize(s);
  T.free_all();
  return st;
}

}  // namespace gp

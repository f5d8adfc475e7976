// dense.hip -- batched dense complex FP64 inverse (Gauss-Jordan, partial
// pivoting) for the coarsest multigrid levels: one matrix per frequency.
//
// The augmented matrices [A | I] (n x 2n per matrix) live in device memory.
// Matrices of at most kGJWorkgroupMax rows are eliminated by one workgroup
// each (k_gj_block), in a single launch.  Larger ones take two launches per
// elimination step covering every matrix of the batch: k_gj_pivot (one block
// per matrix: pivot search, row swap, pivot-row scaling, and a copy of the
// pivot column) and k_gj_elim (all rows other than the pivot row, columns from
// the pivot column on); the whole inverse is then O(n) launches.
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

// Gauss-Jordan with partial pivoting on one augmented matrix per workgroup
// (at most kGJWorkgroupMax rows), in device memory: the blocks in `list`, or,
// without a list, a uniform batch of nu-row matrices.  The steps are those of
// k_gj_pivot / k_gj_elim (the same pivot choice, ties to the smaller row, and
// the same arithmetic), ordered by the workgroup's barriers instead of by
// 2n launches; the matrix stays in L2.  bad[0] is set if a pivot column is
// exactly zero.
__global__ void __launch_bounds__(1024) k_gj_block(const int32_t *__restrict__ list, const int32_t *__restrict__ boff,
                                                   const int64_t *__restrict__ aoff, int nu, C2<double> *__restrict__ aug,
                                                   int *__restrict__ bad) {
  __shared__ double sv[1024];
  __shared__ int si[1024];
  __shared__ C2<double> fcol[kGJWorkgroupMax];
  int b;
  C2<double> *A;
  if (list) {
    const int blk = list[blockIdx.x];
    b = boff[blk + 1] - boff[blk];
    A = aug + aoff[blk];
  } else {
    b = nu;
    A = aug + (int64_t)blockIdx.x * 2 * nu * nu;
  }
  const int w = 2 * b;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int c = 0; c < b; ++c) {
    double best = -1.0;
    int bi = c;
    for (int r = c + tid; r < b; r += nt) {
      const C2<double> v = A[r * w + c];
      const double m = v.x * v.x + v.y * v.y;
      if (m > best) { best = m; bi = r; }
    }
    sv[tid] = best;
    si[tid] = bi;
    __syncthreads();
    for (int h = nt / 2; h > 0; h >>= 1) {
      if (tid < h) {
        const double o = sv[tid + h];
        const int oi = si[tid + h];
        if (o > sv[tid] || (o == sv[tid] && oi < si[tid])) { sv[tid] = o; si[tid] = oi; }
      }
      __syncthreads();
    }
    const int piv = si[0];
    if (!(sv[0] > 0.0)) {              // the same value for every thread: all leave together
      if (tid == 0) bad[list ? 0 : blockIdx.x] = 1;
      return;
    }
    if (piv != c)
      for (int j = c + tid; j < w; j += nt) {
        const C2<double> t = A[c * w + j];
        A[c * w + j] = A[piv * w + j];
        A[piv * w + j] = t;
      }
    __syncthreads();
    const C2<double> p = A[c * w + c];
    const double d = p.x * p.x + p.y * p.y;
    const C2<double> inv = {p.x / d, -p.y / d};
    __syncthreads();
    for (int j = c + tid; j < w; j += nt) {
      const C2<double> v = A[c * w + j];
      A[c * w + j] = {v.x * inv.x - v.y * inv.y, v.x * inv.y + v.y * inv.x};
    }
    __syncthreads();
    for (int r = tid; r < b; r += nt) fcol[r] = A[r * w + c];
    __syncthreads();
    const int wc = w - c;
    for (int t = tid; t < b * wc; t += nt) {
      const int r = t / wc, j = c + (t - r * wc);
      if (r == c) continue;
      const C2<double> f = fcol[r], v = A[c * w + j];
      C2<double> x = A[r * w + j];
      x.x -= f.x * v.x - f.y * v.y;
      x.y -= f.x * v.y + f.y * v.x;
      A[r * w + j] = x;
    }
    __syncthreads();
  }
}

__global__ void k_gj_extract(int n, int nb, const Z2 *__restrict__ aug, Z2 *__restrict__ a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nb * n * n) return;
  const int64_t b = t / ((int64_t)n * n), rem = t - b * n * n;
  const int64_t r = rem / n, c = rem - r * n;
  a[t] = aug[(b * n + r) * 2 * n + n + c];
}

}  // namespace

gp_status gj_workgroup(const int32_t *list, const int32_t *boff, const int64_t *aoff, int nu, int count, C2<double> *aug,
                       int *bad, hipStream_t s) {
  if (count <= 0) return GP_OK;
  k_gj_block<<<count, 1024, 0, s>>>(list, boff, aoff, nu, aug, bad);
  GP_HIP(hipGetLastError());
  return GP_OK;
}

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
    if (n <= kGJWorkgroupMax) {
      k_gj_block<<<nb, 1024, 0, s>>>(nullptr, nullptr, nullptr, n, aug, bad);
    } else {
      for (int c = 0; c < n; ++c) {
        k_gj_pivot<<<nb, 256, 0, s>>>(n, c, aug, fcol, bad);
        k_gj_elim<<<blocks_for((int64_t)nb * n * (2 * (int64_t)n - c)), 256, 0, s>>>(n, nb, c, aug, fcol, bad);
      }
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

// sparse.hip -- sort-based SpGEMM and transpose for the Galerkin hierarchy.
//
// C = A*B is formed by expansion: every product a_ik * b_kj becomes a
// (key = i*nc + j, index) pair, the pairs are radix-sorted by key (stable, so
// equal keys keep expansion order and the sums are deterministic), runs of
// equal keys are summed and the row pointers come from a histogram + scan.
// All arithmetic in FP64.
#include <hipcub/hipcub.hpp>

#include "sparse.hpp"

namespace gp {

namespace {

__global__ void k_expand_count(int64_t nr, const int64_t *__restrict__ arp, const int32_t *__restrict__ aci,
                               const int64_t *__restrict__ brp, int64_t *__restrict__ cnt) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  (void)nr;
  // one thread per nonzero of A: number of products it spawns
  if (p >= arp[nr]) return;
  int32_t k = aci[p];
  cnt[p] = brp[k + 1] - brp[k];
}

// row index of every nonzero of a CSR (binary search would also do; a scatter is simpler)
__global__ void k_row_of(int64_t nr, const int64_t *__restrict__ rp, int32_t *__restrict__ row) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) row[p] = (int32_t)i;
}

__global__ void k_expand(int64_t annz, const int32_t *__restrict__ arow, const int32_t *__restrict__ aci,
                         const int64_t *__restrict__ brp, const int32_t *__restrict__ bci,
                         const int64_t *__restrict__ off, int64_t nc, uint64_t *__restrict__ key,
                         uint32_t *__restrict__ src_a, uint32_t *__restrict__ src_b) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= annz) return;
  int64_t o = off[p];
  const int64_t i = arow[p];
  const int32_t k = aci[p];
  for (int64_t q = brp[k]; q < brp[k + 1]; ++q, ++o) {
    key[o] = (uint64_t)i * (uint64_t)nc + (uint64_t)bci[q];
    src_a[o] = (uint32_t)p;
    src_b[o] = (uint32_t)q;
  }
}

__global__ void k_iota(int64_t n, uint32_t *__restrict__ a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = (uint32_t)i;
}

__global__ void k_heads(int64_t n, const uint64_t *__restrict__ key, int64_t *__restrict__ head) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  head[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

// seg_start[s] = first sorted position of segment s
__global__ void k_seg_start(int64_t n, const int64_t *__restrict__ head, const int64_t *__restrict__ hscan,
                            int64_t *__restrict__ start) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (head[i]) start[hscan[i]] = i;
}

struct ValSrc {
  const double *a[3];
  const double *b[3];
  int na, nb, nout;
  int exact;  // 1: products and sums carried in double-double, rounded once at the end
};

// error-free transformations (Dekker/Knuth): a*b = p + e exactly, a+b = s + t exactly
__device__ __forceinline__ void two_prod(double a, double b, double &p, double &e) { p = a * b; e = fma(a, b, -p); }
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &t) {
  s = a + b;
  double bb = s - a;
  t = (a - (s - bb)) + (b - bb);
}

__global__ void k_reduce(int64_t nseg, int64_t n, const int64_t *__restrict__ start, const uint64_t *__restrict__ key,
                         const uint32_t *__restrict__ perm, const uint32_t *__restrict__ src_a,
                         const uint32_t *__restrict__ src_b, ValSrc vs, int64_t nc, int32_t *__restrict__ ci,
                         int64_t *__restrict__ rowcnt, double *__restrict__ o0, double *__restrict__ o1,
                         double *__restrict__ o2) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  int64_t b = start[s], e = (s + 1 < nseg) ? start[s + 1] : n;
  double acc[3] = {0., 0., 0.}, lo[3] = {0., 0., 0.};
  for (int64_t q = b; q < e; ++q) {
    uint32_t t = perm[q];
    uint32_t pa = src_a ? src_a[t] : t, pb = src_b ? src_b[t] : 0;
    for (int c = 0; c < vs.nout; ++c) {
      double va = vs.a[vs.na > 1 ? c : 0][pa];
      double vb = vs.nb > 0 ? vs.b[vs.nb > 1 ? c : 0][pb] : 1.0;
      if (vs.exact) {
        double p, pe, s2, t2;
        two_prod(va, vb, p, pe);
        two_sum(acc[c], p, s2, t2);
        acc[c] = s2;
        lo[c] += t2 + pe;
      } else {
        acc[c] += va * vb;
      }
    }
  }
  if (vs.exact)
    for (int c = 0; c < vs.nout; ++c) acc[c] += lo[c];
  uint64_t k = key[b];
  int64_t row = (int64_t)(k / (uint64_t)nc);
  ci[s] = (int32_t)(k % (uint64_t)nc);
  atomicAdd((unsigned long long *)&rowcnt[row], 1ULL);
  o0[s] = acc[0];
  if (o1) o1[s] = acc[1];
  if (o2) o2[s] = acc[2];
}

int bits_for(uint64_t maxkey) {
  int b = 1;
  while (b < 64 && (maxkey >> b) != 0) ++b;
  return b;
}

// Sort (key, src) pairs and reduce equal keys into a CSR with nout value arrays.
gp_status sort_reduce(Arena &A, int64_t nr, int64_t nc, int64_t nprod, uint64_t *key, const uint32_t *src_a,
                      const uint32_t *src_b, ValSrc vs, Csr *out, hipStream_t s) {
  Arena tmp;
  gp_status st = GP_OK;
  uint64_t *key2 = nullptr;
  uint32_t *perm = nullptr, *perm2 = nullptr;
  int64_t *head = nullptr, *hscan = nullptr, *start = nullptr, *rowcnt = nullptr;
  if (nprod > 2147483647LL) return fail(GP_ERR_NOMEM, "Galerkin expansion exceeds 2^31 products");
  do {
    if ((st = tmp.alloc(&key2, nprod)) != GP_OK) break;
    if ((st = tmp.alloc(&perm, nprod)) != GP_OK) break;
    if ((st = tmp.alloc(&perm2, nprod)) != GP_OK) break;
    k_iota<<<blocks_for(nprod), 256, 0, s>>>(nprod, perm);
    int endbit = bits_for((uint64_t)nr * (uint64_t)nc);
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, key, key2, perm, perm2, (int)nprod, 0, endbit, s) != hipSuccess) { st = fail(GP_ERR_HIP, "radix sort size"); break; }
    uint8_t *buf = nullptr;
    if ((st = tmp.alloc(&buf, (int64_t)bytes + 16)) != GP_OK) break;
    if (hipcub::DeviceRadixSort::SortPairs(buf, bytes, key, key2, perm, perm2, (int)nprod, 0, endbit, s) != hipSuccess) { st = fail(GP_ERR_HIP, "radix sort"); break; }
    tmp.release(buf, (int64_t)bytes + 16);
    tmp.release(perm, nprod);
    if ((st = tmp.alloc(&head, nprod)) != GP_OK) break;
    if ((st = tmp.alloc(&hscan, nprod + 1)) != GP_OK) break;
    k_heads<<<blocks_for(nprod), 256, 0, s>>>(nprod, key2, head);
    if ((st = exclusive_scan_i64(head, hscan, nprod, s, tmp)) != GP_OK) break;
    int64_t nseg = 0;
    if (hipMemcpyAsync(&nseg, hscan + nprod, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "segment count"); break; }
    if ((st = tmp.alloc(&start, nseg)) != GP_OK) break;
    k_seg_start<<<blocks_for(nprod), 256, 0, s>>>(nprod, head, hscan, start);
    if ((st = csr_alloc(A, nr, nc, nseg, vs.nout, out)) != GP_OK) break;
    if ((st = tmp.alloc(&rowcnt, nr)) != GP_OK) break;
    if (hipMemsetAsync(rowcnt, 0, sizeof(int64_t) * nr, s) != hipSuccess) { st = fail(GP_ERR_HIP, "memset"); break; }
    k_reduce<<<blocks_for(nseg), 256, 0, s>>>(nseg, nprod, start, key2, perm2, src_a, src_b, vs, nc, out->ci, rowcnt,
                                               out->v[0], vs.nout > 1 ? out->v[1] : nullptr,
                                               vs.nout > 2 ? out->v[2] : nullptr);
    if ((st = exclusive_scan_i64(rowcnt, out->rp, nr, s, tmp)) != GP_OK) break;
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) { st = fail(GP_ERR_HIP, "sort_reduce kernels"); break; }
  } while (0);
  (void)hipStreamSynchronize(s);
  tmp.free_all();
  return st;
}

}  // namespace

namespace {
__global__ void k_coo_keys(int64_t n, const int64_t *__restrict__ r, const int64_t *__restrict__ c, int64_t nc,
                           uint64_t *__restrict__ key) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) key[i] = (uint64_t)r[i] * (uint64_t)nc + (uint64_t)c[i];
}
__global__ void k_csr_triplets(int64_t nr, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                               const double *__restrict__ v, double alpha, int64_t base, int64_t *__restrict__ r,
                               int64_t *__restrict__ c, double *__restrict__ ov) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) { r[base + p] = i; c[base + p] = ci[p]; ov[base + p] = alpha * v[p]; }
}
}  // namespace

gp_status csr_from_coo(Arena &A, int64_t nr, int64_t nc, int64_t n, const int64_t *rows, const int64_t *cols,
                       const double *const *vals, int nv, Csr *out, hipStream_t s) {
  Arena tmp;
  uint64_t *key = nullptr;
  gp_status st = GP_OK;
  do {
    if (n == 0) { st = csr_alloc(A, nr, nc, 0, nv, out);
      if (st == GP_OK && hipMemsetAsync(out->rp, 0, sizeof(int64_t) * (nr + 1), s) != hipSuccess) st = fail(GP_ERR_HIP, "memset");
      break; }
    if ((st = tmp.alloc(&key, n)) != GP_OK) break;
    k_coo_keys<<<blocks_for(n), 256, 0, s>>>(n, rows, cols, nc, key);
    ValSrc vs{};
    vs.na = nv; vs.nb = 0; vs.nout = nv;
    for (int k = 0; k < nv; ++k) vs.a[k] = vals[k];
    st = sort_reduce(A, nr, nc, n, key, nullptr, nullptr, vs, out, s);
  } while (0);
  (void)hipStreamSynchronize(s);
  tmp.free_all();
  return st;
}

gp_status csr_add(Arena &A, const Csr &a, double alpha, const Csr &b, Csr *c, hipStream_t s) {
  if (a.nr != b.nr || a.nc != b.nc) return fail(GP_ERR_ARG, "csr_add shape mismatch");
  Arena tmp;
  int64_t *r = nullptr, *cc = nullptr;
  double *v = nullptr;
  const int64_t n = a.nnz + b.nnz;
  gp_status st = GP_OK;
  do {
    if ((st = tmp.alloc(&r, std::max<int64_t>(n, 1))) != GP_OK) break;
    if ((st = tmp.alloc(&cc, std::max<int64_t>(n, 1))) != GP_OK) break;
    if ((st = tmp.alloc(&v, std::max<int64_t>(n, 1))) != GP_OK) break;
    k_csr_triplets<<<blocks_for(a.nr), 256, 0, s>>>(a.nr, a.rp, a.ci, a.v[0], 1.0, 0, r, cc, v);
    k_csr_triplets<<<blocks_for(b.nr), 256, 0, s>>>(b.nr, b.rp, b.ci, b.v[0], alpha, a.nnz, r, cc, v);
    const double *vals[1] = {v};
    st = csr_from_coo(A, a.nr, a.nc, n, r, cc, vals, 1, c, s);
  } while (0);
  (void)hipStreamSynchronize(s);
  tmp.free_all();
  return st;
}

namespace {
__global__ void k_nz_flags(int64_t nnz, const double *__restrict__ v, int64_t *__restrict__ f) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nnz) f[i] = v[i] != 0.0;
}
__global__ void k_nz_compact(int64_t nr, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                             const double *__restrict__ v, const int64_t *__restrict__ pos, int64_t *__restrict__ r,
                             int64_t *__restrict__ c, double *__restrict__ ov) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p)
    if (v[p] != 0.0) { r[pos[p]] = i; c[pos[p]] = ci[p]; ov[pos[p]] = v[p]; }
}
}  // namespace

gp_status csr_drop_zeros(Arena &A, const Csr &a, Csr *out, hipStream_t s) {
  Arena tmp;
  gp_status st = GP_OK;
  int64_t *f = nullptr, *pos = nullptr, *r = nullptr, *c = nullptr;
  double *v = nullptr;
  do {
    if ((st = tmp.alloc(&f, std::max<int64_t>(a.nnz, 1))) != GP_OK) break;
    if ((st = tmp.alloc(&pos, a.nnz + 1)) != GP_OK) break;
    if (a.nnz) k_nz_flags<<<blocks_for(a.nnz), 256, 0, s>>>(a.nnz, a.v[0], f);
    if ((st = exclusive_scan_i64(f, pos, a.nnz, s, tmp)) != GP_OK) break;
    int64_t nk = 0;
    if (hipMemcpyAsync(&nk, pos + a.nnz, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "count"); break; }
    if ((st = tmp.alloc(&r, std::max<int64_t>(nk, 1))) != GP_OK) break;
    if ((st = tmp.alloc(&c, std::max<int64_t>(nk, 1))) != GP_OK) break;
    if ((st = tmp.alloc(&v, std::max<int64_t>(nk, 1))) != GP_OK) break;
    k_nz_compact<<<blocks_for(a.nr), 256, 0, s>>>(a.nr, a.rp, a.ci, a.v[0], pos, r, c, v);
    const double *vals[1] = {v};
    st = csr_from_coo(A, a.nr, a.nc, nk, r, c, vals, 1, out, s);
  } while (0);
  (void)hipStreamSynchronize(s);
  tmp.free_all();
  return st;
}

gp_status csr_alloc(Arena &A, int64_t nr, int64_t nc, int64_t nnz, int nv, Csr *out) {
  *out = Csr();
  out->nr = nr; out->nc = nc; out->nnz = nnz; out->nv = nv;
  GP_TRY(A.alloc(&out->rp, nr + 1));
  GP_TRY(A.alloc(&out->ci, std::max<int64_t>(nnz, 1)));
  for (int k = 0; k < nv; ++k) GP_TRY(A.alloc(&out->v[k], std::max<int64_t>(nnz, 1)));
  return GP_OK;
}

void csr_release(Arena &A, Csr &c) {
  A.release(c.rp, c.nr + 1);
  A.release(c.ci, std::max<int64_t>(c.nnz, 1));
  for (int k = 0; k < c.nv; ++k) A.release(c.v[k], std::max<int64_t>(c.nnz, 1));
  c = Csr();
}

gp_status csr_upload(Arena &A, const HostCsr &h, Csr *out, hipStream_t s) {
  int64_t nnz = (int64_t)h.ci.size();
  GP_TRY(csr_alloc(A, h.nr, h.nc, nnz, 1, out));
  GP_HIP(hipMemcpyAsync(out->rp, h.rp.data(), sizeof(int64_t) * (h.nr + 1), hipMemcpyHostToDevice, s));
  if (nnz) {
    GP_HIP(hipMemcpyAsync(out->ci, h.ci.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice, s));
    GP_HIP(hipMemcpyAsync(out->v[0], h.v.data(), sizeof(double) * nnz, hipMemcpyHostToDevice, s));
  }
  GP_HIP(hipStreamSynchronize(s));
  return GP_OK;
}

gp_status spgemm(Arena &A, const Csr &a, const Csr &b, Csr *c, hipStream_t s) { return spgemm_ex(A, a, b, c, s, 0); }

gp_status spgemm_ex(Arena &A, const Csr &a, const Csr &b, Csr *c, hipStream_t s, int exact) {
  if (a.nc != b.nr) return fail(GP_ERR_ARG, "spgemm shape mismatch");
  if (a.nv > 1 && b.nv > 1 && a.nv != b.nv) return fail(GP_ERR_ARG, "spgemm value arrays mismatch");
  Arena tmp;
  gp_status st = GP_OK;
  int64_t *cnt = nullptr, *off = nullptr;
  int32_t *arow = nullptr;
  uint64_t *key = nullptr;
  uint32_t *sa = nullptr, *sb = nullptr;
  do {
    if (a.nnz == 0) { st = csr_alloc(A, a.nr, b.nc, 0, std::max(a.nv, b.nv), c);
      if (st == GP_OK && hipMemsetAsync(c->rp, 0, sizeof(int64_t) * (a.nr + 1), s) != hipSuccess) st = fail(GP_ERR_HIP, "memset");
      break; }
    if ((st = tmp.alloc(&cnt, a.nnz)) != GP_OK) break;
    if ((st = tmp.alloc(&off, a.nnz + 1)) != GP_OK) break;
    if ((st = tmp.alloc(&arow, a.nnz)) != GP_OK) break;
    k_row_of<<<blocks_for(a.nr), 256, 0, s>>>(a.nr, a.rp, arow);
    k_expand_count<<<blocks_for(a.nnz), 256, 0, s>>>(a.nr, a.rp, a.ci, b.rp, cnt);
    if ((st = exclusive_scan_i64(cnt, off, a.nnz, s, tmp)) != GP_OK) break;
    int64_t nprod = 0;
    if (hipMemcpyAsync(&nprod, off + a.nnz, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "product count"); break; }
    tmp.release(cnt, a.nnz);
    if ((st = tmp.alloc(&key, std::max<int64_t>(nprod, 1))) != GP_OK) break;
    if ((st = tmp.alloc(&sa, std::max<int64_t>(nprod, 1))) != GP_OK) break;
    if ((st = tmp.alloc(&sb, std::max<int64_t>(nprod, 1))) != GP_OK) break;
    k_expand<<<blocks_for(a.nnz), 256, 0, s>>>(a.nnz, arow, a.ci, b.rp, b.ci, off, b.nc, key, sa, sb);
    ValSrc vs{};
    vs.na = a.nv; vs.nb = b.nv; vs.nout = std::max(a.nv, b.nv);
    vs.exact = exact;
    for (int k = 0; k < 3; ++k) { vs.a[k] = a.v[k]; vs.b[k] = b.v[k]; }
    if (nprod == 0) { st = csr_alloc(A, a.nr, b.nc, 0, vs.nout, c);
      if (st == GP_OK && hipMemsetAsync(c->rp, 0, sizeof(int64_t) * (a.nr + 1), s) != hipSuccess) st = fail(GP_ERR_HIP, "memset");
      break; }
    st = sort_reduce(A, a.nr, b.nc, nprod, key, sa, sb, vs, c, s);
  } while (0);
  (void)hipStreamSynchronize(s);
  tmp.free_all();
  return st;
}

namespace {
__global__ void k_transpose_keys(int64_t nr, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                                 int64_t nrow_t, uint64_t *__restrict__ key) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) key[p] = (uint64_t)ci[p] * (uint64_t)nr + (uint64_t)i;
  (void)nrow_t;
}
}  // namespace

gp_status transpose(Arena &A, const Csr &a, Csr *t, hipStream_t s) {
  Arena tmp;
  uint64_t *key = nullptr;
  gp_status st = GP_OK;
  do {
    if (a.nnz == 0) { st = csr_alloc(A, a.nc, a.nr, 0, a.nv, t);
      if (st == GP_OK && hipMemsetAsync(t->rp, 0, sizeof(int64_t) * (a.nc + 1), s) != hipSuccess) st = fail(GP_ERR_HIP, "memset");
      break; }
    if ((st = tmp.alloc(&key, a.nnz)) != GP_OK) break;
    k_transpose_keys<<<blocks_for(a.nr), 256, 0, s>>>(a.nr, a.rp, a.ci, a.nc, key);
    ValSrc vs{};
    vs.na = a.nv; vs.nb = 0; vs.nout = a.nv;
    for (int k = 0; k < 3; ++k) vs.a[k] = a.v[k];
    st = sort_reduce(A, a.nc, a.nr, a.nnz, key, nullptr, nullptr, vs, t, s);
  } while (0);
  (void)hipStreamSynchronize(s);
  tmp.free_all();
  return st;
}

namespace {
__global__ void k_d2f(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
}  // namespace

gp_status csr_to_float(Arena &A, const Csr &a, CsrF *f, hipStream_t s) {
  *f = CsrF();
  f->nr = a.nr; f->nc = a.nc; f->nnz = a.nnz; f->rp = a.rp; f->ci = a.ci; f->nv = a.nv;
  f->nlong = a.nlong; f->lrows = a.lrows; f->lidx = a.lidx;
  f->nchunk = a.nchunk; f->lcb = a.lcb; f->lce = a.lce; f->lqp = a.lqp;
  for (int k = 0; k < a.nv; ++k) {
    GP_TRY(A.alloc(&f->v[k], std::max<int64_t>(a.nnz, 1)));
    if (a.nnz) k_d2f<<<blocks_for(a.nnz), 256, 0, s>>>(a.nnz, a.v[k], f->v[k]);
  }
  GP_HIP(hipStreamSynchronize(s));
  return GP_OK;
}

gp_status csr_mark_long(Arena &A, Csr &a, hipStream_t s) {
  a.nlong = 0; a.lrows = nullptr; a.lidx = nullptr;
  a.nchunk = 0; a.lcb = a.lce = nullptr; a.lqp = nullptr;
  if (a.nr == 0) return GP_OK;
  std::vector<int64_t> rp(a.nr + 1);
  GP_HIP(hipMemcpyAsync(rp.data(), a.rp, sizeof(int64_t) * (a.nr + 1), hipMemcpyDeviceToHost, s));
  GP_HIP(hipStreamSynchronize(s));
  std::vector<int32_t> rows, idx(a.nr, -1), qp(1, 0);
  std::vector<int64_t> cb, ce;
  for (int64_t i = 0; i < a.nr; ++i)
    if (rp[i + 1] - rp[i] > kLongRow) {
      idx[i] = (int32_t)rows.size();
      rows.push_back((int32_t)i);
      for (int64_t p = rp[i]; p < rp[i + 1]; p += kLongChunk) {
        cb.push_back(p);
        ce.push_back(std::min(p + kLongChunk, rp[i + 1]));
      }
      qp.push_back((int32_t)cb.size());
    }
  if (rows.empty()) return GP_OK;
  GP_TRY(A.alloc(&a.lrows, (int64_t)rows.size()));
  GP_TRY(A.alloc(&a.lidx, a.nr));
  GP_TRY(A.alloc(&a.lcb, (int64_t)cb.size()));
  GP_TRY(A.alloc(&a.lce, (int64_t)ce.size()));
  GP_TRY(A.alloc(&a.lqp, (int64_t)qp.size()));
  GP_HIP(hipMemcpyAsync(a.lrows, rows.data(), sizeof(int32_t) * rows.size(), hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(a.lidx, idx.data(), sizeof(int32_t) * a.nr, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(a.lcb, cb.data(), sizeof(int64_t) * cb.size(), hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(a.lce, ce.data(), sizeof(int64_t) * ce.size(), hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(a.lqp, qp.data(), sizeof(int32_t) * qp.size(), hipMemcpyHostToDevice, s));
  GP_HIP(hipStreamSynchronize(s));
  a.nlong = (int32_t)rows.size();
  a.nchunk = (int32_t)cb.size();
  return GP_OK;
}

gp_status csr_host_copy(const Csr &a, HostCsr *h, int which) {
  h->nr = a.nr; h->nc = a.nc;
  h->rp.resize(a.nr + 1);
  h->ci.resize(a.nnz);
  h->v.resize(a.nnz);
  GP_HIP(hipMemcpy(h->rp.data(), a.rp, sizeof(int64_t) * (a.nr + 1), hipMemcpyDeviceToHost));
  if (a.nnz) {
    GP_HIP(hipMemcpy(h->ci.data(), a.ci, sizeof(int32_t) * a.nnz, hipMemcpyDeviceToHost));
    GP_HIP(hipMemcpy(h->v.data(), a.v[which], sizeof(double) * a.nnz, hipMemcpyDeviceToHost));
  }
  return GP_OK;
}

namespace {

__global__ void k_sell_width(int64_t nr, int64_t nslice, const int64_t *__restrict__ rp, const int32_t *__restrict__ lidx,
                             const int32_t *__restrict__ order, int32_t *__restrict__ w, int64_t *__restrict__ cnt) {
  int64_t sl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= nslice) return;
  int64_t m = 0;
  for (int64_t p = sl * 64; p < min(nr, sl * 64 + 64); ++p) {
    const int64_t i = order ? (int64_t)order[p] : p;
    if (!(lidx && lidx[i] >= 0)) m = max(m, rp[i + 1] - rp[i]);
  }
  w[sl] = (int32_t)m;
  cnt[sl] = 64 * m;
}

__global__ void k_sell_fill(int64_t nr, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci, int nv,
                            const double *__restrict__ v0, const double *__restrict__ v1, const double *__restrict__ v2,
                            const int32_t *__restrict__ lidx, const int32_t *__restrict__ order,
                            const int64_t *__restrict__ base, const int32_t *__restrict__ w,
                            int32_t *__restrict__ oci, double *__restrict__ o0, double *__restrict__ o1,
                            double *__restrict__ o2) {
  // one thread per slot position of every slice, so the lanes past the last
  // row (never read by the kernels) hold valid padding too
  int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= ((nr + 63) >> 6) << 6) return;
  const int64_t sl = pos >> 6, lane = pos & 63;
  const int64_t b = base[sl];
  const int64_t i = pos < nr ? (order ? (int64_t)order[pos] : pos) : pos;
  const bool skip = pos >= nr || (lidx && lidx[i] >= 0);
  const int64_t len = skip ? 0 : rp[i + 1] - rp[i];
  for (int64_t t = 0; t < w[sl]; ++t) {
    const int64_t e = b + 64 * t + lane;
    if (t < len) {
      const int64_t p = rp[i] + t;
      oci[e] = ci[p];
      o0[e] = v0[p];
      if (nv > 1) o1[e] = v1[p];
      if (nv > 2) o2[e] = v2[p];
    } else {
      // padding: value 0 at a column that exists (the row's last one, else
      // column 0), so the gather it triggers stays inside the input vector
      oci[e] = len > 0 ? ci[rp[i] + len - 1] : 0;
      o0[e] = 0.;
      if (nv > 1) o1[e] = 0.;
      if (nv > 2) o2[e] = 0.;
    }
  }
}

// bad[0] += 1 for every stored column index outside [0, nc); bad[1], bad[2]:
// smallest and largest index seen.  Each wavefront reduces its 64 lanes
// (a grid-stride loop of them) first and issues one atomic of each kind.
__global__ void __launch_bounds__(256) k_sell_check(int64_t n, int64_t nc, const int32_t *__restrict__ ci,
                                                    int *__restrict__ bad) {
  int cnt = 0, lo = 2147483647, hi = -2147483647 - 1;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int c = ci[e];
    cnt += (c < 0 || (int64_t)c >= nc) ? 1 : 0;
    lo = min(lo, c);
    hi = max(hi, c);
  }
  for (int off = 32; off > 0; off >>= 1) {
    cnt += __shfl_xor(cnt, off, 64);
    lo = min(lo, __shfl_xor(lo, off, 64));
    hi = max(hi, __shfl_xor(hi, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    if (cnt) atomicAdd(bad, cnt);
    atomicMin(bad + 1, lo);
    atomicMax(bad + 2, hi);
  }
}

__global__ void k_d2f_sell(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

}  // namespace

gp_status sell_from_csr(Arena &A, const Csr &a, Sell *out, hipStream_t s, const char *name, const int32_t *order) {
  *out = Sell();
  out->nr = a.nr;
  out->nv = a.nv;
  out->nslice = (a.nr + 63) / 64;
  if (a.nr == 0) return GP_OK;
  Arena tmp;
  gp_status st = GP_OK;
  do {
    int64_t *cnt = nullptr;
    if ((st = A.alloc(&out->w, out->nslice)) != GP_OK) break;
    if ((st = A.alloc(&out->base, out->nslice + 1)) != GP_OK) break;
    if ((st = tmp.alloc(&cnt, out->nslice)) != GP_OK) break;
    if (order) {
      if ((st = A.alloc(&out->rows, a.nr)) != GP_OK) break;
      if (hipMemcpyAsync(out->rows, order, sizeof(int32_t) * a.nr, hipMemcpyDeviceToDevice, s) != hipSuccess) { st = fail(GP_ERR_HIP, "copy"); break; }
    }
    k_sell_width<<<blocks_for(out->nslice), 256, 0, s>>>(a.nr, out->nslice, a.rp, a.lidx, out->rows, out->w, cnt);
    if ((st = exclusive_scan_i64(cnt, out->base, out->nslice, s, tmp)) != GP_OK) break;
    if (hipMemcpyAsync(&out->nent, out->base + out->nslice, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "sell size"); break; }
    const int64_t ne = std::max<int64_t>(out->nent, 1);
    if ((st = A.alloc(&out->ci, ne)) != GP_OK) break;
    for (int q = 0; q < a.nv; ++q)
      if ((st = A.alloc(&out->v[q], ne)) != GP_OK) break;
    if (st != GP_OK) break;
    k_sell_fill<<<blocks_for(out->nslice * 64), 256, 0, s>>>(a.nr, a.rp, a.ci, a.nv, a.v[0], a.v[1], a.v[2], a.lidx,
                                                             out->rows, out->base, out->w, out->ci, out->v[0],
                                                             out->v[1], out->v[2]);
    // every column the kernels will gather from must exist
    int *bad = nullptr, hb[3] = {0, 2147483647, -1};
    if ((st = tmp.alloc(&bad, 3)) != GP_OK) break;
    if (hipMemcpyAsync(bad, hb, sizeof(hb), hipMemcpyHostToDevice, s) != hipSuccess) { st = fail(GP_ERR_HIP, "copy"); break; }
    if (out->nent)
      k_sell_check<<<(unsigned)std::min<int64_t>(blocks_for(out->nent), 4096), 256, 0, s>>>(out->nent, a.nc, out->ci, bad);
    if (hipMemcpyAsync(hb, bad, sizeof(hb), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) { st = fail(GP_ERR_HIP, "sell fill"); break; }
    if (hb[0]) {
      st = fail(GP_ERR_STATE, std::string("SELL copy of ") + name + " (" + std::to_string(a.nr) + "x" +
                                  std::to_string(a.nc) + ", nnz " + std::to_string(a.nnz) + ", " +
                                  std::to_string(out->nent) + " slots) holds " + std::to_string(hb[0]) +
                                  " out-of-range column indices, range [" + std::to_string(hb[1]) + ", " +
                                  std::to_string(hb[2]) + "]");
      break;
    }
  } while (0);
  tmp.free_all();
  return st;
}

gp_status sell_to_float(Arena &A, const Sell &a, SellF *f, hipStream_t s) {
  *f = SellF();
  f->nr = a.nr; f->nslice = a.nslice; f->nent = a.nent; f->base = a.base; f->w = a.w; f->ci = a.ci; f->nv = a.nv;
  f->rows = a.rows;
  for (int q = 0; q < a.nv; ++q) {
    GP_TRY(A.alloc(&f->v[q], std::max<int64_t>(a.nent, 1)));
    if (a.nent) k_d2f_sell<<<blocks_for(a.nent), 256, 0, s>>>(a.nent, a.v[q], f->v[q]);
  }
  GP_HIP(hipStreamSynchronize(s));
  return GP_OK;
}


namespace {
// own(i): the column of row i of the tentative prolongation (one entry per row)
__device__ __forceinline__ int32_t own_col(const int64_t *trp, const int32_t *tci, int64_t i) {
  return trp[i] < trp[i + 1] ? tci[trp[i]] : -1;
}
__device__ __forceinline__ bool keep_entry(double v, int32_t c, double cut, int32_t own) {
  return fabs(v) >= cut || c == own;
}
__global__ void k_trunc_count(int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                              const double *__restrict__ v, const int64_t *__restrict__ trp,
                              const int32_t *__restrict__ tci, double theta, int64_t *__restrict__ cnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double mx = 0.;
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) mx = fmax(mx, fabs(v[p]));
  const int32_t own = own_col(trp, tci, i);
  int64_t c = 0;
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) c += keep_entry(v[p], ci[p], theta * mx, own) ? 1 : 0;
  cnt[i] = c;
}
__global__ void k_trunc_fill(int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                             const double *__restrict__ v, const int64_t *__restrict__ trp,
                             const int32_t *__restrict__ tci, double theta, const int64_t *__restrict__ orp,
                             int32_t *__restrict__ oci, double *__restrict__ ov) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double mx = 0., rs = 0., rk = 0.;
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) { mx = fmax(mx, fabs(v[p])); rs += v[p]; }
  const int32_t own = own_col(trp, tci, i);
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p)
    if (keep_entry(v[p], ci[p], theta * mx, own)) rk += v[p];
  const double sc = rk != 0. ? rs / rk : 1.;
  int64_t q = orp[i];
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p)
    if (keep_entry(v[p], ci[p], theta * mx, own)) { oci[q] = ci[p]; ov[q] = v[p] * sc; ++q; }
}
}  // namespace

gp_status truncate_prolongation(Arena &A, Arena &tmp, const Csr &P, const Csr &Pt, double theta, Csr *out,
                                hipStream_t s) {
  const int64_t n = P.nr;
  if (Pt.nr != n) return fail(GP_ERR_ARG, "truncate_prolongation: P and Pt differ in rows");
  int64_t *cnt, *off;
  GP_TRY(tmp.alloc(&cnt, std::max<int64_t>(n, 1))); GP_TRY(tmp.alloc(&off, n + 1));
  k_trunc_count<<<blocks_for(n), 256, 0, s>>>(n, P.rp, P.ci, P.v[0], Pt.rp, Pt.ci, theta, cnt);
  GP_TRY(exclusive_scan_i64(cnt, off, n, s, tmp));
  int64_t nk = 0;
  GP_HIP(hipMemcpyAsync(&nk, off + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  GP_HIP(hipStreamSynchronize(s));
  GP_TRY(csr_alloc(A, n, P.nc, nk, 1, out));
  GP_HIP(hipMemcpyAsync(out->rp, off, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToDevice, s));
  k_trunc_fill<<<blocks_for(n), 256, 0, s>>>(n, P.rp, P.ci, P.v[0], Pt.rp, Pt.ci, theta, off, out->ci, out->v[0]);
  GP_HIP(hipStreamSynchronize(s));
  return GP_OK;
}

}  // namespace gp

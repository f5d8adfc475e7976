// sparse.hpp -- device sparse matrices used by the multigrid hierarchy:
// CSR with int64 row pointers, int32 columns and up to three FP64 value arrays
// sharing one pattern, plus the sort-based (expand / radix-sort / reduce)
// SpGEMM and transpose that build the Galerkin coarse operators on the GPU.
#pragma once
#include <vector>

#include "gp_dev.hpp"
#include "host_csr.hpp"

namespace gp {

// Rows with more than kLongRow entries (the per-conductor columns of the V0
// basis couple to a whole conductor surface) are listed in lrows; lidx[i] is
// the row's position in that list or -1.  Row-parallel kernels take those rows'
// products from a block-per-row reduction instead of walking them serially.
constexpr int64_t kLongRow = 256;
constexpr int64_t kLongChunk = 256;   // entries per block of the long-row reduction

struct Csr {
  int64_t nr = 0, nc = 0, nnz = 0;
  int64_t *rp = nullptr;
  int32_t *ci = nullptr;
  int nv = 0;
  double *v[3] = {nullptr, nullptr, nullptr};
  int32_t nlong = 0;
  int32_t *lrows = nullptr, *lidx = nullptr;
  // the long rows cut into chunks of kLongChunk entries: chunk c spans
  // [lcb[c], lce[c]); long row q owns chunks lqp[q] .. lqp[q+1]-1
  int32_t nchunk = 0;
  int64_t *lcb = nullptr, *lce = nullptr;
  int32_t *lqp = nullptr;
  // when set, value blocks 1 and 2 are diagonal and live here (one per row)
  double *dg[2] = {nullptr, nullptr};
  // when set (level 0), the diagonal of value block 0; the SELL copy then
  // holds the off-diagonal entries only (nsdg diagonal entries moved here)
  double *sdg = nullptr;
  int64_t nsdg = 0;
};

// FP32 copy of a Csr's values; row pointers and columns are shared.
struct CsrF {
  int64_t nr = 0, nc = 0, nnz = 0;
  const int64_t *rp = nullptr;
  const int32_t *ci = nullptr;
  int nv = 0;
  float *v[3] = {nullptr, nullptr, nullptr};
  int32_t nlong = 0;
  const int32_t *lrows = nullptr, *lidx = nullptr;
  int32_t nchunk = 0;
  const int64_t *lcb = nullptr, *lce = nullptr;
  const int32_t *lqp = nullptr;
  float *dg[2] = {nullptr, nullptr};
  float *sdg = nullptr;
};

// Sliced ELL with 64-row slices (one wavefront): slice s holds w[s] entry
// columns; entry t of row 64 s + l sits at base[s] + 64 t + l, so a wavefront
// reads each entry column of its 64 rows as one coalesced 256-byte line.
// Rows shorter than the slice width are padded with their own index and value 0.
struct Sell {
  int64_t nr = 0, nslice = 0, nent = 0;
  int64_t *base = nullptr;
  int32_t *w = nullptr;
  int32_t *ci = nullptr;
  int nv = 0;
  double *v[3] = {nullptr, nullptr, nullptr};
  int32_t *rows = nullptr;  // slot position -> row (nullptr: position = row)
};
struct SellF {
  int64_t nr = 0, nslice = 0, nent = 0;
  const int64_t *base = nullptr;
  const int32_t *w = nullptr;
  const int32_t *ci = nullptr;
  int nv = 0;
  float *v[3] = {nullptr, nullptr, nullptr};
  const int32_t *rows = nullptr;
};
// SELL copy of a CSR (all value arrays); rows listed as long in a (kLongRow)
// are left empty, their products come from the long-row pass
// order (device, nr entries, a permutation of the rows, or nullptr): slice
// position p holds row order[p], so a launch over positions visits the rows in
// that order; kernels map position -> row through Sell::rows
gp_status sell_from_csr(Arena &A, const Csr &a, Sell *out, hipStream_t s, const char *name,
                        const int32_t *order = nullptr);
gp_status sell_to_float(Arena &A, const Sell &a, SellF *f, hipStream_t s);

gp_status csr_upload(Arena &A, const HostCsr &h, Csr *out, hipStream_t s);
gp_status csr_alloc(Arena &A, int64_t nr, int64_t nc, int64_t nnz, int nv, Csr *out);
void csr_release(Arena &A, Csr &c);
// C = A * B; the value arrays multiply pairwise when both have nv > 1
// (nv must match), otherwise the single array scales every array of the other.
gp_status spgemm(Arena &A, const Csr &a, const Csr &b, Csr *c, hipStream_t s);
// exact = 1: each entry's products and sums are carried in double-double and
// rounded once (the exact product of the FP64 inputs, to 1 ulp)
gp_status spgemm_ex(Arena &A, const Csr &a, const Csr &b, Csr *c, hipStream_t s, int exact);
gp_status transpose(Arena &A, const Csr &a, Csr *t, hipStream_t s);
// copy of a single-valued CSR without its explicit zeros
gp_status csr_drop_zeros(Arena &A, const Csr &a, Csr *out, hipStream_t s);
// fill nlong / lrows / lidx (see kLongRow)
gp_status csr_mark_long(Arena &A, Csr &a, hipStream_t s);
// FP32 copy of every value array (long-row lists shared)
gp_status csr_to_float(Arena &A, const Csr &a, CsrF *f, hipStream_t s);
// CSR from COO triplets (duplicates summed), nv value arrays
gp_status csr_from_coo(Arena &A, int64_t nr, int64_t nc, int64_t n, const int64_t *rows, const int64_t *cols,
                       const double *const *vals, int nv, Csr *out, hipStream_t s);
// C = a + alpha * b (same shape, single value array each)
gp_status csr_add(Arena &A, const Csr &a, double alpha, const Csr &b, Csr *c, hipStream_t s);
// Truncated smoothed prolongation: entries of P below theta times the row's
// largest magnitude are dropped, except the row's own aggregate (the column of
// the tentative prolongation Pt, one entry per row); the kept entries are
// scaled so that every row keeps its sum.
gp_status truncate_prolongation(Arena &A, Arena &tmp, const Csr &P, const Csr &Pt, double theta, Csr *out,
                                hipStream_t s);
// host copy of value array `which` of a
gp_status csr_host_copy(const Csr &a, HostCsr *h, int which);

}  // namespace gp

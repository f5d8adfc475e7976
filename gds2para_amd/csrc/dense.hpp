// dense.hpp -- batched dense complex FP64 inverse for the coarsest levels.
#pragma once
#include <cstdint>

#include "gp_dev.hpp"

namespace gp {

template <class T>
struct C2 { T x, y; };

// a: nb row-major n x n complex matrices, overwritten by their inverses.
gp_status dense_inverse_batched(C2<double> *a, int n, int nb, hipStream_t s);

// Matrices up to this many rows are inverted by one workgroup each.
constexpr int kGJWorkgroupMax = 512;

// Gauss-Jordan with partial pivoting, one workgroup per augmented matrix
// [A | I] (row-major n x 2n, device memory): count blocks listed in `list`
// (block b of boff[b+1] - boff[b] rows at aug + aoff[b]; bad[0] set on a
// zero pivot column), or without a list a uniform batch of count nu-row
// matrices (matrix i at aug + 2 i nu^2; bad[i]).  Asynchronous on s.
gp_status gj_workgroup(const int32_t *list, const int32_t *boff, const int64_t *aoff, int nu, int count, C2<double> *aug,
                       int *bad, hipStream_t s);

}  // namespace gp

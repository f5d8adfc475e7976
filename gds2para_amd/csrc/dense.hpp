// dense.hpp -- batched dense complex FP64 inverse for the coarsest levels.
#pragma once
#include "gp_dev.hpp"

namespace gp {

template <class T>
struct C2 { T x, y; };

// a: nb row-major n x n complex matrices, overwritten by their inverses.
gp_status dense_inverse_batched(C2<double> *a, int n, int nb, hipStream_t s);

}  // namespace gp

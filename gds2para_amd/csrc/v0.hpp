// v0.hpp -- device-resident V0 basis (merge_v0d1 / merge_v0c) and its
// Galerkin Laplacians Ad / Ac, in the reference's COO layouts.
#pragma once
#include "gp_dev.hpp"

namespace gp {

struct V0Coo {  // device COO, reference entry order
  int64_t n = 0;
  int64_t *row = nullptr, *col = nullptr;
  double *val = nullptr, *aval = nullptr;
};

struct V0Data {
  V0Coo v0d, v0c;           // row = original edge id
  V0Coo ad, ac;             // row-major sorted
  int64_t v0d_cols = 0, v0c_cols = 0, ndiel = 0;
  int64_t ad_len = 0, ac_len = 0;
  int64_t *mapd = nullptr;  // node -> V0d column + 1 (count pass) or 0
  int64_t *mapc = nullptr;  // node -> V0c column + 1 or 0
  int64_t *dnode_of_col = nullptr;  // V0d dielectric column -> node
  int64_t *cnode_of_col = nullptr;  // V0c column -> node
};

gp_status build_v0(gp_layout *lay);

}  // namespace gp

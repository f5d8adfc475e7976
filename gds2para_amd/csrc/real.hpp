// real.hpp -- device-resident real AMG + FGMRES solve, the engine under
// gp_solve_real (the hypreSolve replacement, src/hypreSolve.cpp:9-283) and
// under the V0 quasi-static path, which reuses one setup for all the solves
// with Ad or Ac (src/matrixCon.cpp:392, 511, 544).
#pragma once
#include "sparse.hpp"

namespace gp {

struct RealAMG;

// Builds the hierarchy of the n x n matrix given as device COO triplets
// (duplicates summed), smoothing with smoother (gp_real_smoother).  The
// solver owns its copy of the matrix.
gp_status real_amg_create(int64_t n, int64_t nnz, const int64_t *rows, const int64_t *cols, const double *vals,
                          int smoother, RealAMG **out, hipStream_t s);
// x = A^-1 b (device vectors of n) by FGMRES(30) preconditioned with one
// V-cycle, from x = 0, until ||b - A x|| <= tol ||b|| or maxit iterations.
// Returns GP_ERR_NOCONV when the tolerance is not met (x holds the last iterate).
gp_status real_amg_solve(RealAMG *G, const double *b, double *x, double tol, int maxit, int32_t *iters,
                         double *relres, hipStream_t s);
int64_t real_amg_rows(const RealAMG *G);
int32_t real_amg_levels(const RealAMG *G);
int64_t real_amg_bytes(const RealAMG *G);
void real_amg_destroy(RealAMG *G);

}  // namespace gp

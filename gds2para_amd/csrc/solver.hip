// solver.hip -- full-wave solve orchestration.
#include "gp_dev.hpp"

namespace gp {
struct Solver {};
void destroy_solver(Solver *s) { delete s; }
}  // namespace gp
using namespace gp;

extern "C" void gp_solve_opts_default(gp_solve_opts *o) {
  o->rtol = 1e-8; o->max_iters = 2000; o->restart = 30; o->precision = GP_PREC_MIXED; o->batch = 0; o->verbose = 0;
}
extern "C" gp_status gp_prepare_solver(gp_layout *, int32_t) { return fail(GP_ERR_STATE, "solver not built yet"); }
extern "C" gp_status gp_solve_fullwave(gp_layout *, const double *, int32_t, const int32_t *, int32_t, const gp_solve_opts *, double *, double *, gp_solve_stats *) { return fail(GP_ERR_STATE, "solver not built yet"); }
extern "C" gp_status gp_solve_real(const int64_t *, const int64_t *, const double *, int64_t, int64_t, const double *, double *, double, int32_t *, double *) { return fail(GP_ERR_STATE, "solver not built yet"); }
extern "C" gp_status gp_bench_spmv(gp_layout *, int32_t, int32_t, int32_t, double *) { return fail(GP_ERR_STATE, "solver not built yet"); }

// v0.hip -- V0 basis (merge_v0d1 / merge_v0c) and nodal Laplacians (generateAd / generateAc).
#include "gp_dev.hpp"

namespace gp {
struct V0Data {};
void destroy_v0(V0Data *v) { delete v; }
}  // namespace gp
using namespace gp;

extern "C" gp_status gp_build_v0(gp_layout *lay) { return fail(GP_ERR_STATE, "V0 stage not built yet"); }
extern "C" gp_status gp_v0_sizes_get(const gp_layout *, gp_v0_sizes *) { return fail(GP_ERR_STATE, "V0 stage not built yet"); }
extern "C" gp_status gp_v0d_copy(const gp_layout *, int64_t *, int64_t *, double *, double *) { return fail(GP_ERR_STATE, "V0 stage not built yet"); }
extern "C" gp_status gp_ad_copy(const gp_layout *, int64_t *, int64_t *, double *) { return fail(GP_ERR_STATE, "V0 stage not built yet"); }
extern "C" gp_status gp_v0c_copy(const gp_layout *, int64_t *, int64_t *, double *, double *) { return fail(GP_ERR_STATE, "V0 stage not built yet"); }
extern "C" gp_status gp_ac_copy(const gp_layout *, int64_t *, int64_t *, double *) { return fail(GP_ERR_STATE, "V0 stage not built yet"); }

// gp_dev.hpp -- device-side internals shared by the HIP translation units:
// error macros, tracked device buffers and the gp_layout handle.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "gds2para.h"
#include "gp_grid.hpp"
#include "gp_status.hpp"

#define GP_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t _e = (call);                                                             \
    if (_e != hipSuccess)                                                               \
      return gp::fail(GP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define GP_TRY(call)                    \
  do {                                  \
    gp_status _s = (call);              \
    if (_s != GP_OK) return _s;         \
  } while (0)

namespace gp {

// Grid-stride launch helper: blocks of 256 threads, enough blocks to cover n.
inline unsigned blocks_for(int64_t n, int threads = 256) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > 2147483647LL) b = 2147483647LL;
  return (unsigned)b;
}

// Device allocation with a byte counter (the layout's footprint).
struct Arena {
  int64_t bytes = 0;
  std::vector<void *> owned;
  template <class T>
  gp_status alloc(T **p, int64_t n) {
    *p = nullptr;
    if (n <= 0) return GP_OK;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * (size_t)n);
    if (e != hipSuccess) {
      *p = nullptr;
      (void)hipGetLastError();
      return fail(GP_ERR_NOMEM, "hipMalloc of " + std::to_string(sizeof(T) * (size_t)n) + " bytes failed");
    }
    bytes += (int64_t)(sizeof(T) * (size_t)n);
    owned.push_back(*p);
    return GP_OK;
  }
  template <class T>
  void release(T *&p, int64_t n) {
    if (!p) return;
    for (auto &q : owned)
      if (q == (void *)p) { q = nullptr; break; }
    (void)hipFree(p);
    bytes -= (int64_t)(sizeof(T) * (size_t)n);
    p = nullptr;
  }
  void free_all() {
    for (void *q : owned)
      if (q) (void)hipFree(q);
    owned.clear();
    bytes = 0;
  }
};

// Deep host copy of a gp_mesh so the caller's arrays may be released.
struct HostMesh {
  gp_mesh view{};
  std::vector<double> xn, yn, zn, epsn, sign, side_area;
  std::vector<int32_t> mark_edge, mark_node, cdt_mark_port, port_side_ptr, side_dir, side_cnd;
  std::vector<int64_t> bd_edges, bd_nodes, side_edge_ptr, side_edges;
  void copy_from(const gp_mesh &m);
};

struct CsrD {            // FP64 CSR, int64 row pointers, int32 columns
  int64_t n = 0, nnz = 0;
  int64_t *rp = nullptr;
  int32_t *ci = nullptr;
  double *v = nullptr;
};

struct V0Data;           // defined in v0.hip
struct V0Path;           // defined in v0path.hip
void destroy_v0path(V0Path *);
int64_t v0path_bytes(const V0Path *);
struct Solver;           // defined in solver.hip
void destroy_v0(V0Data *);
void destroy_solver(Solver *);
int64_t solver_bytes(const Solver *);  // hierarchy, coarse cache and cached slot workspace
struct Hierarchy;
const Hierarchy *solver_hierarchy(const Solver *);

}  // namespace gp

namespace gp {
// Device time of one kernel class, measured with HIP events recorded on the
// launching stream around each launch, together with the algorithmic bytes of
// each launch (for the achieved-bandwidth figure).
struct KTimer {
  bool on = false;
  std::vector<hipEvent_t> ev;     // begin/end pairs
  std::vector<double> bytes;      // per pair
  size_t used = 0;
  void begin(hipStream_t s) {
    if (!on) return;
    if (2 * used + 2 > ev.size()) {
      for (int q = 0; q < 256; ++q) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) { on = false; return; }
        ev.push_back(e);
      }
    }
    (void)hipEventRecord(ev[2 * used], s);
  }
  void end(hipStream_t s, double b) {
    if (!on) return;
    (void)hipEventRecord(ev[2 * used + 1], s);
    if (bytes.size() <= used) bytes.resize(used + 1);
    bytes[used] = b;
    ++used;
  }
  void reset() { used = 0; }
  ~KTimer() { for (hipEvent_t e : ev) (void)hipEventDestroy(e); }
};
}  // namespace gp

struct gp_layout {
  gp::KTimer ktimer;                    // level-0 edge smoother launches
  gp::HostMesh hm;
  gp::Grid g{};
  int64_t n = 0;                        // unknowns after PEC removal
  hipStream_t stream = nullptr;
  int32_t device = 0;                   // the device current when the layout was created
  gp::Arena arena;
  // device copy of the mesh contract
  double *xn = nullptr, *yn = nullptr, *zn = nullptr, *epsn = nullptr, *sign = nullptr;
  int32_t *mark_edge = nullptr, *mark_node = nullptr;
  uint8_t *rm_edge = nullptr, *bd_node = nullptr;
  int64_t *map_edge = nullptr, *map_edge_r = nullptr;
  // stiffness
  bool s_ready = false;
  gp::CsrD S;
  // V0 basis
  gp::V0Data *v0 = nullptr;
  // V0 quasi-static path: V0 operators in CSR and the Ad / Ac hierarchies,
  // built with the smoother v0_smoother names (gp_real_smoother)
  gp::V0Path *v0path = nullptr;
  int32_t v0_smoother = GP_SMOOTH_CHEBYSHEV;
  // solver hierarchy
  gp::Solver *solver = nullptr;
};

namespace gp {
gp_status ensure_device();
// exclusive scan of int64 counts (n entries) into offsets (n+1 entries) on the layout stream
gp_status exclusive_scan_i64(const int64_t *cnt, int64_t *off, int64_t n, hipStream_t s, Arena &tmp);
}  // namespace gp

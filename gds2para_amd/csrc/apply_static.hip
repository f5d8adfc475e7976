// apply_static.hip -- y = A(w) x (gp_apply_fullwave) without any device
// allocation, for callers that cannot allocate: code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc.
//
// gp_apply_fullwave stages the caller's host columns, their interleaved copy,
// the product and the per-column angular frequencies in device buffers it
// allocates per call.  Here those staging buffers live in one file-scope
// device array, `scratch`, of kScratchBytes, which each GPU the library runs
// on holds until the process ends.  Its device address comes from
// hipGetSymbolAddress, once per device with that device current; host code
// uses only that address (the array's name is a host-side stand-in).  The
// kernels are the ones gp_apply_fullwave launches, in the same order, so the
// results are the same bits.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one GPU
// (two threads, or two layouts on one device) race and must be serialised by
// the caller.  Every byte a call reads is written earlier in the same call,
// so nothing has to start at zero.
#include <mutex>

#include "mg.hpp"

namespace gp {

namespace {

// 96 MiB: 256 bytes of per-column frequencies, then three n*k complex FP64
// blocks (host column layout, interleaved input, interleaved product).
constexpr int64_t kScratchBytes = (int64_t)96 << 20;
constexpr int64_t kScratchHead = 256;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, looked up once per device with that
// device current.
gp_status scratch_on(int device, void **out) {
  if (device < 0 || device >= kMaxDevices) return fail(GP_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(GP_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    }
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return GP_OK;
}

// Makes `dev` current for the scope and restores the caller's device.
struct DeviceScope {
  int prev = -1;
  bool switched = false;
  gp_status enter(int dev) {
    GP_HIP(hipGetDevice(&prev));
    if (prev != dev) {
      GP_HIP(hipSetDevice(dev));
      switched = true;
    }
    return GP_OK;
  }
  ~DeviceScope() {
    if (switched) (void)hipSetDevice(prev);
  }
};

}  // namespace

}  // namespace gp

using namespace gp;

extern "C" gp_status gp_apply_fullwave_static(gp_layout *lay, const double *freqs, int32_t k, const double *x,
                                              double *y) {
  if (!lay || !freqs || !x || !y || k < 1 || k > 64)
    return fail(GP_ERR_ARG, "gp_apply_fullwave_static: bad arguments (1 <= k <= 64)");
  for (int32_t j = 0; j < k; ++j)
    if (!(freqs[j] > 0.)) return fail(GP_ERR_ARG, "gp_apply_fullwave_static: frequencies must be positive");
  if (!lay->s_ready || !lay->solver)
    return fail(GP_ERR_STATE, "gp_apply_fullwave_static: prepare the layout first (gp_assemble_stiffness and "
                              "gp_prepare_solver allocate; this entry point does not)");
  const int64_t n = lay->n;
  const int64_t need = kScratchHead + 3 * n * (int64_t)k * (int64_t)sizeof(C2<double>);
  if (need > kScratchBytes)
    return fail(GP_ERR_ARG, "gp_apply_fullwave_static: n*k = " + std::to_string(n * k) + " needs " +
                                std::to_string(need) + " bytes of staging, above the static buffer's " +
                                std::to_string(kScratchBytes) + " (n*k <= " +
                                std::to_string((kScratchBytes - kScratchHead) / (3 * 16)) + ")");
  const Hierarchy *H = solver_hierarchy(lay->solver);
  DeviceScope ds;
  GP_TRY(ds.enter(lay->device));
  void *base = nullptr;
  GP_TRY(scratch_on(lay->device, &base));
  hipStream_t s = lay->stream;
  double *om = (double *)base;
  C2<double> *xc = (C2<double> *)((unsigned char *)base + kScratchHead);
  C2<double> *xr = xc + n * k, *yr = xr + n * k;
  double omh[64];
  for (int32_t j = 0; j < k; ++j) omh[j] = 2 * kPi * freqs[j];
  GP_HIP(hipMemcpyAsync(om, omh, sizeof(double) * k, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(xc, x, sizeof(C2<double>) * n * k, hipMemcpyHostToDevice, s));
  k_cols_to_rows<<<blocks_for(n * k), 256, 0, s>>>(n, k, xc, xr);
  amul(*H, n, k, om, xr, yr, s);
  k_rows_to_cols<<<blocks_for(n * k), 256, 0, s>>>(n, k, yr, xc);
  GP_HIP(hipMemcpyAsync(y, xc, sizeof(C2<double>) * n * k, hipMemcpyDeviceToHost, s));
  GP_HIP(hipStreamSynchronize(s));
  GP_HIP(hipGetLastError());
  return GP_OK;
}

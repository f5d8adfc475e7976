// vh_solve.hip -- paraGenerator's per-frequency Vh solve (src/matrixCon.cpp:608-801):
// the full-wave solution at each frequency as the V0 solution plus a Galerkin
// correction in the span of the Vh basis.
//
// For each frequency w, with D(w) = diag(-w^2 eps_i + i w sigma_i):
//   Vh' = Vh - u0 (u0a^H D u0)^-1 (u0a^H D Vh)          (:611-690, deflation)
//   m_h = Vh'^H (S + D) Vh'                             (:692-724)
//   y_w = Re yd + i Im yd f0 / f                        (:757-770, :784)
//   rhs = Vh'^H (J - D y_w),  J_i = -i w on the port     (:726-776)
//   x   = y_w + Vh' (m_h^-1 rhs)                        (:777-786)
// then Construct_Z_V0_Vh(x) (:789).  The coefficients are the ones this loop
// uses, which differ from the full-wave system's: eps_i from the layer
// (e + N_edge_v) / (N_edge_s + N_edge_v) and sigma_i = SIGMA (5.8e7) on every
// conductor edge (:614-622); J_i = -i w on every edge of the source port
// whatever its direction (:727-731).
//
// The dense products run in rocBLAS (ZGEMM / ZGEMV, conjugate transposes as
// the reference's matrix_multi 'T', src/findVh.cpp:1205-1218), the sparse
// product S Vh' in one kernel, the L x L system in rocSOLVER's ZGESV; the 2 x 2
// deflation system on the host.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <vector>

#include "mg.hpp"

namespace gp {

namespace {

constexpr double kSigmaVh = 5.8e+7;  // the SIGMA macro of the reference (src/fdtd.hpp:50)

// out[c n + i] = d_i in[c n + i], d_i = (-w^2 eps_i, w sigma_i)
__global__ void k_vh_dmul(int64_t n, int cols, const double *__restrict__ eps, const double *__restrict__ sig, double w,
                          const double2 *__restrict__ in, double2 *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n || c >= cols) return;
  const double dr = -(w * w) * eps[i], di = w * sig[i];
  const double2 v = in[(int64_t)c * n + i];
  out[(int64_t)c * n + i] = {dr * v.x - di * v.y, dr * v.y + di * v.x};
}

// T[c n + i] = sum_p S_ip V[c n + ci_p] + d_i V[c n + i]
__global__ void k_vh_sd(int64_t n, int cols, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
                        const double *__restrict__ sv, const double *__restrict__ eps, const double *__restrict__ sig,
                        double w, const double2 *__restrict__ V, double2 *__restrict__ T) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n || c >= cols) return;
  const double2 *vc = V + (int64_t)c * n;
  double re = 0, im = 0;
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
    const double a = sv[p];
    const double2 v = vc[ci[p]];
    re += a * v.x;
    im += a * v.y;
  }
  const double dr = -(w * w) * eps[i], di = w * sig[i];
  const double2 v = vc[i];
  T[(int64_t)c * n + i] = {re + (dr * v.x - di * v.y), im + (dr * v.y + di * v.x)};
}

// y_w = (Re yd, Im yd f0 / f); r = J - D y_w with J_i = (0, -w) on the port
__global__ void k_vh_rhs(int64_t n, const double *__restrict__ eps, const double *__restrict__ sig, double w,
                         double ratio, const uint8_t *__restrict__ jmask, const double2 *__restrict__ yd,
                         double2 *__restrict__ yw, double2 *__restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 y = {yd[i].x, yd[i].y * ratio};
  const double dr = -(w * w) * eps[i], di = w * sig[i];
  const double jr = 0.0, ji = jmask[i] ? -w : 0.0;
  yw[i] = y;
  r[i] = {jr - (dr * y.x - di * y.y), ji - (dr * y.y + di * y.x)};
}

// complex interleaved host copy of a real n x cols block
void real_to_complex(const double *in, int64_t len, std::vector<double> &out) {
  out.assign(2 * (size_t)len, 0.0);
  for (int64_t i = 0; i < len; ++i) out[2 * i] = in[i];
}

#define RB_TRY(call)                                                                       \
  do {                                                                                     \
    const rocblas_status _r = (call);                                                      \
    if (_r != rocblas_status_success)                                                      \
      return fail(GP_ERR_HIP, std::string(#call) + " failed: " + rocblas_status_to_string(_r)); \
  } while (0)

}  // namespace

}  // namespace gp

using namespace gp;

extern "C" gp_status gp_solve_vh(gp_layout *lay, const double *vh, int64_t leng_vh, const double *u0,
                                 const double *u0a, const double *yd, int32_t source_port, const double *freqs,
                                 int32_t nf, double *X, double *Z) {
  if (!lay || !vh || leng_vh < 1 || !u0 || !u0a || !yd || !freqs || nf < 1 || !Z)
    return fail(GP_ERR_ARG, "gp_solve_vh: bad arguments");
  const gp_mesh &m = lay->hm.view;
  if (source_port < 0 || source_port >= m.num_ports) return fail(GP_ERR_ARG, "gp_solve_vh: source port out of range");
  for (int32_t k = 0; k < nf; ++k)
    if (!(freqs[k] > 0)) return fail(GP_ERR_ARG, "gp_solve_vh: frequencies must be positive");
  if (leng_vh > 65535 || lay->n > 0x7fffffff) return fail(GP_ERR_ARG, "gp_solve_vh: sizes beyond 32-bit BLAS");
  if (!lay->s_ready) GP_TRY(gp_assemble_stiffness(lay));
  const int64_t n = lay->n, L = leng_vh, ne = lay->g.Ne;
  const Grid &g = lay->g;
  hipStream_t s = lay->stream;
  // per-row coefficients of this loop, the port mask and yd in row order (host)
  std::vector<int64_t> map(ne), mr(n);
  GP_TRY(gp_edge_map(&m, map.data(), mr.data()));
  std::vector<double> eps(n), sig(n), ydr(2 * n);
  std::vector<uint8_t> jm(n, 0);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t e = mr[i];
    eps[i] = m.stack_epsn[(e + g.Ev) / (g.Es + g.Ev)] * kEps0;
    sig[i] = m.mark_edge[e] != 0 ? kSigmaVh : 0.0;
    ydr[2 * i] = yd[2 * e];
    ydr[2 * i + 1] = yd[2 * e + 1];
  }
  for (int32_t sd = m.port_side_ptr[source_port]; sd < m.port_side_ptr[source_port + 1]; ++sd)
    for (int64_t q = m.side_edge_ptr[sd]; q < m.side_edge_ptr[sd + 1]; ++q) {
      const int64_t row = map[m.side_edges[q]];
      if (row >= 0 && m.side_dir[sd] != 0) jm[row] = 1;
    }
  std::vector<double> uc, uac;
  real_to_complex(u0, 2 * n, uc);
  real_to_complex(u0a, 2 * n, uac);

  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) return fail(GP_ERR_HIP, "rocblas_create_handle failed");
  struct HandleGuard { rocblas_handle h; ~HandleGuard() { rocblas_destroy_handle(h); } } hg{h};
  RB_TRY(rocblas_set_stream(h, s));
  Arena T;
  struct Free { Arena &a; hipStream_t s; ~Free() { (void)hipStreamSynchronize(s); a.free_all(); } } fr{T, s};
  double2 *dVh, *dVp, *dT, *dU, *dUa, *dDU, *dY, *dM, *dYd, *dYw, *dR, *dRh, *d4;
  double *dEps, *dSig;
  uint8_t *dJ;
  rocblas_int *ipiv, *info;
  GP_TRY(T.alloc(&dVh, n * L)); GP_TRY(T.alloc(&dVp, n * L)); GP_TRY(T.alloc(&dT, n * L));
  GP_TRY(T.alloc(&dU, 2 * n)); GP_TRY(T.alloc(&dUa, 2 * n)); GP_TRY(T.alloc(&dDU, 2 * n));
  GP_TRY(T.alloc(&dY, 2 * L)); GP_TRY(T.alloc(&dM, L * L)); GP_TRY(T.alloc(&d4, 4));
  GP_TRY(T.alloc(&dYd, n)); GP_TRY(T.alloc(&dYw, n)); GP_TRY(T.alloc(&dR, n)); GP_TRY(T.alloc(&dRh, L));
  GP_TRY(T.alloc(&dEps, n)); GP_TRY(T.alloc(&dSig, n)); GP_TRY(T.alloc(&dJ, n));
  GP_TRY(T.alloc(&ipiv, L)); GP_TRY(T.alloc(&info, 1));
  GP_HIP(hipMemcpyAsync(dVh, vh, sizeof(double2) * n * L, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(dU, uc.data(), sizeof(double2) * 2 * n, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(dUa, uac.data(), sizeof(double2) * 2 * n, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(dYd, ydr.data(), sizeof(double2) * n, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(dEps, eps.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(dSig, sig.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  GP_HIP(hipMemcpyAsync(dJ, jm.data(), n, hipMemcpyHostToDevice, s));
  auto cz = [](double2 *p) { return reinterpret_cast<rocblas_double_complex *>(p); };
  const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0}, mone = {-1.0, 0.0};
  const rocblas_operation C = rocblas_operation_conjugate_transpose, N = rocblas_operation_none;
  const int in = (int)n, iL = (int)L;
  const double f0 = m.freq_start * m.freq_unit;
  const int np = m.num_ports;
  std::vector<double2> yre(2 * L), t4(4), x(n);
  std::vector<double> Zk(2 * (size_t)np * np);
  const unsigned gn = blocks_for(n);
  for (int32_t k = 0; k < nf; ++k) {
    const double f = freqs[k], w = 2 * kPi * f;
    // deflation: y_re = u0a^H D Vh (2 x L), t4 = u0a^H D u0 (2 x 2); Y = t4^-1 y_re
    k_vh_dmul<<<dim3(gn, (unsigned)L), 256, 0, s>>>(n, iL, dEps, dSig, w, dVh, dT);
    k_vh_dmul<<<dim3(gn, 2), 256, 0, s>>>(n, 2, dEps, dSig, w, dU, dDU);
    RB_TRY(rocblas_zgemm(h, C, N, 2, iL, in, &one, cz(dUa), in, cz(dT), in, &zero, cz(dY), 2));
    RB_TRY(rocblas_zgemm(h, C, N, 2, 2, in, &one, cz(dUa), in, cz(dDU), in, &zero, cz(d4), 2));
    GP_HIP(hipMemcpyAsync(yre.data(), dY, sizeof(double2) * 2 * L, hipMemcpyDeviceToHost, s));
    GP_HIP(hipMemcpyAsync(t4.data(), d4, sizeof(double2) * 4, hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    {
      // 2 x 2 complex solve by elimination with partial pivoting (column-major t4)
      typedef std::complex<double> cd;
      cd a00(t4[0].x, t4[0].y), a10(t4[1].x, t4[1].y), a01(t4[2].x, t4[2].y), a11(t4[3].x, t4[3].y);
      const bool sw = std::abs(a10) > std::abs(a00);
      if (sw) { std::swap(a00, a10); std::swap(a01, a11); }
      if (a00 == cd(0.0)) return fail(GP_ERR_STATE, "gp_solve_vh: singular deflation system u0a^H D u0");
      const cd l = a10 / a00, u11 = a11 - l * a01;
      if (u11 == cd(0.0)) return fail(GP_ERR_STATE, "gp_solve_vh: singular deflation system u0a^H D u0");
      for (int64_t c = 0; c < L; ++c) {
        cd b0(yre[2 * c].x, yre[2 * c].y), b1(yre[2 * c + 1].x, yre[2 * c + 1].y);
        if (sw) std::swap(b0, b1);
        b1 -= l * b0;
        const cd x1 = b1 / u11, x0 = (b0 - a01 * x1) / a00;
        yre[2 * c] = {x0.real(), x0.imag()};
        yre[2 * c + 1] = {x1.real(), x1.imag()};
      }
    }
    GP_HIP(hipMemcpyAsync(dY, yre.data(), sizeof(double2) * 2 * L, hipMemcpyHostToDevice, s));
    GP_HIP(hipMemcpyAsync(dVp, dVh, sizeof(double2) * n * L, hipMemcpyDeviceToDevice, s));
    RB_TRY(rocblas_zgemm(h, N, N, in, iL, 2, &mone, cz(dU), in, cz(dY), 2, &one, cz(dVp), in));
    // m_h = Vh'^H (S + D) Vh'
    k_vh_sd<<<dim3(gn, (unsigned)L), 256, 0, s>>>(n, iL, lay->S.rp, lay->S.ci, lay->S.v, dEps, dSig, w, dVp, dT);
    RB_TRY(rocblas_zgemm(h, C, N, iL, iL, in, &one, cz(dVp), in, cz(dT), in, &zero, cz(dM), iL));
    // rhs = Vh'^H (J - D y_w); solve; x = y_w + Vh' y
    k_vh_rhs<<<gn, 256, 0, s>>>(n, dEps, dSig, w, f0 / f, dJ, dYd, dYw, dR);
    RB_TRY(rocblas_zgemv(h, C, in, iL, &one, cz(dVp), in, cz(dR), 1, &zero, cz(dRh), 1));
    RB_TRY(rocsolver_zgesv(h, iL, 1, cz(dM), iL, ipiv, cz(dRh), iL, info));
    RB_TRY(rocblas_zgemv(h, N, in, iL, &one, cz(dVp), in, cz(dRh), 1, &one, cz(dYw), 1));
    rocblas_int hinfo = 0;
    GP_HIP(hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost, s));
    GP_HIP(hipMemcpyAsync(x.data(), dYw, sizeof(double2) * n, hipMemcpyDeviceToHost, s));
    GP_HIP(hipStreamSynchronize(s));
    GP_HIP(hipGetLastError());
    if (hinfo != 0) return fail(GP_ERR_STATE, "gp_solve_vh: singular reduced system Vh'^H (S + D) Vh' at " +
                                              std::to_string(f) + " Hz");
    if (X) std::copy((const double *)x.data(), (const double *)x.data() + 2 * n, X + 2 * (size_t)k * n);
    // Construct_Z_V0_Vh(final_x, k, sourcePort)
    double *zk = Z + 2 * (size_t)k * np * np;
    GP_TRY(gp_extract_z(&m, (const double *)x.data(), source_port, zk));
  }
  return GP_OK;
}

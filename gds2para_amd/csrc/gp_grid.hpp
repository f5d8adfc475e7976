// gp_grid.hpp -- structured Yee-grid index arithmetic shared by host C++ and
// HIP device code.  The numbering is the reference's grow-along-z ordering
// (src/fdtd.hpp:4107-4130 compute_edgelink, src/mesh.cpp:395-410 sizes,
// src/generateStiff.cpp:36-245 face/edge indexing).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define GP_HD __host__ __device__ __forceinline__
#else
#define GP_HD inline
#endif

namespace gp {

// Physical constants exactly as the reference's macros evaluate them
// (src/fdtd.hpp:45-47): MU = (4*M_PI*1.e-7), EPSILON0 = 1./(CSPED*CSPED*MU).
constexpr double kPi = 3.14159265358979323846;
constexpr double kMu = (4 * kPi * 1.e-7);
constexpr double kCsped = 299792458.;
constexpr double kEps0 = (1. / (kCsped * kCsped * kMu));

struct Grid {
  int64_t Nx, Ny, Nz;   // cells
  int64_t Es, Ev, El;   // surface edges / plane, vertical edges / layer, Es+Ev
  int64_t Ne, Nn, Nns;  // edges, nodes, nodes per plane

  GP_HD static Grid make(int64_t nx_nodes, int64_t ny_nodes, int64_t nz_nodes) {
    Grid g;
    g.Nx = nx_nodes - 1; g.Ny = ny_nodes - 1; g.Nz = nz_nodes - 1;
    g.Es = g.Ny * (g.Nx + 1) + g.Nx * (g.Ny + 1);
    g.Ev = (g.Nx + 1) * (g.Ny + 1);
    g.El = g.Es + g.Ev;
    g.Ne = g.Es * (g.Nz + 1) + g.Ev * g.Nz;
    g.Nns = g.Ev;
    g.Nn = (g.Nz + 1) * g.Nns;
    return g;
  }
  GP_HD int64_t npatch() const { return Nx * Ny * (Nz + 1) + ((Nx + 1) * Ny + (Ny + 1) * Nx) * Nz; }
  // edge ids
  GP_HD int64_t ey(int64_t iz, int64_t ix, int64_t iy) const { return iz * El + ix * Ny + iy; }
  GP_HD int64_t ex(int64_t iz, int64_t ix, int64_t iy) const { return iz * El + Ny * (Nx + 1) + ix * (Ny + 1) + iy; }
  GP_HD int64_t ez(int64_t iz, int64_t ix, int64_t iy) const { return iz * El + Es + ix * (Ny + 1) + iy; }
  GP_HD int64_t node(int64_t iz, int64_t ix, int64_t iy) const { return iz * Nns + ix * (Ny + 1) + iy; }

  // Decode an edge id: dir 0 = y, 1 = x, 2 = z (matches compute_edgelink's branches).
  GP_HD void decode(int64_t e, int &dir, int64_t &iz, int64_t &ix, int64_t &iy) const {
    iz = e / El;
    int64_t r = e - iz * El;
    if (r >= Es) { dir = 2; r -= Es; ix = r / (Ny + 1); iy = r - ix * (Ny + 1); }
    else if (r >= Ny * (Nx + 1)) { dir = 1; r -= Ny * (Nx + 1); ix = r / (Ny + 1); iy = r - ix * (Ny + 1); }
    else { dir = 0; ix = r / Ny; iy = r - ix * Ny; }
  }
  // The two end nodes of an edge (compute_edgelink, src/fdtd.hpp:4107).
  GP_HD void ends(int64_t e, int64_t &n1, int64_t &n2) const {
    int d; int64_t iz, ix, iy;
    decode(e, d, iz, ix, iy);
    n1 = node(iz, ix, iy);
    n2 = (d == 2) ? node(iz + 1, ix, iy) : (d == 1) ? node(iz, ix + 1, iy) : node(iz, ix, iy + 1);
  }

  // Face ids in the reference's Se/Sh column (emission) order
  // (src/generateStiff.cpp:33-245): cells below the top layer emit
  // [xy-bottom, yz, xz]; the top cell layer emits [xy-bottom, yz, xz, xy-top];
  // then the x = Nx yz-plane, then the y = Ny xz-plane.
  GP_HD int64_t f_xy(int64_t iz, int64_t ix, int64_t iy) const {  // xy face at plane iz, cell (ix,iy)
    if (iz < Nz - 1) return iz * Nx * Ny * 3 + (ix * Ny + iy) * 3;
    if (iz == Nz - 1) return (Nz - 1) * Nx * Ny * 3 + (ix * Ny + iy) * 4;
    return (Nz - 1) * Nx * Ny * 3 + (ix * Ny + iy) * 4 + 3;  // iz == Nz (top plane)
  }
  GP_HD int64_t f_yz(int64_t iz, int64_t ix, int64_t iy) const {  // yz face at x = ix, cells (iy, iz)
    if (ix < Nx) {
      if (iz < Nz - 1) return iz * Nx * Ny * 3 + (ix * Ny + iy) * 3 + 1;
      return (Nz - 1) * Nx * Ny * 3 + (ix * Ny + iy) * 4 + 1;
    }
    return (Nz - 1) * Nx * Ny * 3 + 4 * Nx * Ny + iz * Ny + iy;
  }
  GP_HD int64_t f_xz(int64_t iz, int64_t ix, int64_t iy) const {  // xz face at y = iy, cells (ix, iz)
    if (iy < Ny) {
      if (iz < Nz - 1) return iz * Nx * Ny * 3 + (ix * Ny + iy) * 3 + 2;
      return (Nz - 1) * Nx * Ny * 3 + (ix * Ny + iy) * 4 + 2;
    }
    return (Nz - 1) * Nx * Ny * 3 + 4 * Nx * Ny + Nz * Ny + iz * Nx + ix;
  }
};

// avg_length (src/fdtd.hpp:4074): node-centred dual lengths, full first/last cell.
GP_HD double avg_len(const double *c, int64_t i, int64_t n) {
  if (i == 0) return c[1] - c[0];
  if (i == n - 1) return c[i] - c[i - 1];
  return (c[i + 1] - c[i - 1]) / 2;
}

}  // namespace gp

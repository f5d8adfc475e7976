// host_mesh.cpp -- host-side mesh services of the C ABI: closed-form sizes,
// validation, the PEC edge map, the frequency grid, port-Z extraction and the
// seeded synthetic layout generator that stands in for the GDSII front end.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <queue>
#include <random>
#include <string>
#include <vector>

#include "gds2para.h"
#include "gp_grid.hpp"
#include "gp_status.hpp"

using gp::Grid;

extern "C" gp_status gp_mesh_sizes_of(const gp_mesh *m, gp_mesh_sizes *out) {
  if (!m || !out || m->nx < 2 || m->ny < 2 || m->nz < 2) return gp::fail(GP_ERR_ARG, "mesh needs >= 2 nodes per axis");
  Grid g = Grid::make(m->nx, m->ny, m->nz);
  out->n_edge = g.Ne;
  out->n_edge_s = g.Es;
  out->n_edge_v = g.Ev;
  out->n_node = g.Nn;
  out->n_node_s = g.Nns;
  out->n_patch = g.npatch();
  out->n_unknowns = g.Ne - m->num_bd_edges;
  out->num_sides = m->port_side_ptr ? m->port_side_ptr[m->num_ports] : 0;
  // nnz of S: every kept edge couples to itself and to the kept edges of each
  // face it bounds (two distinct edges share at most one face).
  std::vector<uint8_t> rm(g.Ne, 0);
  for (int64_t i = 0; i < m->num_bd_edges; ++i) rm[m->bd_edges[i]] = 1;
  int64_t nnz = 0;
  for (int64_t e = 0; e < g.Ne; ++e) {
    if (rm[e]) continue;
    int d; int64_t iz, ix, iy;
    g.decode(e, d, iz, ix, iy);
    int64_t cnt = 1;
    auto face = [&](int64_t a, int64_t b, int64_t c) { cnt += !rm[a] + !rm[b] + !rm[c]; };
    if (d == 0) {  // y-edge: xy faces at cells (ix-1,iy),(ix,iy); yz faces at z cells iz-1, iz
      if (ix > 0) face(g.ey(iz, ix - 1, iy), g.ex(iz, ix - 1, iy), g.ex(iz, ix - 1, iy + 1));
      if (ix < g.Nx) face(g.ey(iz, ix + 1, iy), g.ex(iz, ix, iy), g.ex(iz, ix, iy + 1));
      if (iz > 0) face(g.ey(iz - 1, ix, iy), g.ez(iz - 1, ix, iy), g.ez(iz - 1, ix, iy + 1));
      if (iz < g.Nz) face(g.ey(iz + 1, ix, iy), g.ez(iz, ix, iy), g.ez(iz, ix, iy + 1));
    } else if (d == 1) {
      if (iy > 0) face(g.ex(iz, ix, iy - 1), g.ey(iz, ix, iy - 1), g.ey(iz, ix + 1, iy - 1));
      if (iy < g.Ny) face(g.ex(iz, ix, iy + 1), g.ey(iz, ix, iy), g.ey(iz, ix + 1, iy));
      if (iz > 0) face(g.ex(iz - 1, ix, iy), g.ez(iz - 1, ix, iy), g.ez(iz - 1, ix + 1, iy));
      if (iz < g.Nz) face(g.ex(iz + 1, ix, iy), g.ez(iz, ix, iy), g.ez(iz, ix + 1, iy));
    } else {
      if (iy > 0) face(g.ez(iz, ix, iy - 1), g.ey(iz, ix, iy - 1), g.ey(iz + 1, ix, iy - 1));
      if (iy < g.Ny) face(g.ez(iz, ix, iy + 1), g.ey(iz, ix, iy), g.ey(iz + 1, ix, iy));
      if (ix > 0) face(g.ez(iz, ix - 1, iy), g.ex(iz, ix - 1, iy), g.ex(iz + 1, ix - 1, iy));
      if (ix < g.Nx) face(g.ez(iz, ix + 1, iy), g.ex(iz, ix, iy), g.ex(iz + 1, ix, iy));
    }
    nnz += cnt;
  }
  out->nnz_s = nnz;
  return GP_OK;
}

extern "C" gp_status gp_mesh_validate(const gp_mesh *m) {
  if (!m) return gp::fail(GP_ERR_ARG, "null mesh");
  if (m->nx < 2 || m->ny < 2 || m->nz < 2) return gp::fail(GP_ERR_ARG, "mesh needs >= 2 nodes per axis");
  if (!m->xn || !m->yn || !m->zn || !m->stack_epsn || !m->stack_sign || !m->mark_edge || !m->mark_node)
    return gp::fail(GP_ERR_ARG, "mesh array missing");
  const double *c[3] = {m->xn, m->yn, m->zn};
  const int64_t n[3] = {m->nx, m->ny, m->nz};
  for (int a = 0; a < 3; ++a)
    for (int64_t i = 1; i < n[a]; ++i)
      if (!(c[a][i] > c[a][i - 1])) return gp::fail(GP_ERR_ARG, "coordinates must increase strictly");
  Grid g = Grid::make(m->nx, m->ny, m->nz);
  for (int64_t i = 0; i < m->num_bd_edges; ++i) {
    int64_t e = m->bd_edges[i];
    if (e < 0 || e >= g.Ne || (i && e <= m->bd_edges[i - 1])) return gp::fail(GP_ERR_ARG, "bd_edges must be sorted, unique, in range");
    // setMapEdge only removes plane-0 and top-plane surface edges (src/fdtd.hpp:1508-1529)
    bool low = e < g.Es, top = e >= g.Ne - g.Es;
    if (!low && !top) return gp::fail(GP_ERR_ARG, "boundary edge not on plane 0 or the top plane");
  }
  for (int64_t i = 0; i < m->num_bd_nodes; ++i) {
    int64_t v = m->bd_nodes[i];
    if (v < 0 || v >= g.Nn || (i && v <= m->bd_nodes[i - 1])) return gp::fail(GP_ERR_ARG, "bd_nodes must be sorted, unique, in range");
  }
  for (int32_t c2 = 0; c2 < m->num_cdt; ++c2) {
    int32_t mp = m->cdt_mark_port[c2];
    if (mp < -1 || mp > m->num_ports) return gp::fail(GP_ERR_ARG, "cdt_mark_port out of range");
  }
  for (int64_t v = 0; v < g.Nn; ++v)
    if (m->mark_node[v] < 0 || m->mark_node[v] > m->num_cdt) return gp::fail(GP_ERR_ARG, "mark_node out of range");
  for (int64_t e = 0; e < g.Ne; ++e)
    if (m->mark_edge[e] < 0) return gp::fail(GP_ERR_ARG, "mark_edge negative");
  if (m->num_ports < 0 || (m->num_ports > 0 && !m->port_side_ptr)) return gp::fail(GP_ERR_ARG, "ports malformed");
  if (m->num_ports > 0) {
    if (m->port_side_ptr[0] != 0) return gp::fail(GP_ERR_ARG, "port_side_ptr[0] != 0");
    for (int32_t p = 0; p < m->num_ports; ++p)
      if (m->port_side_ptr[p + 1] <= m->port_side_ptr[p]) return gp::fail(GP_ERR_ARG, "port with no side");
    int32_t ns = m->port_side_ptr[m->num_ports];
    if (m->side_edge_ptr[0] != 0) return gp::fail(GP_ERR_ARG, "side_edge_ptr[0] != 0");
    for (int32_t s = 0; s < ns; ++s) {
      if (m->side_edge_ptr[s + 1] < m->side_edge_ptr[s]) return gp::fail(GP_ERR_ARG, "side_edge_ptr decreasing");
      if (!(m->side_area[s] > 0)) return gp::fail(GP_ERR_ARG, "port side area must be positive");
    }
    std::vector<uint8_t> rm(g.Ne, 0);
    for (int64_t i = 0; i < m->num_bd_edges; ++i) rm[m->bd_edges[i]] = 1;
    for (int64_t k = 0; k < m->side_edge_ptr[ns]; ++k) {
      int64_t e = m->side_edges[k];
      if (e < 0 || e >= g.Ne || rm[e]) return gp::fail(GP_ERR_ARG, "port edge out of range or on a PEC boundary");
    }
  }
  return GP_OK;
}

extern "C" gp_status gp_edge_map(const gp_mesh *m, int64_t *map_edge, int64_t *map_edge_r) {
  if (!m || !map_edge) return gp::fail(GP_ERR_ARG, "null argument");
  Grid g = Grid::make(m->nx, m->ny, m->nz);
  int64_t k = 0, count = 0;
  for (int64_t e = 0; e < g.Ne; ++e) {
    if (k < m->num_bd_edges && m->bd_edges[k] == e) {
      map_edge[e] = -1;
      ++count;
      ++k;
    } else {
      map_edge[e] = e - count;
      if (map_edge_r) map_edge_r[e - count] = e;
    }
  }
  return GP_OK;
}

extern "C" double gp_freq_of(const gp_mesh *m, int32_t k) {
  // freqNo2freq (src/fdtd.hpp:4319-4334)
  if (m->nfreq == 1) return m->freq_start * m->freq_unit;
  if (m->freq_scale == 1)
    return (m->freq_start + k * (m->freq_end - m->freq_start) / (m->nfreq - 1)) * m->freq_unit;
  return m->freq_start * m->freq_unit * std::pow(m->freq_end / m->freq_start, (k * 1.0 / (m->nfreq - 1)));
}

extern "C" gp_status gp_extract_z(const gp_mesh *m, const double *x, int32_t src, double *Zk) {
  // Construct_Z_V0_Vh (src/fdtd.hpp:4133-4170): response line integral over the
  // first side of each port, divided by the summed source-side areas.
  if (!m || !x || !Zk || src < 0 || src >= m->num_ports) return gp::fail(GP_ERR_ARG, "bad argument");
  Grid g = Grid::make(m->nx, m->ny, m->nz);
  std::vector<int64_t> map(g.Ne);
  gp_edge_map(m, map.data(), nullptr);
  const int32_t np = m->num_ports;
  double cur = 0.;
  for (int32_t s = m->port_side_ptr[src]; s < m->port_side_ptr[src + 1]; ++s) cur += m->side_area[s];
  for (int32_t p = 0; p < np; ++p) {
    std::complex<double> z(Zk[2 * (p + np * src)], Zk[2 * (p + np * src) + 1]);
    int32_t s0 = m->port_side_ptr[p];
    for (int64_t k = m->side_edge_ptr[s0]; k < m->side_edge_ptr[s0 + 1]; ++k) {
      int64_t e = m->side_edges[k];
      int d; int64_t iz, ix, iy;
      g.decode(e, d, iz, ix, iy);
      double leng = d == 2 ? m->zn[iz + 1] - m->zn[iz] : d == 1 ? m->xn[ix + 1] - m->xn[ix] : m->yn[iy + 1] - m->yn[iy];
      std::complex<double> xv(x[2 * map[e]], x[2 * map[e] + 1]);
      z -= xv * leng * (m->side_dir[s0] * 1.0);
    }
    z /= cur;
    Zk[2 * (p + np * src)] = z.real();
    Zk[2 * (p + np * src) + 1] = z.imag();
  }
  return GP_OK;
}

// ---------------------------------------------------------------------------
// synthetic layouts
// ---------------------------------------------------------------------------
struct gp_owned_mesh {
  gp_mesh view;
  std::vector<double> xn, yn, zn, epsn, sign, side_area;
  std::vector<int32_t> mark_edge, mark_node, cdt_mark_port, port_side_ptr, side_dir, side_cnd;
  std::vector<int64_t> bd_edges, bd_nodes, side_edge_ptr, side_edges;
};

namespace {

struct Box { int64_t x0, x1, y0, y1, z0, z1; int32_t layer; };  // inclusive node ranges

struct Rng {
  std::mt19937_64 r;
  explicit Rng(uint64_t s) : r(s) {}
  int64_t uni(int64_t lo, int64_t hi) { return lo + (int64_t)(r() % (uint64_t)(hi - lo + 1)); }
  double unif() { return (double)(r() >> 11) * (1.0 / 9007199254740992.0); }
};

std::vector<double> axis(int64_t ncell, double h, double grade) {
  std::vector<double> c(ncell + 1);
  c[0] = 0.;
  for (int64_t i = 0; i < ncell; ++i) {
    double w = h * (1. + grade * std::sin(2. * gp::kPi * (double)i / 16.));
    c[i + 1] = c[i] + w;
  }
  return c;
}

// Marks nodes/edges covered by boxes, labels isolated conductors by BFS in
// ascending node order (src/mesh.cpp:640-800) and recolours conductor edges
// with the conductor id (src/mesh.cpp:1004-1015).  Plane 0 is a PEC ground
// (LOWER_BOUNDARY_PEC, src/mesh.cpp:474-483).
void mark_conductors(gp_owned_mesh &M, const Grid &g, const std::vector<Box> &boxes) {
  std::vector<int32_t> raw_node(g.Nn, 0), raw_edge(g.Ne, 0);
  for (int64_t e = 0; e < g.Es; ++e) raw_edge[e] = 1;
  for (int64_t v = 0; v < g.Nns; ++v) raw_node[v] = 1;
  for (const Box &b : boxes) {
    for (int64_t iz = b.z0; iz <= b.z1; ++iz)
      for (int64_t ix = b.x0; ix <= b.x1; ++ix)
        for (int64_t iy = b.y0; iy <= b.y1; ++iy) {
          raw_node[g.node(iz, ix, iy)] = 1;
          if (ix < b.x1) raw_edge[g.ex(iz, ix, iy)] = 1;
          if (iy < b.y1) raw_edge[g.ey(iz, ix, iy)] = 1;
          if (iz < b.z1) raw_edge[g.ez(iz, ix, iy)] = 1;
        }
  }
  std::vector<int32_t> vis(g.Nn, 0);
  int32_t count = 0;
  std::vector<int64_t> qu;
  for (int64_t s = 0; s < g.Nn; ++s) {
    if (!raw_node[s] || vis[s]) continue;
    vis[s] = ++count;
    qu.clear();
    qu.push_back(s);
    for (size_t h = 0; h < qu.size(); ++h) {
      int64_t v = qu[h];
      int64_t iz = v / g.Nns, ix = (v % g.Nns) / (g.Ny + 1), iy = (v % g.Nns) % (g.Ny + 1);
      int64_t ed[6] = {-1, -1, -1, -1, -1, -1};
      if (iz > 0) ed[0] = g.ez(iz - 1, ix, iy);
      if (iz < g.Nz) ed[1] = g.ez(iz, ix, iy);
      if (ix > 0) ed[2] = g.ex(iz, ix - 1, iy);
      if (ix < g.Nx) ed[3] = g.ex(iz, ix, iy);
      if (iy > 0) ed[4] = g.ey(iz, ix, iy - 1);
      if (iy < g.Ny) ed[5] = g.ey(iz, ix, iy);
      for (int k = 0; k < 6; ++k) {
        if (ed[k] < 0 || !raw_edge[ed[k]]) continue;
        int64_t a, b;
        g.ends(ed[k], a, b);
        int64_t o = a == v ? b : a;
        if (!vis[o]) { vis[o] = count; qu.push_back(o); }
      }
    }
  }
  M.mark_node = vis;
  M.mark_edge.assign(g.Ne, 0);
  for (int64_t e = 0; e < g.Ne; ++e) {
    if (!raw_edge[e]) continue;
    int64_t a, b;
    g.ends(e, a, b);
    M.mark_edge[e] = (vis[a] == vis[b] && vis[a] != 0) ? vis[a] : raw_edge[e];
  }
  M.cdt_mark_port.assign(count, 0);
}

// Port sides as portSet builds them for a vertical side (src/mesh.cpp:1318-1360):
// edges from the return plane up to the supply plane, area from half-cells.
void add_z_port(gp_owned_mesh &M, const Grid &g, int64_t ix, int64_t iy, int64_t z_ret, int64_t z_sup, int32_t dir) {
  for (int64_t iz = z_ret; iz < z_sup; ++iz) M.side_edges.push_back(g.ez(iz, ix, iy));
  M.side_edge_ptr.push_back((int64_t)M.side_edges.size());
  const double *xn = M.xn.data(), *yn = M.yn.data();
  double x1 = xn[ix], y1 = yn[iy];
  double area = 1.;
  if (ix == 0) area *= (xn[ix + 1] - x1);
  else if (ix == g.Nx) area *= (x1 - xn[ix - 1]);
  else area *= ((xn[ix + 1] - x1) / 2 + (x1 - xn[ix - 1]) / 2);
  if (iy == 0) area *= (yn[iy + 1] - y1);
  else if (iy == g.Ny) area *= (y1 - yn[iy - 1]);
  else area *= ((yn[iy + 1] - y1) / 2 + (y1 - yn[iy - 1]) / 2);
  M.side_area.push_back(area);
  M.side_dir.push_back(dir);
  M.side_cnd.push_back(M.mark_node[g.node(z_sup, ix, iy)]);
}

// PEC boundary sets: findBoundNodeEdge over the reference conductors
// (src/fdtd.hpp:1444-1505) plus the whole lower plane (src/mesh.cpp:963-974).
void boundary_sets(gp_owned_mesh &M, const Grid &g) {
  std::vector<uint8_t> be(g.Ne, 0), bn(g.Nn, 0);
  for (int64_t e = 0; e < g.Es; ++e) be[e] = 1;
  for (int64_t v = 0; v < g.Nns; ++v) bn[v] = 1;
  int64_t ztop = g.Nz;
  for (int64_t ix = 0; ix <= g.Nx; ++ix)
    for (int64_t iy = 0; iy <= g.Ny; ++iy) {
      int64_t v = g.node(ztop, ix, iy);
      int32_t c = M.mark_node[v];
      if (!c || M.cdt_mark_port[c - 1] != -1) continue;
      bn[v] = 1;
      if (ix > 0 && M.mark_edge[g.ex(ztop, ix - 1, iy)]) be[g.ex(ztop, ix - 1, iy)] = 1;
      if (ix < g.Nx && M.mark_edge[g.ex(ztop, ix, iy)]) be[g.ex(ztop, ix, iy)] = 1;
      if (iy > 0 && M.mark_edge[g.ey(ztop, ix, iy - 1)]) be[g.ey(ztop, ix, iy - 1)] = 1;
      if (iy < g.Ny && M.mark_edge[g.ey(ztop, ix, iy)]) be[g.ey(ztop, ix, iy)] = 1;
    }
  for (int64_t e = 0; e < g.Ne; ++e) if (be[e]) M.bd_edges.push_back(e);
  for (int64_t v = 0; v < g.Nn; ++v) if (bn[v]) M.bd_nodes.push_back(v);
}

void finish_view(gp_owned_mesh &M, const gp_gen_params &p) {
  gp_mesh &v = M.view;
  std::memset(&v, 0, sizeof(v));
  v.nx = (int64_t)M.xn.size(); v.ny = (int64_t)M.yn.size(); v.nz = (int64_t)M.zn.size();
  v.xn = M.xn.data(); v.yn = M.yn.data(); v.zn = M.zn.data();
  v.stack_epsn = M.epsn.data(); v.stack_sign = M.sign.data();
  v.mark_edge = M.mark_edge.data(); v.mark_node = M.mark_node.data();
  v.num_cdt = (int32_t)M.cdt_mark_port.size(); v.cdt_mark_port = M.cdt_mark_port.data();
  v.num_bd_edges = (int64_t)M.bd_edges.size(); v.bd_edges = M.bd_edges.data();
  v.num_bd_nodes = (int64_t)M.bd_nodes.size(); v.bd_nodes = M.bd_nodes.data();
  v.num_ports = (int32_t)M.port_side_ptr.size() - 1;
  v.port_side_ptr = M.port_side_ptr.data();
  v.side_edge_ptr = M.side_edge_ptr.data(); v.side_edges = M.side_edges.data();
  v.side_area = M.side_area.data(); v.side_dir = M.side_dir.data(); v.side_cnd = M.side_cnd.data();
  v.freq_unit = 1.0; v.freq_start = p.f_start; v.freq_end = p.f_end;
  v.nfreq = p.nfreq; v.freq_scale = p.freq_scale;
}

gp_status gen_microstrip(const gp_gen_params &p, gp_owned_mesh &M) {
  // One strip over the ground plane: strip spans x cells [Nx/6, Nx), the two
  // middle y cells, and the z cell Nz/2; two dielectric layers split at Nz/2.
  int64_t Nx = p.nx_cells, Ny = p.ny_cells, Nz = p.nz_cells;
  if (Nx < 4 || Ny < 4 || Nz < 2) return gp::fail(GP_ERR_ARG, "microstrip needs at least 4x4x2 cells");
  M.xn = axis(Nx, p.dx, p.grade); M.yn = axis(Ny, p.dy, p.grade * 0.5); M.zn = axis(Nz, p.dz, p.grade * 0.25);
  Grid g = Grid::make(Nx + 1, Ny + 1, Nz + 1);
  int64_t zs = Nz / 2, xs = Nx / 6, ys = Ny / 2 - 1;
  std::vector<Box> boxes{{xs, Nx, ys, ys + 2, zs, zs + 1, 0}};
  mark_conductors(M, g, boxes);
  M.epsn.resize(Nz + 1);
  for (int64_t iz = 0; iz <= Nz; ++iz) M.epsn[iz] = iz < zs ? p.eps_lo : p.eps_hi;
  M.sign.assign(Nz + 1, p.sigma);
  int32_t strip = M.mark_node[g.node(zs, xs, ys + 1)], ground = M.mark_node[0];
  M.cdt_mark_port[ground - 1] = -1;
  M.cdt_mark_port[strip - 1] = 1;
  M.port_side_ptr = {0, 1};
  M.side_edge_ptr = {0};
  add_z_port(M, g, xs, ys + 1, 0, zs, +1);
  boundary_sets(M, g);
  finish_view(M, p);
  return GP_OK;
}

// Parallel plate: a conductor plate one cell thick over the whole domain at
// cell zp = Nz / 2 above the PEC ground plane, two dielectric layers between
// them (eps_lo for cells [0, zp / 2), eps_hi for [zp / 2, zp); eps_hi above
// the plate), one z-port from the ground to the plate at the domain centre.
// With no fringing (the plate spans the lateral walls, which are natural
// boundaries) its capacitance is eps0 * area / sum(d_i / eps_i) exactly, for
// the discrete operator as for the continuum.
gp_status gen_plate(const gp_gen_params &p, gp_owned_mesh &M) {
  int64_t Nx = p.nx_cells, Ny = p.ny_cells, Nz = p.nz_cells;
  if (Nx < 2 || Ny < 2 || Nz < 6) return gp::fail(GP_ERR_ARG, "parallel plate needs at least 2x2x6 cells");
  M.xn = axis(Nx, p.dx, p.grade); M.yn = axis(Ny, p.dy, p.grade * 0.5); M.zn = axis(Nz, p.dz, 0.0);
  Grid g = Grid::make(Nx + 1, Ny + 1, Nz + 1);
  const int64_t zp = Nz / 2, z1 = zp / 2;
  std::vector<Box> boxes{{0, Nx, 0, Ny, zp, zp + 1, 0}};
  mark_conductors(M, g, boxes);
  M.epsn.resize(Nz + 1);
  for (int64_t iz = 0; iz <= Nz; ++iz) M.epsn[iz] = iz < z1 ? p.eps_lo : p.eps_hi;
  M.sign.assign(Nz + 1, p.sigma);
  const int64_t xc = Nx / 2, yc = Ny / 2;
  int32_t plate = M.mark_node[g.node(zp, xc, yc)], ground = M.mark_node[0];
  M.cdt_mark_port[ground - 1] = -1;
  M.cdt_mark_port[plate - 1] = 1;
  M.port_side_ptr = {0, 1};
  M.side_edge_ptr = {0};
  add_z_port(M, g, xc, yc, 0, zp, +1);
  boundary_sets(M, g);
  finish_view(M, p);
  return GP_OK;
}

gp_status gen_manhattan(const gp_gen_params &p, gp_owned_mesh &M) {
  int64_t Nx = p.nx_cells, Ny = p.ny_cells, Nz = p.nz_cells;
  int nm = p.n_metal;
  if (Nx < 8 || Ny < 8 || nm < 1 || Nz < 2 * nm + 2) return gp::fail(GP_ERR_ARG, "manhattan grid too small for the metal stack");
  if (p.wmin < 1 || p.wmax < p.wmin || p.wmax >= Nx || p.wmax >= Ny) return gp::fail(GP_ERR_ARG, "bad rectangle width range");
  Rng rng(p.seed);
  M.xn = axis(Nx, p.dx, p.grade); M.yn = axis(Ny, p.dy, p.grade * 0.5); M.zn = axis(Nz, p.dz, p.grade * 0.25);
  Grid g = Grid::make(Nx + 1, Ny + 1, Nz + 1);
  // metal k occupies z cell mz[k] (node planes mz..mz+1)
  std::vector<int64_t> mz(nm);
  for (int k = 0; k < nm; ++k) mz[k] = (int64_t)std::floor((k + 1) * (double)Nz / (nm + 1));
  // relative permittivity per stack layer: dielectric gaps and metal cells alternate
  M.epsn.resize(Nz + 1);
  {
    // stack layers bottom-up: dielectric gap, metal cell, ..., top dielectric
    std::vector<std::pair<int64_t, int64_t>> groups;
    int64_t lo = 0;
    for (int k = 0; k < nm; ++k) {
      if (mz[k] > lo) groups.push_back({lo, mz[k]});
      groups.push_back({mz[k], mz[k] + 1});
      lo = mz[k] + 1;
    }
    if (lo < Nz) groups.push_back({lo, Nz});
    for (auto &gr : groups) {
      double er = p.eps_lo + (p.eps_hi - p.eps_lo) * rng.unif();
      for (int64_t j = gr.first; j < gr.second; ++j) M.epsn[j] = er;
    }
    M.epsn[Nz] = M.epsn[Nz - 1];  // top surface entry (src/mesh.cpp:327-328)
  }
  M.sign.assign(Nz + 1, p.sigma);
  std::vector<Box> boxes;
  std::vector<std::vector<uint8_t>> cov(nm, std::vector<uint8_t>(Nx * Ny, 0));
  for (int k = 0; k < nm; ++k) {
    int64_t covered = 0;
    int guard = 0;
    while ((double)covered < p.fill * (double)(Nx * Ny) && guard++ < 1000000) {
      int64_t wx = rng.uni(p.wmin, p.wmax), wy = rng.uni(p.wmin, p.wmax);
      int64_t x0 = rng.uni(0, Nx - wx), y0 = rng.uni(0, Ny - wy);
      boxes.push_back({x0, x0 + wx, y0, y0 + wy, mz[k], mz[k] + 1, k});
      for (int64_t i = x0; i < x0 + wx; ++i)
        for (int64_t j = y0; j < y0 + wy; ++j)
          if (!cov[k][i * Ny + j]) { cov[k][i * Ny + j] = 1; ++covered; }
    }
  }
  // one-cell vias between consecutive metals where both carry metal
  for (int k = 0; k + 1 < nm; ++k) {
    std::vector<int64_t> cand;
    for (int64_t c = 0; c < Nx * Ny; ++c)
      if (cov[k][c] && cov[k + 1][c]) cand.push_back(c);
    if (cand.empty()) continue;
    double want = p.via_density * (double)(Nx * Ny);
    double prob = std::min(1.0, want / (double)cand.size());
    for (int64_t c : cand) {
      if (rng.unif() >= prob) continue;
      int64_t cx = c / Ny, cy = c % Ny;
      boxes.push_back({cx, cx + 1, cy, cy + 1, mz[k] + 1, mz[k + 1], -1});
    }
  }
  mark_conductors(M, g, boxes);
  int32_t ground = M.mark_node[0];
  M.cdt_mark_port[ground - 1] = -1;
  // ports: vertical lines from the ground plane up to the bottom of the top
  // metal, through dielectric only, preferably on distinct conductors
  int64_t ztop = mz[nm - 1];
  M.port_side_ptr = {0};
  M.side_edge_ptr = {0};
  std::vector<int32_t> used;
  int placed = 0;
  for (int attempt = 0; attempt < 200000 && placed < p.n_ports; ++attempt) {
    int64_t ix = rng.uni(1, Nx - 1), iy = rng.uni(1, Ny - 1);
    int32_t c = M.mark_node[g.node(ztop, ix, iy)];
    if (c == 0 || c == ground) continue;
    bool clear = true;
    for (int64_t iz = 1; iz < ztop && clear; ++iz) clear = M.mark_node[g.node(iz, ix, iy)] == 0;
    if (!clear) continue;
    bool dup = std::find(used.begin(), used.end(), c) != used.end();
    if (dup && attempt < 100000) continue;
    used.push_back(c);
    M.cdt_mark_port[c - 1] = placed + 1;
    add_z_port(M, g, ix, iy, 0, ztop, +1);
    M.port_side_ptr.push_back((int32_t)M.side_area.size());
    ++placed;
  }
  if (placed < p.n_ports) return gp::fail(GP_ERR_ARG, "could not place the requested ports");
  boundary_sets(M, g);
  finish_view(M, p);
  return GP_OK;
}

}  // namespace

extern "C" gp_status gp_generate(const gp_gen_params *p, gp_owned_mesh **out) {
  if (!p || !out) return gp::fail(GP_ERR_ARG, "null argument");
  if (!(p->dx > 0 && p->dy > 0 && p->dz > 0) || p->nfreq < 1 || !(p->f_start > 0) || !(p->f_end > 0))
    return gp::fail(GP_ERR_ARG, "bad generator parameters");
  auto *M = new gp_owned_mesh();
  gp_status st = p->kind == GP_GEN_MICROSTRIP ? gen_microstrip(*p, *M)
               : p->kind == GP_GEN_MANHATTAN ? gen_manhattan(*p, *M)
               : p->kind == GP_GEN_PLATE ? gen_plate(*p, *M)
               : gp::fail(GP_ERR_ARG, "unknown generator kind");
  if (st != GP_OK) { delete M; return st; }
  *out = M;
  return GP_OK;
}

extern "C" const gp_mesh *gp_owned_mesh_view(const gp_owned_mesh *m) { return m ? &m->view : nullptr; }
extern "C" void gp_owned_mesh_free(gp_owned_mesh *m) { delete m; }

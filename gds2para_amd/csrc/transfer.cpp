// transfer.cpp -- host construction of the structured multigrid transfer
// operators (see transfer.hpp).  O(edges) loops, run once per layout.
#include "transfer.hpp"

#include <algorithm>
#include <cmath>

namespace gp {

namespace {

struct Wt { int64_t line; double w; };

// Interpolation of fine node-line i from the coarse lines (indices into the
// fine coordinate array): one parent when i is a coarse line, else two with
// linear weights.
int line_interp(int64_t i, const std::vector<int64_t> &lines, const std::vector<double> &cf, Wt out[2]) {
  auto it = std::upper_bound(lines.begin(), lines.end(), i);
  int64_t c0 = (int64_t)(it - lines.begin()) - 1;
  if (lines[c0] == i) { out[0] = {c0, 1.0}; return 1; }
  double a = cf[lines[c0]], b = cf[lines[c0 + 1]];
  double t = (cf[i] - a) / (b - a);
  out[0] = {c0, 1.0 - t};
  out[1] = {c0 + 1, t};
  return 2;
}

// Coarse interval containing fine interval i (fine cell i spans lines i..i+1).
int64_t interval_of(int64_t i, const std::vector<int64_t> &lines) {
  auto it = std::upper_bound(lines.begin(), lines.end(), i);
  return (int64_t)(it - lines.begin()) - 1;
}

std::vector<int64_t> pick_lines(int64_t ncell, bool coarsen) {
  std::vector<int64_t> l;
  if (!coarsen) {
    for (int64_t i = 0; i <= ncell; ++i) l.push_back(i);
    return l;
  }
  for (int64_t i = 0; i <= ncell; i += 2) l.push_back(i);
  if (l.back() != ncell) l.push_back(ncell);
  return l;
}

double mean_cell(const std::vector<double> &c) { return (c.back() - c.front()) / (double)(c.size() - 1); }

}  // namespace

void level0_grid(const gp_mesh &m, LevelGrid *L) {
  L->x.assign(m.xn, m.xn + m.nx);
  L->y.assign(m.yn, m.yn + m.ny);
  L->z.assign(m.zn, m.zn + m.nz);
  L->g = Grid::make(m.nx, m.ny, m.nz);
  const Grid &g = L->g;
  std::vector<uint8_t> rm(g.Ne, 0), fixed(g.Nn, 0);
  for (int64_t i = 0; i < m.num_bd_edges; ++i) rm[m.bd_edges[i]] = 1;
  for (int64_t i = 0; i < m.num_bd_nodes; ++i) fixed[m.bd_nodes[i]] = 1;
  // a node touching a removed edge is held at zero potential as well
  for (int64_t e = 0; e < g.Ne; ++e)
    if (rm[e]) { int64_t a, b; g.ends(e, a, b); fixed[a] = fixed[b] = 1; }
  L->emap.assign(g.Ne, -1);
  L->erev.clear();
  for (int64_t e = 0; e < g.Ne; ++e)
    if (!rm[e]) { L->emap[e] = (int64_t)L->erev.size(); L->erev.push_back(e); }
  L->nmap.assign(g.Nn, -1);
  L->nrev.clear();
  for (int64_t v = 0; v < g.Nn; ++v)
    if (!fixed[v]) { L->nmap[v] = (int64_t)L->nrev.size(); L->nrev.push_back(v); }
}

bool coarsen_grid(const LevelGrid &f, LevelGrid *c) {
  const Grid &gf = f.g;
  const double hx = mean_cell(f.x), hy = mean_cell(f.y), hz = mean_cell(f.z);
  const double hmin = std::min(hx, std::min(hy, hz));
  // coarsen an axis only where it is not already much coarser than the finest
  // axis (semi-coarsening keeps the coarse cells from turning into slabs)
  bool cx = gf.Nx >= 2 && hx <= 2.5 * hmin;
  bool cy = gf.Ny >= 2 && hy <= 2.5 * hmin;
  bool cz = gf.Nz >= 2 && hz <= 2.5 * hmin;
  if (!cx && !cy && !cz) {
    cx = gf.Nx >= 2; cy = gf.Ny >= 2; cz = gf.Nz >= 2;
    if (!cx && !cy && !cz) return false;
  }
  c->lx = pick_lines(gf.Nx, cx);
  c->ly = pick_lines(gf.Ny, cy);
  c->lz = pick_lines(gf.Nz, cz);
  c->x.clear(); c->y.clear(); c->z.clear();
  for (int64_t i : c->lx) c->x.push_back(f.x[i]);
  for (int64_t i : c->ly) c->y.push_back(f.y[i]);
  for (int64_t i : c->lz) c->z.push_back(f.z[i]);
  c->g = Grid::make((int64_t)c->x.size(), (int64_t)c->y.size(), (int64_t)c->z.size());
  const Grid &g = c->g;
  // coarse node is free iff the fine node at the same place is free
  c->nmap.assign(g.Nn, -1);
  c->nrev.clear();
  for (int64_t iz = 0; iz <= g.Nz; ++iz)
    for (int64_t ix = 0; ix <= g.Nx; ++ix)
      for (int64_t iy = 0; iy <= g.Ny; ++iy) {
        int64_t v = g.node(iz, ix, iy);
        if (f.nmap[gf.node(c->lz[iz], c->lx[ix], c->ly[iy])] >= 0) { c->nmap[v] = (int64_t)c->nrev.size(); c->nrev.push_back(v); }
      }
  // coarse edge is removed iff every fine edge along it is removed
  c->emap.assign(g.Ne, -1);
  c->erev.clear();
  for (int64_t e = 0; e < g.Ne; ++e) {
    int d; int64_t iz, ix, iy;
    g.decode(e, d, iz, ix, iy);
    bool any = false;
    if (d == 1) {
      for (int64_t i = c->lx[ix]; i < c->lx[ix + 1] && !any; ++i) any = f.emap[gf.ex(c->lz[iz], i, c->ly[iy])] >= 0;
    } else if (d == 0) {
      for (int64_t i = c->ly[iy]; i < c->ly[iy + 1] && !any; ++i) any = f.emap[gf.ey(c->lz[iz], c->lx[ix], i)] >= 0;
    } else {
      for (int64_t i = c->lz[iz]; i < c->lz[iz + 1] && !any; ++i) any = f.emap[gf.ez(i, c->lx[ix], c->ly[iy])] >= 0;
    }
    if (any) { c->emap[e] = (int64_t)c->erev.size(); c->erev.push_back(e); }
  }
  return true;
}

void build_edge_prolongation(const LevelGrid &f, const LevelGrid &c, HostCsr *P) {
  const Grid &gf = f.g, &gc = c.g;
  P->nr = f.ne();
  P->nc = c.ne();
  P->rp.assign(P->nr + 1, 0);
  P->ci.clear();
  P->v.clear();
  std::vector<std::pair<int64_t, double>> row;
  for (int64_t k = 0; k < f.ne(); ++k) {
    int64_t e = f.erev[k];
    int d; int64_t iz, ix, iy;
    gf.decode(e, d, iz, ix, iy);
    row.clear();
    Wt wa[2], wb[2];
    if (d == 1) {  // x-edge: constant along x, interpolated in y and z
      int64_t cxi = interval_of(ix, c.lx);
      int na = line_interp(iy, c.ly, f.y, wa), nb = line_interp(iz, c.lz, f.z, wb);
      for (int b = 0; b < nb; ++b)
        for (int a = 0; a < na; ++a) {
          int64_t ce = c.emap[gc.ex(wb[b].line, cxi, wa[a].line)];
          if (ce >= 0) row.push_back({ce, wa[a].w * wb[b].w});
        }
    } else if (d == 0) {  // y-edge
      int64_t cyi = interval_of(iy, c.ly);
      int na = line_interp(ix, c.lx, f.x, wa), nb = line_interp(iz, c.lz, f.z, wb);
      for (int b = 0; b < nb; ++b)
        for (int a = 0; a < na; ++a) {
          int64_t ce = c.emap[gc.ey(wb[b].line, wa[a].line, cyi)];
          if (ce >= 0) row.push_back({ce, wa[a].w * wb[b].w});
        }
    } else {  // z-edge
      int64_t czi = interval_of(iz, c.lz);
      int na = line_interp(ix, c.lx, f.x, wa), nb = line_interp(iy, c.ly, f.y, wb);
      for (int b = 0; b < nb; ++b)
        for (int a = 0; a < na; ++a) {
          int64_t ce = c.emap[gc.ez(czi, wa[a].line, wb[b].line)];
          if (ce >= 0) row.push_back({ce, wa[a].w * wb[b].w});
        }
    }
    std::sort(row.begin(), row.end());
    for (auto &p : row) { P->ci.push_back((int32_t)p.first); P->v.push_back(p.second); }
    P->rp[k + 1] = (int64_t)P->ci.size();
  }
}

void build_node_prolongation(const LevelGrid &f, const LevelGrid &c, HostCsr *P) {
  const Grid &gf = f.g, &gc = c.g;
  P->nr = f.nn();
  P->nc = c.nn();
  P->rp.assign(P->nr + 1, 0);
  P->ci.clear();
  P->v.clear();
  std::vector<std::pair<int64_t, double>> row;
  for (int64_t k = 0; k < f.nn(); ++k) {
    int64_t v = f.nrev[k];
    int64_t iz = v / gf.Nns, ix = (v % gf.Nns) / (gf.Ny + 1), iy = (v % gf.Nns) % (gf.Ny + 1);
    Wt wx[2], wy[2], wz[2];
    int nx = line_interp(ix, c.lx, f.x, wx), ny = line_interp(iy, c.ly, f.y, wy), nz = line_interp(iz, c.lz, f.z, wz);
    row.clear();
    for (int a = 0; a < nz; ++a)
      for (int b = 0; b < nx; ++b)
        for (int q = 0; q < ny; ++q) {
          int64_t cn = c.nmap[gc.node(wz[a].line, wx[b].line, wy[q].line)];
          if (cn >= 0) row.push_back({cn, wz[a].w * wx[b].w * wy[q].w});
        }
    std::sort(row.begin(), row.end());
    for (auto &p : row) { P->ci.push_back((int32_t)p.first); P->v.push_back(p.second); }
    P->rp[k + 1] = (int64_t)P->ci.size();
  }
}

void build_gradient(const LevelGrid &L, HostCsr *G) {
  const Grid &g = L.g;
  G->nr = L.ne();
  G->nc = L.nn();
  G->rp.assign(G->nr + 1, 0);
  G->ci.clear();
  G->v.clear();
  for (int64_t k = 0; k < L.ne(); ++k) {
    int64_t e = L.erev[k];
    int d; int64_t iz, ix, iy;
    g.decode(e, d, iz, ix, iy);
    double len = d == 2 ? L.z[iz + 1] - L.z[iz] : d == 1 ? L.x[ix + 1] - L.x[ix] : L.y[iy + 1] - L.y[iy];
    int64_t a, b;
    g.ends(e, a, b);
    if (L.nmap[a] >= 0) { G->ci.push_back((int32_t)L.nmap[a]); G->v.push_back(-1.0 / len); }
    if (L.nmap[b] >= 0) { G->ci.push_back((int32_t)L.nmap[b]); G->v.push_back(1.0 / len); }
    G->rp[k + 1] = (int64_t)G->ci.size();
  }
}

void edge_dual_volume(const LevelGrid &L, std::vector<double> *W) {
  const Grid &g = L.g;
  auto dual = [](const std::vector<double> &c, int64_t i) {
    int64_t N = (int64_t)c.size() - 1;
    if (i == 0) return c[1] - c[0];
    if (i == N) return c[N] - c[N - 1];
    return (c[i + 1] - c[i - 1]) / 2;
  };
  W->resize(L.ne());
  for (int64_t k = 0; k < L.ne(); ++k) {
    int d; int64_t iz, ix, iy;
    g.decode(L.erev[k], d, iz, ix, iy);
    double w;
    if (d == 0) w = dual(L.x, ix) * dual(L.z, iz) * (L.y[iy + 1] - L.y[iy]);
    else if (d == 1) w = dual(L.y, iy) * dual(L.z, iz) * (L.x[ix + 1] - L.x[ix]);
    else w = dual(L.x, ix) * dual(L.y, iy) * (L.z[iz + 1] - L.z[iz]);
    (*W)[k] = w;
  }
}

}  // namespace gp

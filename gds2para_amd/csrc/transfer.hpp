// transfer.hpp -- structured multigrid levels of the Yee mesh.
//
// Level 0 is the mesh itself (kept edges after PEC removal, free nodes = nodes
// off the PEC boundary).  A coarse level keeps every other node line of each
// coarsened axis (plus the last line), so coarse nodes are fine nodes and the
// lowest-order edge (Nedelec) prolongation P_e and the trilinear node
// prolongation P_n commute with the discrete gradient: P_e G_c = G_f P_n.
// That identity carries the gradient null space of the curl-curl operator
// exactly through the Galerkin hierarchy.
#pragma once
#include <vector>

#include "gds2para.h"
#include "gp_grid.hpp"
#include "host_csr.hpp"

namespace gp {

struct LevelGrid {
  std::vector<double> x, y, z;     // node coordinates of this level
  Grid g{};
  std::vector<int64_t> emap, erev;  // edge -> kept index (-1 removed), kept -> edge
  std::vector<int64_t> nmap, nrev;  // node -> free index (-1 fixed), free -> node
  std::vector<int64_t> lx, ly, lz;  // coarse line -> line index on the finer level
  int64_t ne() const { return (int64_t)erev.size(); }
  int64_t nn() const { return (int64_t)nrev.size(); }
};

// Level 0 from the mesh contract.
void level0_grid(const gp_mesh &m, LevelGrid *L);
// Next coarser level; returns false when no axis can be coarsened.
bool coarsen_grid(const LevelGrid &f, LevelGrid *c);
// Edge prolongation (fine kept edges x coarse kept edges), node prolongation
// (fine free nodes x coarse free nodes), gradient (kept edges x free nodes).
void build_edge_prolongation(const LevelGrid &f, const LevelGrid &c, HostCsr *P);
void build_node_prolongation(const LevelGrid &f, const LevelGrid &c, HostCsr *P);
void build_gradient(const LevelGrid &L, HostCsr *G);
// Dual volume (dual area x length) of every kept edge of level 0, the
// symmetrizer of S: diag(W) S is symmetric.
void edge_dual_volume(const LevelGrid &L, std::vector<double> *W);

}  // namespace gp

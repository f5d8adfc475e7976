/*
 * gp_oracle.c -- CPU restatement of gds2Para's simulate-mode hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is
 * compared against; it is linked only by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg, never by the product.
 *
 * Each function follows the reference file:line it restates (paths relative to
 * the gds2Para tree).  The mesh comes in through the same gp_mesh struct as the
 * product's C ABI (include/gds2para.h).  Plain C99, FP64, no FMA contraction
 * (built with -ffp-contract=off) so the products and sums round exactly as the
 * written expressions say.
 *
 * Pinning (how each function is checked beyond itself; see tests/test_oracle.py):
 *   or_stiffness   nnz closed form N_edge + 12 N_patch - removed; dense Sh*Se^T
 *                  brute force in numpy on tiny grids; curl(grad) = 0 (S G ~ 0);
 *                  W S symmetric with W the edge dual volume.
 *   or_v0d/or_ad   column count = dielectric non-boundary nodes + conductor
 *                  columns; V0d = -grad so Se^T V0d = 0; Ad rows sum to the
 *                  boundary-coupling terms (zero row sums away from boundaries).
 *   or_v0c/or_ac   singleton columns per conductor node minus the omitted ones;
 *                  Ac rows sum to zero inside conductors.
 *   or_diag/or_rhs A(w) diagonal and J as reference() builds them; checked by
 *                  direct solves (SuperLU) against closed-form low-frequency
 *                  capacitance limits and reciprocity.
 *   or_vh_coeffs   eps / sig per unknown; the Ritz values built on them match
 *                  the eigenvalues of find_reference_Vh's dense pencil and
 *                  the Ritz vectors satisfy the quadratic eigen-equation.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/gds2para.h"

#define OR_PI 3.14159265358979323846
#define OR_MU (4 * OR_PI * 1.e-7)                     /* src/fdtd.hpp:45 */
#define OR_CSPED (299792458.)                         /* src/fdtd.hpp:46 */
#define OR_EPSILON0 (1. / (OR_CSPED * OR_CSPED * OR_MU)) /* src/fdtd.hpp:47 */

typedef struct {
  int64_t Nx, Ny, Nz, Es, Ev, El, Ne, Nns, Nn, Np;
} or_grid;

static or_grid or_make_grid(const gp_mesh *m) {
  or_grid g;
  g.Nx = m->nx - 1; g.Ny = m->ny - 1; g.Nz = m->nz - 1;
  g.Es = g.Ny * (g.Nx + 1) + g.Nx * (g.Ny + 1); /* src/mesh.cpp:397 */
  g.Ev = (g.Nx + 1) * (g.Ny + 1);               /* src/mesh.cpp:398 */
  g.El = g.Es + g.Ev;
  g.Ne = g.Es * (g.Nz + 1) + g.Ev * g.Nz;       /* src/mesh.cpp:399 */
  g.Nns = g.Ev;
  g.Nn = (g.Nz + 1) * g.Nns;
  g.Np = g.Nx * g.Ny * (g.Nz + 1) + ((g.Nx + 1) * g.Ny + (g.Ny + 1) * g.Nx) * g.Nz; /* src/mesh.cpp:405 */
  return g;
}

/* removed-edge flags (lbde u ubde) */
static unsigned char *or_removed(const gp_mesh *m, const or_grid *g) {
  unsigned char *rm = (unsigned char *)calloc((size_t)g->Ne, 1);
  for (int64_t i = 0; i < m->num_bd_edges; ++i) rm[m->bd_edges[i]] = 1;
  return rm;
}

/* setMapEdge (src/fdtd.hpp:1508-1529) */
static int64_t *or_map_edge(const gp_mesh *m, const or_grid *g) {
  int64_t *map = (int64_t *)malloc(sizeof(int64_t) * (size_t)g->Ne);
  unsigned char *rm = or_removed(m, g);
  int64_t count = 0;
  for (int64_t e = 0; e < g->Ne; ++e) {
    if (rm[e]) { map[e] = -1; ++count; }
    else map[e] = e - count;
  }
  free(rm);
  return map;
}

/* ------------------------------------------------------------------------- */
/* Se / Sh in COO, in the reference's emission order                          */
/* (src/generateStiff.cpp:33-245 for Se, :249-484 for Sh).                    */
/* ------------------------------------------------------------------------- */
typedef struct { int64_t row, col; double se, sh; } or_fe; /* one (edge, face) incidence */

static or_fe *or_build_se_sh(const gp_mesh *m, const or_grid *g, int64_t *count) {
  const int64_t Nx = g->Nx, Ny = g->Ny, Nz = g->Nz, El = g->El, Es = g->Es;
  const double *xn = m->xn, *yn = m->yn, *zn = m->zn;
  or_fe *t = (or_fe *)malloc(sizeof(or_fe) * (size_t)(g->Np * 4));
  int64_t k = 0, f = 0;
  /* dual lengths dxa/dya/dza (src/generateStiff.cpp:250-267) */
  double *dxa = (double *)malloc(sizeof(double) * (size_t)(Nx + 1));
  double *dya = (double *)malloc(sizeof(double) * (size_t)(Ny + 1));
  double *dza = (double *)malloc(sizeof(double) * (size_t)(Nz + 1));
  dxa[0] = xn[1] - xn[0]; dya[0] = yn[1] - yn[0]; dza[0] = zn[1] - zn[0];
  dxa[Nx] = xn[Nx] - xn[Nx - 1]; dya[Ny] = yn[Ny] - yn[Ny - 1]; dza[Nz] = zn[Nz] - zn[Nz - 1];
  for (int64_t i = 1; i < Nx; i++) dxa[i] = (xn[i + 1] - xn[i - 1]) / 2;
  for (int64_t i = 1; i < Ny; i++) dya[i] = (yn[i + 1] - yn[i - 1]) / 2;
  for (int64_t i = 1; i < Nz; i++) dza[i] = (zn[i + 1] - zn[i - 1]) / 2;

#define EMIT(R, SE, SH) do { t[k].row = (R); t[k].col = f; t[k].se = (SE); t[k].sh = (SH); ++k; } while (0)
  for (int64_t indz = 1; indz <= Nz; indz++) {
    for (int64_t indx = 1; indx <= Nx; indx++) {
      for (int64_t indy = 1; indy <= Ny; indy++) {
        const int64_t b = (indz - 1) * El;
        const double hx = xn[indx] - xn[indx - 1], hy = yn[indy] - yn[indy - 1], hz = zn[indz] - zn[indz - 1];
        /* xy face at plane indz-1 */
        EMIT(b + (indx - 1) * Ny + indy - 1, -1 / hx, -dza[indz - 1] / (dxa[indx - 1] * dza[indz - 1]));
        EMIT(b + indx * Ny + indy - 1, 1 / hx, dza[indz - 1] / (dxa[indx] * dza[indz - 1]));
        EMIT(b + Ny * (Nx + 1) + (indx - 1) * (Ny + 1) + indy - 1, 1 / hy, dza[indz - 1] / (dza[indz - 1] * dya[indy - 1]));
        EMIT(b + Ny * (Nx + 1) + (indx - 1) * (Ny + 1) + indy, -1 / hy, -dza[indz - 1] / (dza[indz - 1] * dya[indy]));
        ++f;
        /* yz face at x = indx-1 */
        EMIT(b + (indx - 1) * Ny + indy - 1, 1 / hz, dxa[indx - 1] / (dxa[indx - 1] * dza[indz - 1]));
        EMIT(b + Es + (indx - 1) * (Ny + 1) + indy - 1, -1 / hy, -dxa[indx - 1] / (dxa[indx - 1] * dya[indy - 1]));
        EMIT(b + Es + (indx - 1) * (Ny + 1) + indy, 1 / hy, dxa[indx - 1] / (dxa[indx - 1] * dya[indy]));
        EMIT(indz * El + (indx - 1) * Ny + indy - 1, -1 / hz, -dxa[indx - 1] / (dxa[indx - 1] * dza[indz]));
        ++f;
        /* xz face at y = indy-1 */
        EMIT(b + Ny * (Nx + 1) + (indx - 1) * (Ny + 1) + indy - 1, -1 / hz, -dya[indy - 1] / (dza[indz - 1] * dya[indy - 1]));
        EMIT(b + Es + (indx - 1) * (Ny + 1) + indy - 1, 1 / hx, dya[indy - 1] / (dxa[indx - 1] * dya[indy - 1]));
        EMIT(b + Es + indx * (Ny + 1) + indy - 1, -1 / hx, -dya[indy - 1] / (dya[indy - 1] * dxa[indx]));
        EMIT(indz * El + Ny * (Nx + 1) + (indx - 1) * (Ny + 1) + indy - 1, 1 / hz, dya[indy - 1] / (dya[indy - 1] * dza[indz]));
        ++f;
        if (indz == Nz) { /* top xy face (src/generateStiff.cpp:171-189, 408-428) */
          const int64_t bt = indz * El;
          EMIT(bt + (indx - 1) * Ny + indy - 1, -1 / hx, -dza[indz - 1] / (dza[indz - 1] * dxa[indx - 1]));
          EMIT(bt + indx * Ny + indy - 1, 1 / hx, dza[indz - 1] / (dza[indz - 1] * dxa[indx]));
          EMIT(bt + Ny * (Nx + 1) + (indx - 1) * (Ny + 1) + indy - 1, 1 / hy, dza[indz - 1] / (dza[indz - 1] * dya[indy - 1]));
          EMIT(bt + Ny * (Nx + 1) + (indx - 1) * (Ny + 1) + indy, -1 / hy, -dza[indz - 1] / (dza[indz - 1] * dya[indy]));
          ++f;
        }
      }
    }
  }
  { /* rightmost yz plane (src/generateStiff.cpp:193-218, 430-456) */
    const int64_t indx = Nx + 1;
    for (int64_t indz = 1; indz <= Nz; indz++)
      for (int64_t indy = 1; indy <= Ny; indy++) {
        const int64_t b = (indz - 1) * El;
        const double hy = yn[indy] - yn[indy - 1], hz = zn[indz] - zn[indz - 1];
        EMIT(b + (indx - 1) * Ny + indy - 1, 1 / hz, dxa[indx - 1] / (dxa[indx - 1] * dza[indz - 1]));
        EMIT(b + Es + (indx - 1) * (Ny + 1) + indy - 1, -1 / hy, -dxa[indx - 1] / (dxa[indx - 1] * dya[indy - 1]));
        EMIT(b + Es + (indx - 1) * (Ny + 1) + indy, 1 / hy, dxa[indx - 1] / (dxa[indx - 1] * dya[indy]));
        EMIT(indz * El + (indx - 1) * Ny + indy - 1, -1 / hz, -dxa[indx - 1] / (dxa[indx - 1] * dza[indz]));
        ++f;
      }
  }
  { /* farthest xz plane (src/generateStiff.cpp:221-245, 458-484) */
    const int64_t indy = Ny + 1;
    for (int64_t indz = 1; indz <= Nz; indz++)
      for (int64_t indx = 1; indx <= Nx; indx++) {
        const int64_t b = (indz - 1) * El;
        const double hx = xn[indx] - xn[indx - 1], hz = zn[indz] - zn[indz - 1];
        EMIT(b + Ny * (Nx + 1) + (indx - 1) * (Ny + 1) + indy - 1, -1 / hz, -dya[indy - 1] / (dya[indy - 1] * dza[indz - 1]));
        EMIT(b + Es + (indx - 1) * (Ny + 1) + indy - 1, 1 / hx, dya[indy - 1] / (dya[indy - 1] * dxa[indx - 1]));
        EMIT(b + Es + indx * (Ny + 1) + indy - 1, -1 / hx, -dya[indy - 1] / (dya[indy - 1] * dxa[indx]));
        EMIT(indz * El + Ny * (Nx + 1) + (indx - 1) * (Ny + 1) + indy - 1, 1 / hz, dya[indy - 1] / (dya[indy - 1] * dza[indz]));
        ++f;
      }
  }
#undef EMIT
  free(dxa); free(dya); free(dza);
  *count = k;
  return t;
}

/* exported for the tests: raw Se/Sh triplets in emission order */
int64_t or_se_sh_count(const gp_mesh *m) { or_grid g = or_make_grid(m); return g.Np * 4; }
int or_se_sh(const gp_mesh *m, int64_t *row, int64_t *col, double *se, double *sh) {
  or_grid g = or_make_grid(m);
  int64_t cnt = 0;
  or_fe *t = or_build_se_sh(m, &g, &cnt);
  for (int64_t i = 0; i < cnt; ++i) { row[i] = t[i].row; col[i] = t[i].col; se[i] = t[i].se; sh[i] = t[i].sh; }
  free(t);
  return cnt == g.Np * 4 ? 0 : 1;
}

typedef struct { int64_t key; int64_t seq; double v; } or_kv;
static int or_kv_cmp(const void *a, const void *b) {
  const or_kv *x = (const or_kv *)a, *y = (const or_kv *)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->seq < y->seq ? -1 : (x->seq > y->seq);
}

/*
 * S = Sh * Se^T / MU with PEC rows/columns removed and renumbered
 * (src/generateStiff.cpp:511-540, mklMatrixMulti_nt :897-980).  The product
 * is formed row by row over the faces of the row edge in ascending face
 * order; each entry accumulates its products in that order starting from 0.
 * Off-diagonal entries hold exactly one product (two edges share at most one
 * face); diagonal entries hold up to four.
 */
int64_t or_stiffness_nnz(const gp_mesh *m) {
  or_grid g = or_make_grid(m);
  int64_t cnt = 0;
  or_fe *t = or_build_se_sh(m, &g, &cnt);
  unsigned char *rm = or_removed(m, &g);
  /* faces hold 4 consecutive incidences; count per row the distinct kept partners */
  int64_t nnz = 0;
  int64_t *deg = (int64_t *)calloc((size_t)g.Ne, sizeof(int64_t));
  for (int64_t f = 0; f < cnt / 4; ++f)
    for (int a = 0; a < 4; ++a) {
      int64_t r = t[4 * f + a].row;
      if (rm[r]) continue;
      for (int b2 = 0; b2 < 4; ++b2)
        if (b2 != a && !rm[t[4 * f + b2].row]) deg[r]++;
    }
  for (int64_t e = 0; e < g.Ne; ++e) if (!rm[e]) nnz += 1 + deg[e];
  free(deg); free(rm); free(t);
  return nnz;
}

int or_stiffness(const gp_mesh *m, int64_t *row_ptr, int32_t *colidx, double *val) {
  or_grid g = or_make_grid(m);
  int64_t cnt = 0;
  or_fe *t = or_build_se_sh(m, &g, &cnt);
  unsigned char *rm = or_removed(m, &g);
  int64_t *map = or_map_edge(m, &g);
  /* transpose the incidences: faces of each edge in ascending face order */
  int64_t *ptr = (int64_t *)calloc((size_t)g.Ne + 1, sizeof(int64_t));
  for (int64_t i = 0; i < cnt; ++i) ptr[t[i].row + 1]++;
  for (int64_t e = 0; e < g.Ne; ++e) ptr[e + 1] += ptr[e];
  int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)g.Ne);
  memcpy(pos, ptr, sizeof(int64_t) * (size_t)g.Ne);
  int64_t *inc = (int64_t *)malloc(sizeof(int64_t) * (size_t)cnt); /* incidence index, by edge */
  for (int64_t i = 0; i < cnt; ++i) inc[pos[t[i].row]++] = i;    /* faces ascend with i */
  or_kv buf[64];
  int64_t r = 0, nz = 0;
  row_ptr[0] = 0;
  for (int64_t e = 0; e < g.Ne; ++e) {
    if (rm[e]) continue;
    int nb = 0;
    for (int64_t q = ptr[e]; q < ptr[e + 1]; ++q) {
      const or_fe *me = &t[inc[q]];
      int64_t f0 = inc[q] - (inc[q] % 4);
      for (int b2 = 0; b2 < 4; ++b2) {
        const or_fe *o = &t[f0 + b2];
        buf[nb].key = o->row;
        buf[nb].seq = nb;
        buf[nb].v = me->sh * o->se;
        ++nb;
      }
    }
    qsort(buf, (size_t)nb, sizeof(or_kv), or_kv_cmp);
    for (int i = 0; i < nb;) {
      int j = i;
      double s = 0.;
      while (j < nb && buf[j].key == buf[i].key) { s += buf[j].v; ++j; }
      if (!rm[buf[i].key]) {
        colidx[nz] = (int32_t)map[buf[i].key];
        val[nz] = s / OR_MU;
        ++nz;
      }
      i = j;
    }
    row_ptr[++r] = nz;
  }
  free(inc); free(pos); free(ptr); free(map); free(rm); free(t);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* V0d basis and Ad (merge_v0d1 src/fdtd.hpp:1823-2774, generateAd :2778-2862) */
/* ------------------------------------------------------------------------- */
static double or_avg(const double *c, int64_t i, int64_t n) { /* avg_length src/fdtd.hpp:4074 */
  if (i == 0) return c[1] - c[0];
  if (i == n - 1) return c[i] - c[i - 1];
  return (c[i + 1] - c[i - 1]) / 2;
}

static int or_is_bd_node(const gp_mesh *m, int64_t v) {
  int64_t lo = 0, hi = m->num_bd_nodes;
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (m->bd_nodes[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo < m->num_bd_nodes && m->bd_nodes[lo] == v;
}

typedef struct {
  int64_t n, cap;
  int64_t *row, *col;
  double *val, *aval;
} or_coo;

static void or_push(or_coo *c, int64_t r, int64_t col, double v, double av) {
  if (c->n == c->cap) {
    c->cap = c->cap ? 2 * c->cap : 1024;
    c->row = (int64_t *)realloc(c->row, sizeof(int64_t) * (size_t)c->cap);
    c->col = (int64_t *)realloc(c->col, sizeof(int64_t) * (size_t)c->cap);
    c->val = (double *)realloc(c->val, sizeof(double) * (size_t)c->cap);
    c->aval = (double *)realloc(c->aval, sizeof(double) * (size_t)c->cap);
  }
  c->row[c->n] = r; c->col[c->n] = col; c->val[c->n] = v; c->aval[c->n] = av; c->n++;
}
static void or_coo_free(or_coo *c) { free(c->row); free(c->col); free(c->val); free(c->aval); memset(c, 0, sizeof(*c)); }

/* the six incident edges of a node in the reference's order: lower z, upper z,
 * left x, right x, front y, back y; value +-1/h toward/away from the node */
typedef struct { int64_t e, other; double val; int axis, sgn; } or_inc;
static int or_incident(const gp_mesh *m, const or_grid *g, int64_t iz, int64_t ix, int64_t iy, or_inc out[6]) {
  int k = 0;
  const double *xn = m->xn, *yn = m->yn, *zn = m->zn;
  const int64_t El = g->El, Es = g->Es, Ny = g->Ny, Nx = g->Nx, Nns = g->Nns;
  if (iz != 0) { out[k].e = (iz - 1) * El + Es + ix * (Ny + 1) + iy; out[k].other = (iz - 1) * Nns + ix * (Ny + 1) + iy; out[k].val = -1 / (zn[iz] - zn[iz - 1]); out[k].axis = 2; out[k].sgn = -1; ++k; }
  if (iz != m->nz - 1) { out[k].e = iz * El + Es + ix * (Ny + 1) + iy; out[k].other = (iz + 1) * Nns + ix * (Ny + 1) + iy; out[k].val = 1 / (zn[iz + 1] - zn[iz]); out[k].axis = 2; out[k].sgn = 1; ++k; }
  if (ix != 0) { out[k].e = iz * El + Ny * (Nx + 1) + (ix - 1) * (Ny + 1) + iy; out[k].other = iz * Nns + (ix - 1) * (Ny + 1) + iy; out[k].val = -1 / (xn[ix] - xn[ix - 1]); out[k].axis = 0; out[k].sgn = -1; ++k; }
  if (ix != m->nx - 1) { out[k].e = iz * El + Ny * (Nx + 1) + ix * (Ny + 1) + iy; out[k].other = iz * Nns + (ix + 1) * (Ny + 1) + iy; out[k].val = 1 / (xn[ix + 1] - xn[ix]); out[k].axis = 0; out[k].sgn = 1; ++k; }
  if (iy != 0) { out[k].e = iz * El + ix * Ny + iy - 1; out[k].other = iz * Nns + ix * (Ny + 1) + iy - 1; out[k].val = -1 / (yn[iy] - yn[iy - 1]); out[k].axis = 1; out[k].sgn = -1; ++k; }
  if (iy != m->ny - 1) { out[k].e = iz * El + ix * Ny + iy; out[k].other = iz * Nns + ix * (Ny + 1) + iy + 1; out[k].val = 1 / (yn[iy + 1] - yn[iy]); out[k].axis = 1; out[k].sgn = 1; ++k; }
  return k;
}

typedef struct {
  or_coo v0;          /* row = original edge id, col, val, aval */
  int64_t cols;       /* leng_v0d1 / leng_v0c */
  int64_t *map;       /* [N_node] column + 1 or 0 */
} or_basis;

static void or_basis_free(or_basis *b) { or_coo_free(&b->v0); free(b->map); memset(b, 0, sizeof(*b)); }

static void or_build_v0d(const gp_mesh *m, or_basis *B) {
  or_grid g = or_make_grid(m);
  memset(B, 0, sizeof(*B));
  B->map = (int64_t *)calloc((size_t)g.Nn, sizeof(int64_t));
  const int64_t nx = m->nx, ny = m->ny, nz = m->nz, Nns = g.Nns, Ny = g.Ny;
  /* singleton dielectric columns in (iz, ix, iy) scan order; the block sizes
   * are 0 (src/matrixCon.cpp:35-40) so every group holds one node */
  int64_t count = 1;
  for (int64_t iz = 0; iz < nz; iz++)
    for (int64_t ix = 0; ix < nx; ix++)
      for (int64_t iy = 0; iy < ny; iy++) {
        int64_t nno = iz * Nns + ix * (Ny + 1) + iy;
        if (!((iz != 0 && iz != nz - 1) || !or_is_bd_node(m, nno))) continue;
        if (m->mark_node[nno] != 0) continue;
        B->map[nno] = count;
        double lx = or_avg(m->xn, ix, nx), ly = or_avg(m->yn, iy, ny), lz = or_avg(m->zn, iz, nz);
        or_inc in[6];
        int k = or_incident(m, &g, iz, ix, iy, in);
        for (int q = 0; q < k; ++q) {
          double l = in[q].axis == 0 ? lx : in[q].axis == 1 ? ly : lz;
          or_push(&B->v0, in[q].e, count - 1, in[q].val, in[q].sgn > 0 ? 1 / l : -1 / l);
        }
        count++;
      }
  /* conductor columns (V0d2).  Count pass: map = running count, the count
   * advances only for conductors with dielectric-facing edges
   * (src/fdtd.hpp:2430-2500).  Fill pass: `mark` is not reset per conductor
   * (src/fdtd.hpp:2673-2760), so once one conductor had entries every later
   * non-reference conductor takes a column. */
  const int64_t ndiel = count - 1;
  int mark = 0;
  for (int32_t c = 0; c < m->num_cdt; ++c) {
    if (m->cdt_mark_port[c] == -1) continue;
    mark = 0;
    for (int64_t v = 0; v < g.Nn; ++v) {
      if (m->mark_node[v] != c + 1) continue;
      B->map[v] = count;
      int64_t iz = v / Nns, ix = (v % Nns) / (Ny + 1), iy = (v % Nns) % (Ny + 1);
      or_inc in[6];
      int k = or_incident(m, &g, iz, ix, iy, in);
      for (int q = 0; q < k; ++q)
        if (m->mark_edge[in[q].e] == 0 && m->mark_node[in[q].other] != m->mark_node[v]) mark = 1;
    }
    if (mark == 1) count++;
  }
  /* fill pass */
  int64_t leng = ndiel;
  double lxw = (m->xn[nx - 1] - m->xn[0]) / (nx - 1);
  double lyw = (m->yn[ny - 1] - m->yn[0]) / (ny - 1);
  double lzw = (m->zn[nz - 1] - m->zn[0]) / (nz - 1);
  for (int32_t c = 0; c < m->num_cdt; ++c) {
    if (m->cdt_mark_port[c] == -1) continue;
    for (int64_t v = 0; v < g.Nn; ++v) {
      if (m->mark_node[v] != c + 1) continue;
      int64_t iz = v / Nns, ix = (v % Nns) / (Ny + 1), iy = (v % Nns) % (Ny + 1);
      double lx = or_avg(m->xn, ix, nx), ly = or_avg(m->yn, iy, ny), lz = or_avg(m->zn, iz, nz);
      or_inc in[6];
      int k = or_incident(m, &g, iz, ix, iy, in);
      for (int q = 0; q < k; ++q) {
        if (!(m->mark_edge[in[q].e] == 0 && m->mark_node[in[q].other] != m->mark_node[v])) continue;
        double av;
        if (in[q].axis == 2) av = in[q].sgn > 0 ? lx * ly / (lxw * lyw * lzw) : -lx * ly / (lxw * lyw * lzw);
        else if (in[q].axis == 0) av = in[q].sgn > 0 ? ly * lz / (lxw * lyw * lzw) : -ly * lz / (lxw * lyw * lzw);
        else av = in[q].sgn > 0 ? lx * lz / (lxw * lyw * lzw) : -lx * lz / (lxw * lyw * lzw);
        or_push(&B->v0, in[q].e, leng, in[q].val, av);
        mark = 1;
      }
    }
    if (mark == 1) leng++;
  }
  B->cols = leng;
}

/* Galerkin accumulation shared by generateAd / generateAc: entries of one
 * column are consecutive, so row `col` of the product is accumulated over that
 * run in entry order (unordered_map += semantics), then sorted by column. */
typedef double (*or_weight_fn)(const gp_mesh *m, const or_grid *g, int64_t e, int64_t iz, int *use);

static double or_w_eps(const gp_mesh *m, const or_grid *g, int64_t e, int64_t iz, int *use) {
  (void)g; (void)e; *use = 1;
  return m->stack_epsn[iz];
}
static double or_w_sig(const gp_mesh *m, const or_grid *g, int64_t e, int64_t iz, int *use) {
  (void)g; *use = m->mark_edge[e] != 0;
  return m->stack_sign[iz];
}

static void or_galerkin(const gp_mesh *m, const or_basis *B, int with_eps0, int filter,
                        or_coo *out, int64_t *len_unfiltered) {
  or_grid g = or_make_grid(m);
  memset(out, 0, sizeof(*out));
  *len_unfiltered = 0;
  or_weight_fn wfn = with_eps0 ? or_w_eps : or_w_sig;
  int64_t i = 0, cap = 64;
  or_kv *buf = (or_kv *)malloc(sizeof(or_kv) * (size_t)cap);
  while (i < B->v0.n) {
    int64_t c = B->v0.col[i], j = i, nb = 0;
    for (; j < B->v0.n && B->v0.col[j] == c; ++j) {
      int64_t e = B->v0.row[j];
      int d; int64_t iz, r = e % g.El, ix, iy, n1, n2;
      iz = e / g.El;
      double h;
      if (r >= g.Es) { d = 2; ix = (r - g.Es) / (g.Ny + 1); iy = (r - g.Es) % (g.Ny + 1);
        n1 = iz * g.Nns + (g.Ny + 1) * ix + iy; n2 = (iz + 1) * g.Nns + (g.Ny + 1) * ix + iy; h = m->zn[iz + 1] - m->zn[iz]; }
      else if (r >= g.Ny * (g.Nx + 1)) { d = 1; ix = (r - g.Ny * (g.Nx + 1)) / (g.Ny + 1); iy = (r - g.Ny * (g.Nx + 1)) % (g.Ny + 1);
        n1 = iz * g.Nns + ix * (g.Ny + 1) + iy; n2 = iz * g.Nns + (ix + 1) * (g.Ny + 1) + iy; h = m->xn[ix + 1] - m->xn[ix]; }
      else { d = 0; ix = r / g.Ny; iy = r % g.Ny;
        n1 = iz * g.Nns + ix * (g.Ny + 1) + iy; n2 = iz * g.Nns + ix * (g.Ny + 1) + iy + 1; h = m->yn[iy + 1] - m->yn[iy]; }
      (void)d;
      int use;
      double w = wfn(m, &g, e, iz, &use);
      double av = B->v0.aval[j];
      if (nb + 2 > cap) { cap *= 2; buf = (or_kv *)realloc(buf, sizeof(or_kv) * (size_t)cap); }
      /* generateAd: v0d1aval * 1 / h * stackEpsn * EPSILON0 (src/fdtd.hpp:2782-2837)
       * generateAc: v0caval * 1 / h * stackSign, conductor edges only (src/fdtd.hpp:3489-3541) */
#define TERM(S) (with_eps0 ? av * (S) / h * w * OR_EPSILON0 : av * (S) / h * w)
      if (B->map[n1] != c + 1 && B->map[n1] != 0 && use) {
        buf[nb].key = B->map[n1] - 1; buf[nb].seq = nb; buf[nb].v = TERM(1); nb++;
        buf[nb].key = c; buf[nb].seq = nb; buf[nb].v = TERM(-1); nb++;
      } else if (B->map[n2] != c + 1 && B->map[n2] != 0 && use) {
        buf[nb].key = B->map[n2] - 1; buf[nb].seq = nb; buf[nb].v = TERM(-1); nb++;
        buf[nb].key = c; buf[nb].seq = nb; buf[nb].v = TERM(1); nb++;
      } else if (use) {
        buf[nb].key = c; buf[nb].seq = nb; buf[nb].v = fabs(TERM(1)); nb++;
      }
#undef TERM
    }
    qsort(buf, (size_t)nb, sizeof(or_kv), or_kv_cmp);
    for (int64_t a = 0; a < nb;) {
      int64_t b2 = a;
      double s = 0.;
      while (b2 < nb && buf[b2].key == buf[a].key) { s += buf[b2].v; ++b2; }
      (*len_unfiltered)++;
      if (!filter || fabs(s) > 1e-8) or_push(out, c, buf[a].key, s, 0.);
      a = b2;
    }
    i = j;
  }
  free(buf);
}

/* ------------------------------------------------------------------------- */
/* V0c basis and Ac (merge_v0c src/fdtd.hpp:2866-3479, generateAc :3481-3575) */
/* ------------------------------------------------------------------------- */
static void or_build_v0c(const gp_mesh *m, or_basis *B) {
  or_grid g = or_make_grid(m);
  memset(B, 0, sizeof(*B));
  B->map = (int64_t *)calloc((size_t)g.Nn, sizeof(int64_t));
  const int64_t nx = m->nx, ny = m->ny, nz = m->nz, Nns = g.Nns, Ny = g.Ny;
  int64_t *nodes = (int64_t *)malloc(sizeof(int64_t) * (size_t)g.Nn);
  int64_t col = 0;
  for (int32_t c = 0; c < m->num_cdt; ++c) {
    int64_t cnt = 0;
    for (int64_t v = 0; v < g.Nn; ++v) if (m->mark_node[v] == c + 1) nodes[cnt++] = v;
    int32_t mp = m->cdt_mark_port[c];
    /* floating and excited conductors omit their last node; the reference
     * conductor keeps all nodes but skips boundary nodes (src/fdtd.hpp:2900-2912, 3098-3104) */
    int64_t n = cnt - 1;
    if (mp <= -1) n = cnt;
    for (int64_t jc = 0; jc < n; ++jc) {
      int64_t v = nodes[jc];
      if (mp <= 0 && or_is_bd_node(m, v)) continue;
      B->map[v] = col + 1;
      int64_t iz = v / Nns, ix = (v % Nns) / (Ny + 1), iy = (v % Nns) % (Ny + 1);
      double lx = or_avg(m->xn, ix, nx), ly = or_avg(m->yn, iy, ny), lz = or_avg(m->zn, iz, nz);
      or_inc in[6];
      int k = or_incident(m, &g, iz, ix, iy, in);
      for (int q = 0; q < k; ++q) {
        double l = in[q].axis == 0 ? lx : in[q].axis == 1 ? ly : lz;
        or_push(&B->v0, in[q].e, col, in[q].val, in[q].sgn > 0 ? 1 / l : -1 / l);
      }
      col++;
    }
  }
  free(nodes);
  B->cols = col;
}

/* ---- exported V0 entry points -------------------------------------------- */
int or_v0_sizes(const gp_mesh *m, int64_t *sz /* [7] */) {
  or_basis d, c;
  or_coo ad, ac;
  int64_t lad, lac;
  or_build_v0d(m, &d);
  or_galerkin(m, &d, 1, 1, &ad, &lad);
  or_build_v0c(m, &c);
  or_galerkin(m, &c, 0, 0, &ac, &lac);
  sz[0] = d.v0.n; sz[1] = d.cols; sz[2] = ad.n; sz[3] = lad;
  sz[4] = c.v0.n; sz[5] = c.cols; sz[6] = ac.n;
  or_basis_free(&d); or_basis_free(&c); or_coo_free(&ad); or_coo_free(&ac);
  return 0;
}

int or_v0d(const gp_mesh *m, int64_t *row, int64_t *col, double *val, double *aval,
           int64_t *ad_row, int64_t *ad_col, double *ad_val) {
  or_basis d;
  or_coo ad;
  int64_t lad;
  or_build_v0d(m, &d);
  or_galerkin(m, &d, 1, 1, &ad, &lad);
  memcpy(row, d.v0.row, sizeof(int64_t) * (size_t)d.v0.n);
  memcpy(col, d.v0.col, sizeof(int64_t) * (size_t)d.v0.n);
  memcpy(val, d.v0.val, sizeof(double) * (size_t)d.v0.n);
  memcpy(aval, d.v0.aval, sizeof(double) * (size_t)d.v0.n);
  memcpy(ad_row, ad.row, sizeof(int64_t) * (size_t)ad.n);
  memcpy(ad_col, ad.col, sizeof(int64_t) * (size_t)ad.n);
  memcpy(ad_val, ad.val, sizeof(double) * (size_t)ad.n);
  or_basis_free(&d); or_coo_free(&ad);
  return 0;
}

int or_v0c(const gp_mesh *m, int64_t *row, int64_t *col, double *val, double *aval,
           int64_t *ac_row, int64_t *ac_col, double *ac_val) {
  or_basis c;
  or_coo ac;
  int64_t lac;
  or_build_v0c(m, &c);
  or_galerkin(m, &c, 0, 0, &ac, &lac);
  memcpy(row, c.v0.row, sizeof(int64_t) * (size_t)c.v0.n);
  memcpy(col, c.v0.col, sizeof(int64_t) * (size_t)c.v0.n);
  memcpy(val, c.v0.val, sizeof(double) * (size_t)c.v0.n);
  memcpy(aval, c.v0.aval, sizeof(double) * (size_t)c.v0.n);
  memcpy(ac_row, ac.row, sizeof(int64_t) * (size_t)ac.n);
  memcpy(ac_col, ac.col, sizeof(int64_t) * (size_t)ac.n);
  memcpy(ac_val, ac.val, sizeof(double) * (size_t)ac.n);
  or_basis_free(&c); or_coo_free(&ac);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* full-wave system pieces (reference(), src/generateStiff.cpp:568-636)        */
/* ------------------------------------------------------------------------- */
/* diagonal addition of A(w) per renumbered row: (2 pi f) * complex(-(2 pi f) eps_r eps0, sigma) */
int or_fullwave_diag(const gp_mesh *m, double freq, double *diag /* 2n interleaved */) {
  or_grid g = or_make_grid(m);
  int64_t *map = or_map_edge(m, &g);
  const double w = 2. * OR_PI * freq;
  for (int64_t e = 0; e < g.Ne; ++e) {
    if (map[e] < 0) continue;
    int64_t L = e / (g.Es + g.Ev);
    double re = -(w) * m->stack_epsn[L] * OR_EPSILON0;
    double im = m->mark_edge[e] != 0 ? m->stack_sign[L] : 0.;
    diag[2 * map[e]] = w * re;
    diag[2 * map[e] + 1] = w * im;
  }
  free(map);
  return 0;
}

/* J column of one port: J[mapEdge[e]] = 0. - 1i * dir * freq * 2. * M_PI (src/generateStiff.cpp:583-593) */
int or_rhs(const gp_mesh *m, double freq, int32_t port, double *b /* 2n, zeroed by the caller */) {
  or_grid g = or_make_grid(m);
  int64_t *map = or_map_edge(m, &g);
  for (int32_t s = m->port_side_ptr[port]; s < m->port_side_ptr[port + 1]; ++s)
    for (int64_t k = m->side_edge_ptr[s]; k < m->side_edge_ptr[s + 1]; ++k) {
      int64_t e = m->side_edges[k];
      double v = (double)m->side_dir[s] * freq * 2. * OR_PI;
      b[2 * map[e]] = 0.;          /* 0. - (+-0) */
      b[2 * map[e] + 1] = 0. - v;
    }
  free(map);
  return 0;
}

/* Construct_Z_V0_Vh (src/fdtd.hpp:4133-4170) for solution x of source port src */
int or_extract_z(const gp_mesh *m, const double *x, int32_t src, double *Zk) {
  or_grid g = or_make_grid(m);
  int64_t *map = or_map_edge(m, &g);
  const int32_t np = m->num_ports;
  double cur = 0.;
  for (int32_t s = m->port_side_ptr[src]; s < m->port_side_ptr[src + 1]; ++s) cur += m->side_area[s];
  for (int32_t p = 0; p < np; ++p) {
    double zr = Zk[2 * (p + np * src)], zi = Zk[2 * (p + np * src) + 1];
    int32_t s0 = m->port_side_ptr[p];
    for (int64_t k = m->side_edge_ptr[s0]; k < m->side_edge_ptr[s0 + 1]; ++k) {
      int64_t e = m->side_edges[k];
      int64_t iz = e / g.El, r = e % g.El;
      double leng;
      if (r >= g.Es) leng = m->zn[iz + 1] - m->zn[iz];
      else if (r >= g.Ny * (g.Nx + 1)) leng = m->xn[(r - g.Ny * (g.Nx + 1)) / (g.Ny + 1) + 1] - m->xn[(r - g.Ny * (g.Nx + 1)) / (g.Ny + 1)];
      else leng = m->yn[r % g.Ny + 1] - m->yn[r % g.Ny];
      double dir = m->side_dir[s0] * 1.0;
      /* complex<double> * double * double, then -= */
      double tr = x[2 * map[e]] * leng, ti = x[2 * map[e] + 1] * leng;
      tr = tr * dir; ti = ti * dir;
      zr -= tr; zi -= ti;
    }
    zr /= cur; zi /= cur;
    Zk[2 * (p + np * src)] = zr; Zk[2 * (p + np * src) + 1] = zi;
  }
  free(map);
  return 0;
}

/* freqNo2freq (src/fdtd.hpp:4319-4334) */
double or_freq(const gp_mesh *m, int32_t k) {
  if (m->nfreq == 1) return m->freq_start * m->freq_unit;
  if (m->freq_scale == 1) return (m->freq_start + k * (m->freq_end - m->freq_start) / (m->nfreq - 1)) * m->freq_unit;
  return m->freq_start * m->freq_unit * pow(m->freq_end / m->freq_start, (k * 1.0 / (m->nfreq - 1)));
}

/* ---- V0 quasi-static path (paraGenerator's V0 section) ------------------- */

/* J of one source port over original edges: J[e] = portDirection of the side,
 * assigned side by side (src/matrixCon.cpp:367-374); J zeroed by the caller */
int or_port_j(const gp_mesh *m, int32_t port, double *J) {
  for (int32_t s = m->port_side_ptr[port]; s < m->port_side_ptr[port + 1]; ++s)
    for (int64_t k = m->side_edge_ptr[s]; k < m->side_edge_ptr[s + 1]; ++k) J[m->side_edges[k]] = m->side_dir[s];
  return 0;
}

/* stackEpsn[e / (N_edge_s + N_edge_v)] * EPSILON0 per original edge
 * (src/matrixCon.cpp:443, 497) */
int or_edge_eps(const gp_mesh *m, double *eps) {
  or_grid g = or_make_grid(m);
  for (int64_t e = 0; e < g.Ne; ++e) eps[e] = m->stack_epsn[e / g.El] * OR_EPSILON0;
  return 0;
}

/* find_Vh's per-unknown coefficients (src/fdtd.hpp:3601-3605): for unknown
 * i = mapEdge[e], eps[i] = stackEpsn[L] * EPSILON0 and sig[i] = stackSign[L]
 * if markEdge[e] != 0 else 0, L = e / (N_edge_s + N_edge_v). */
int or_vh_coeffs(const gp_mesh *m, double *eps, double *sig) {
  or_grid g = or_make_grid(m);
  int64_t *map = or_map_edge(m, &g);
  for (int64_t e = 0; e < g.Ne; ++e) {
    if (map[e] < 0) continue;
    const int64_t L = e / (g.Es + g.Ev);
    eps[map[e]] = m->stack_epsn[L] * OR_EPSILON0;
    sig[map[e]] = m->mark_edge[e] != 0 ? m->stack_sign[L] : 0.;
  }
  free(map);
  return 0;
}

/* mapEdge / mapEdgeR (setMapEdge, src/fdtd.hpp:1508-1529) as arrays */
int or_edge_map(const gp_mesh *m, int64_t *map_out, int64_t *map_r) {
  or_grid g = or_make_grid(m);
  int64_t *map = or_map_edge(m, &g);
  for (int64_t e = 0; e < g.Ne; ++e) {
    map_out[e] = map[e];
    if (map[e] >= 0) map_r[map[e]] = e;
  }
  free(map);
  return 0;
}

/* The per-frequency Vh solve's coefficients (src/matrixCon.cpp:614-622):
 * eps[i] = stackEpsn[(e + N_edge_v) / (N_edge_s + N_edge_v)] * EPSILON0 and
 * sig[i] = SIGMA (5.8e7, src/fdtd.hpp:50) if markEdge[e] != 0, else 0, for
 * unknown i = mapEdge[e]. */
int or_vh_solve_coeffs(const gp_mesh *m, double *eps, double *sig) {
  or_grid g = or_make_grid(m);
  int64_t *map = or_map_edge(m, &g);
  for (int64_t e = 0; e < g.Ne; ++e) {
    if (map[e] < 0) continue;
    eps[map[e]] = m->stack_epsn[(e + g.Ev) / (g.Es + g.Ev)] * OR_EPSILON0;
    sig[map[e]] = m->mark_edge[e] != 0 ? 5.8e+7 : 0.;
  }
  free(map);
  return 0;
}

/* Construct_Z_V0 (src/fdtd.hpp:4178-4215): x over ORIGINAL edge ids (N_edge
 * complex, interleaved), accumulated into Z0[resp + np * src] */
int or_extract_z_v0(const gp_mesh *m, const double *x, int32_t src, double *Z0) {
  or_grid g = or_make_grid(m);
  const int32_t np = m->num_ports;
  double cur = 0.;
  for (int32_t s = m->port_side_ptr[src]; s < m->port_side_ptr[src + 1]; ++s) cur += m->side_area[s];
  for (int32_t p = 0; p < np; ++p) {
    double zr = Z0[2 * (p + np * src)], zi = Z0[2 * (p + np * src) + 1];
    int32_t s0 = m->port_side_ptr[p];
    for (int64_t k = m->side_edge_ptr[s0]; k < m->side_edge_ptr[s0 + 1]; ++k) {
      int64_t e = m->side_edges[k];
      int64_t iz = e / g.El, r = e % g.El;
      double leng;
      if (r >= g.Es) leng = m->zn[iz + 1] - m->zn[iz];
      else if (r >= g.Ny * (g.Nx + 1)) leng = m->xn[(r - g.Ny * (g.Nx + 1)) / (g.Ny + 1) + 1] - m->xn[(r - g.Ny * (g.Nx + 1)) / (g.Ny + 1)];
      else leng = m->yn[r % g.Ny + 1] - m->yn[r % g.Ny];
      double dir = m->side_dir[s0] * 1.0;
      double tr = x[2 * e] * leng, ti = x[2 * e + 1] * leng;
      tr = tr * dir; ti = ti * dir;
      zr -= tr; zi -= ti;
    }
    zr /= cur; zi /= cur;
    Z0[2 * (p + np * src)] = zr; Z0[2 * (p + np * src) + 1] = zi;
  }
  return 0;
}

/* print_z_V0's store (src/fdtd.hpp:4249-4316): slot 0 of X (nf * np * np
 * complex, interleaved) holds Z0; slot id >= 1 receives
 * Re Z0[j + i np] + i Im Z0[j + i np] f0 / f at index i + j np (transposed),
 * with f = freqEnd * unit for the last slot and the linear / log grid between */
int or_print_z_v0(const gp_mesh *m, double *X) {
  const int32_t np = m->num_ports, nf = m->nfreq;
  if (nf <= 1) return 0;
  for (int32_t id = 1; id < nf; ++id) {
    double f;
    if (id == nf - 1) f = m->freq_end * m->freq_unit;
    else if (m->freq_scale == 1) f = (m->freq_start + id * (m->freq_end - m->freq_start) / (nf - 1)) * m->freq_unit;
    else f = m->freq_start * m->freq_unit * pow(m->freq_end / m->freq_start, (id * 1.0 / (nf - 1)));
    for (int32_t i = 0; i < np; ++i)
      for (int32_t j = 0; j < np; ++j) {
        const double re = X[2 * (j + i * np)], im = X[2 * (j + i * np) + 1];
        /* x.real() + (1i) * x.imag() * freqStart * freqUnit / freq */
        const double zi = im * m->freq_start * m->freq_unit / f;
        const size_t o = (size_t)id * np * np + i + (size_t)j * np;
        X[2 * o] = re;
        X[2 * o + 1] = zi;
      }
  }
  return 0;
}

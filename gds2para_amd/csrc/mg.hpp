// mg.hpp -- geometric-algebraic multigrid for A(w) = S + diag(-w^2 eps + i w sigma).
//
// The V-cycle works on the symmetrized operator K(w) = W A(w) (W = edge dual
// volume, diag(W) S symmetric) and carries every iterate in split form
// x = u + G phi.  Because S G = 0, K G = W D(w) G is assembled directly (the
// B blocks) and residuals r - K u - B phi never cancel the huge gradient
// component against the curl-curl part.  Each level holds three real
// Galerkin blocks sharing one pattern:
//   K = [K_S, K_eps, K_sig]  -> K(w)  = K_S - w^2 K_eps + i w K_sig
//   B = [B_eps, B_sig]       -> B(w)  = -w^2 B_eps + i w B_sig      (= K(w) G)
//   N = [N_eps, N_sig]       -> N(w)  = -w^2 N_eps + i w N_sig      (= G^T K(w) G)
// built with the sort-based SpGEMM from the structured transfer operators.
// Smoothing is hybrid: l1-Jacobi on the edges, l1-Jacobi on the nodal
// auxiliary problem, both with per-column frequencies.  The coarsest level is
// solved exactly in FP64 (dense inverses of N_L(w) and K_L(w) per frequency).
#pragma once
#include <complex>
#include <vector>

#include "dense.hpp"
#include "sparse.hpp"
#include "transfer.hpp"

namespace gp {

// Workgroups go to the 8 XCDs round-robin (block b to XCD b % 8).  This
// renumbering gives each XCD one contiguous range of blocks, so the rows one
// XCD works on are neighbours in memory and their stencil gathers hit its L2.
__device__ __forceinline__ int64_t xcd_block_id() {
  const unsigned nb = gridDim.x, b = blockIdx.x;
  const unsigned per = nb >> 3, rem = nb & 7u, x = b & 7u, q = b >> 3;
  return (int64_t)(x * per + (x < rem ? x : rem) + q);
}

// KS consecutive slots of one row, with 16-byte vector loads (the row
// offset times 8 bytes must be a multiple of 16 for the float4 path: k and the
// first slot even)
template <class T, int KS>
__device__ __forceinline__ void load_slots(const C2<T> *__restrict__ p, C2<T> (&x)[KS]) {
  if constexpr (sizeof(T) == 4 && KS % 2 == 0) {
    const float4 *q = reinterpret_cast<const float4 *>(p);
#pragma unroll
    for (int h = 0; h < KS / 2; ++h) {
      const float4 v = q[h];
      x[2 * h] = {(T)v.x, (T)v.y};
      x[2 * h + 1] = {(T)v.z, (T)v.w};
    }
  } else if constexpr (sizeof(T) == 8) {
    const double2 *q = reinterpret_cast<const double2 *>(p);
#pragma unroll
    for (int h = 0; h < KS; ++h) {
      const double2 v = q[h];
      x[h] = {(T)v.x, (T)v.y};
    }
  } else {
#pragma unroll
    for (int h = 0; h < KS; ++h) x[h] = p[h];
  }
}

template <class T, int KS>
__device__ __forceinline__ void store_slots(C2<T> *__restrict__ p, const C2<T> (&x)[KS]) {
  if constexpr (sizeof(T) == 4 && KS % 2 == 0) {
    float4 *q = reinterpret_cast<float4 *>(p);
#pragma unroll
    for (int h = 0; h < KS / 2; ++h) q[h] = make_float4(x[2 * h].x, x[2 * h].y, x[2 * h + 1].x, x[2 * h + 1].y);
  } else {
#pragma unroll
    for (int h = 0; h < KS; ++h) p[h] = x[h];
  }
}

template <class T>
struct MatV {  // read-only view of a CSR in precision T
  int64_t nr = 0, nc = 0;
  const int64_t *rp = nullptr;
  const int32_t *ci = nullptr;
  const T *v[3] = {nullptr, nullptr, nullptr};
  int nv = 0;
  int32_t nlong = 0;                 // rows reduced block-wise beforehand (kLongRow)
  const int32_t *lrows = nullptr, *lidx = nullptr;
  int32_t nchunk = 0;
  const int64_t *lcb = nullptr, *lce = nullptr;
  const int32_t *lqp = nullptr;
  const T *dg[2] = {nullptr, nullptr};  // diagonal value blocks 1, 2 (then v[1], v[2] unused)
};

template <class T>
struct SellV {  // read-only view of a Sell in precision T
  int64_t nr = 0, nslice = 0, nent = 0;   // nent: stored entries incl. padding
  const int64_t *base = nullptr;
  const int32_t *w = nullptr;
  const int32_t *ci = nullptr;
  const T *v[3] = {nullptr, nullptr, nullptr};
  int nv = 0;
  const T *dg[2] = {nullptr, nullptr};  // diagonal value blocks 1, 2 (level 0)
  const T *dgs = nullptr;               // diagonal of block 0, kept out of the entries (level 0)
  const int32_t *rows = nullptr;        // slot position -> row (nullptr: identity)
};

struct MgLevel {
  int64_t ne = 0, nn = 0;
  Csr K, B, N, G, GT, P, PT, Pn, PnT;
  CsrF Kf, Bf, Nf, Gf, GTf, Pf, PTf, Pnf, PnTf;
  Sell Ks, Bs, Ns, GTs, Ps, PTs, Pns;    // SELL copies used by the V-cycle kernels
  SellF Ksf, Bsf, Nsf, GTsf, Psf, PTsf, Pnsf;
  double *kd[3] = {}, *kl[3] = {}, *nd[2] = {}, *nl[2] = {};
  float *kdf[3] = {}, *klf[3] = {}, *ndf[2] = {}, *nlf[2] = {};
};

struct AuxLevel {
  int64_t n = 0;
  Csr N, P, PT;                  // N = [N_eps, N_sig]; P to the next coarser level
  CsrF Nf, Pf, PTf;
  Sell Ns, Ps, PTs;              // SELL copies (long rows left to the long-row pass)
  SellF Nsf, Psf, PTsf;
  double *nd[2] = {}, *nl[2] = {};
  float *ndf[2] = {}, *nlf[2] = {};
};

// Blocks of the auxiliary coarsest level up to this many rows are inverted by
// one workgroup each (one launch for all of them); larger ones by the batched
// multi-launch Gauss-Jordan of dense.hip.
constexpr int32_t kBlockGJMax = kGJWorkgroupMax;

// Auxiliary nodal space [V0d | V0c] and its class-aware smoothed-aggregation hierarchy.
struct AuxAMG {
  int64_t naux = 0, ndiel = 0, nL = 0;
  Csr Gx, GxT, Bx;               // Gx: kept edges x naux; Bx = W [eps_r eps0, sigma] Gx
  CsrF Gxf, GxTf, Bxf;
  Sell GxTs, Bxs;
  SellF GxTsf, Bxsf;
  std::vector<AuxLevel> L;
  std::vector<double> NL[2];     // coarsest level, dense (host)
  double *NLd[2] = {};           // same, on the device
  // The coarsest level is block diagonal: V0d lives on dielectric edges and
  // V0c on conductor edges, and conductors share no edge, so the dielectric
  // aggregates couple only among themselves and each conductor's only within
  // that conductor.  Its connected components (of the pattern NL[0] | NL[1])
  // are found once; per frequency each block is inverted on its own.
  int32_t nblk = 0, max_blk = 0;
  int64_t inv_len = 0, aug_len = 0;        // sum of b^2 and of 2 b^2 over the blocks
  std::vector<int32_t> perm_h, boff_h;     // position -> coarse index (ascending within a block); offsets [nblk+1]
  std::vector<int32_t> small_h, big_h;     // blocks inverted by one workgroup each / by the batched launches
  int32_t *perm = nullptr, *rowblk = nullptr, *boff = nullptr, *small = nullptr;  // device copies
  int64_t *ioff = nullptr, *aoff = nullptr;  // per block: offset of its packed inverse / augmented matrix
  // Frequency-independent form of the coarsest inverse, when the pencil
  // NL[1] v = lambda NL[0] v could be solved (coarse_eig.hip):
  // N(w)^-1 = V diag(1 / (-w^2 + i w lambda)) V^T.  V and V^T row-major.
  bool eig = false;
  double *V = nullptr, *VT = nullptr, *lam = nullptr;
};

// Plane march of the level-0 edge smoother (k_edge_march).  The level-0 slot
// positions fall into blocks (16 x 16-cell tile, plane, half of the tile's ix
// columns), numbered b = (tile * nplane + plane) * 2 + half in position order,
// each a contiguous position range [start[b], start[b + 1]).  A workgroup
// walks one tile half up through the planes holding u of planes z - 1, z,
// z + 1 in LDS; code[e] tells, for each stored entry e of the level-0 K SELL
// copy, where its column sits: (plane - z + 1) << 16 | offset in that plane's
// block, or INT32_MIN | column for a column outside those three blocks (read
// from global memory).
struct EdgeMarch {
  int32_t *code = nullptr;
  int32_t *start = nullptr;      // [nblk + 1]
  int64_t ntile = 0, nblk = 0;
  int32_t nplane = 0, maxb = 0;  // planes; most rows in one block
  int32_t lds_max = 0;           // LDS one workgroup may take (0: the march is off)
  int64_t outside = 0;           // entries of real rows whose column is read from global memory
  int64_t entries = 0;           // entries of real rows
  bool tried = false;            // built (or found not to apply) once
};

struct Hierarchy {
  std::vector<MgLevel> L;
  EdgeMarch march;
  AuxAMG aux;
  // S G and S Gx of the rounded FP64 stiffness, formed exactly (double-double):
  // A (u + G phi + Gx a) = S u + SG phi + SGx a + D (u + G phi + Gx a)
  Csr SG, SGx;
  Sell Ss;                       // SELL copy of S for the outer operator
  Sell Ss1;                      // the same in the reference's row order, for single-column products
  bool has_f32 = false;
  // level-0 data for the FP64 outer operator and the V-cycle input scaling
  double *W = nullptr;        // edge dual volume
  float *Wf = nullptr;
  double *epsr = nullptr;     // stack_epsn of each kept edge's layer
  double *sig = nullptr;      // stack_sign if the edge is a conductor edge, else 0
  // the same two values as a 2-byte code per edge, layer * 2 + (conductor edge),
  // into per-layer tables (for kernels that stream the diagonal per row)
  int16_t *mcode = nullptr;
  double *tab_eps = nullptr, *tab_sig = nullptr;
  double *sdiag = nullptr;    // S_ii: the system diagonal is the FP64 sum S_ii + D_ii(w)
  // level-0 slot orders of the SELL copies whose rows are level-0 edges /
  // level-0 free nodes: (16 x 16-cell xy tile, plane, ix, iy[, direction])
  int32_t *ord_e = nullptr, *ord_n = nullptr;
  // coarsest level, dense on the host (row-major), per block
  int64_t neL = 0, nnL = 0;
  std::vector<double> KL[3], BL[2], NL[2], GTL;
  double *KLd[3] = {}, *BLd[2] = {}, *NLd[2] = {}, *GTLd = nullptr;  // device copies
  double setup_ms = 0;
};

// Coarsest-level operators per frequency, formed and inverted on the device
// once per distinct frequency.  tab holds device pointers
// [kinv | ninv | bl | ainv] x cap, indexed by an entry's slot.  A solve pins
// the entry of every device slot's frequency until that slot takes another
// column or the solve ends; unpinned entries are evicted least recently used
// first once the cache holds more than max_entries entries or max_bytes bytes
// (pinned entries are never evicted, so the bound is exceeded only while more
// distinct frequencies are in flight than it allows).
struct CoarseEntry {
  double omega = 0;
  void *mem = nullptr;           // one allocation: kinv, ninv, bl, ainv
  int64_t bytes = 0;
  int pins = 0;
  uint64_t used = 0;
};
struct CoarseCache {
  std::vector<CoarseEntry> e;    // index = slot; mem == nullptr marks a free slot
  const double **tab = nullptr;
  int cap = 0;
  int64_t bytes = 0;             // entries plus the pointer table
  uint64_t clock = 0;
  int max_entries = 64;
  int64_t max_bytes = (int64_t)2 << 30;
  int64_t evictions = 0;
  double form_ms = 0;            // host wall time spent forming and inverting entries
};

// Slots of the given angular frequencies, adding (and inverting) new ones;
// pin != 0 also pins each returned slot once.
gp_status coarse_cache_lookup(const Hierarchy &H, CoarseCache &C, const std::vector<double> &omegas,
                              std::vector<int32_t> *slots, hipStream_t s, int pin = 0);
void coarse_cache_unpin(CoarseCache &C, int32_t slot);
// The auxiliary coarsest level at angular frequency w as a dense matrix and
// its inverse assembled from the per-block inverses (n x n row-major, host);
// for checking the block inverse.
gp_status aux_coarse_dense(const Hierarchy &H, double w, std::vector<std::complex<double>> *mat,
                           std::vector<std::complex<double>> *inv, hipStream_t s);
// Device bytes of one entry of this hierarchy.
int64_t coarse_entry_bytes(const Hierarchy &H);
void coarse_cache_free(CoarseCache &C);

// y = A(w) x for k interleaved columns (row i, column j at i*k + j) at the
// angular frequencies om[k] (device), A(w) the matrix reference() hands
// PARDISO; one launch of the FP64 SELL product.  The transposition kernels
// convert between that layout and the host's column-major one (column j at
// j*n).
void amul(const Hierarchy &H, int64_t n, int k, const double *om, const C2<double> *x, C2<double> *y, hipStream_t s);
__global__ void k_cols_to_rows(int64_t n, int k, const C2<double> *__restrict__ a, C2<double> *__restrict__ b);
__global__ void k_rows_to_cols(int64_t n, int k, const C2<double> *__restrict__ a, C2<double> *__restrict__ b);

// Builds the hierarchy into A (the solver's arena, freed with the solver).
gp_status build_hierarchy(gp_layout *lay, Arena &A, Hierarchy *H, bool want_f32, int max_levels);
gp_status build_aux(gp_layout *lay, Arena &A, Hierarchy *H);
// Block structure of the auxiliary coarsest level (from its host dense copy).
gp_status aux_coarse_blocks(Arena &A, AuxAMG &X, hipStream_t s);
// Solves the auxiliary coarsest level's pencil once (sets X.eig on success).
gp_status aux_coarse_eigen(Arena &A, AuxAMG &X, hipStream_t s);
// x = N(w_j)^-1 r for k columns through the pencil's eigenvectors (X.eig);
// om[k] the columns' angular frequencies (device, FP64), yp an nL x k FP64
// work block.  Asynchronous on s.
template <class T>
gp_status aux_eig_apply(const AuxAMG &X, int k, const double *om, const C2<T> *r, C2<T> *x, C2<double> *yp,
                        hipStream_t s);
// FP32 copies for the mixed-precision V-cycle, added to an FP64 hierarchy.
gp_status hierarchy_add_f32(gp_layout *lay, Arena &A, Hierarchy *H, hipStream_t s);
// Block starts and entry codes of the level-0 plane march (EdgeMarch), built on
// the first solve that asks for it; leaves H->march.code null where it does not
// apply (rows out of block order, no tile order, too many rows).
gp_status build_edge_march(gp_layout *lay, Arena &A, Hierarchy *H, hipStream_t s);
gp_status aux_add_f32(Arena &A, Hierarchy *H, hipStream_t s);
gp_status aux_smoother_data(Arena &A, const Csr &M, double **d, double **l, hipStream_t s);
gp_status d2f_copy(const double *src, float *dst, int64_t n, hipStream_t s);

// Per-batch work space: per-slot frequencies and every level's vectors.
template <class T>
struct VWork {
  int k = 0;
  T *c1 = nullptr, *c2 = nullptr;     // per slot: -w^2 and w
  int32_t *fidx = nullptr;            // per slot: frequency slot in the CoarseCache
  const double *const *tab = nullptr; // CoarseCache::tab
  int cap = 0;
  const double *gtl = nullptr;        // coarsest G^T (nnL x neL), real
  std::vector<C2<T> *> r, u, phi, res, tmp, rn, dn, dn2;
  // auxiliary AMG: per-level rhs / iterate / scratch, level-0 output, saved input
  std::vector<C2<T> *> ar, ax, at;
  const double *om64 = nullptr;       // per slot: angular frequency in FP64 (device)
  C2<double> *eig_y = nullptr;        // coarsest auxiliary projection (eigen form)
  C2<T> *a = nullptr, *r0 = nullptr;
  C2<T> *lp = nullptr;                      // long-row products [nlong][nv][k]
  C2<T> *lpart = nullptr;                   // per-chunk partial products [nchunk][nv][k]
  int nu_aux = 2;
  bool aux_post = false;  // auxiliary correction after the edge V-cycle too (symmetric)
  bool u0_ready = false;  // level 0's first edge sweep already written (set by precond)
  int64_t sell_min_rows = 1;  // operators with fewer rows use the (row, slot) CSR kernels
  bool edge_march = false;    // level-0 K u gathers through the LDS plane march (FP32, k % 8 == 0)
  KTimer *kt = nullptr;               // times the level-0 edge smoother launches
  Arena arena;
};

template <class T>
gp_status vwork_create(gp_layout *lay, const Hierarchy &H, int k, VWork<T> *w);
// per-slot angular frequencies and CoarseCache slots (host arrays of length k)
template <class T>
gp_status vwork_set_columns(VWork<T> &w, const double *omega_cols, const int32_t *fslot_cols, const CoarseCache &C,
                            hipStream_t s);
template <class T>
void vwork_free(VWork<T> *w);

// Largest lanes-per-row the V-cycle's sliced-ELL kernels use on the calling
// thread (16 by default; the autotuner may lower it); returns the previous
// cap; cap < 1 only reads it.
int set_lanes_cap(int cap);

// One V-cycle: input w->r[0] (K-form residual, n x k), output w->u[0], w->phi[0].
template <class T>
gp_status vcycle(const Hierarchy &H, VWork<T> &w, hipStream_t s, int nu_pre, int nu_post);
// Full preconditioner: auxiliary-space AMG correction, edge V-cycle, auxiliary
// correction again.  Input w->r[0]; output (w->u[0], w->phi[0], w->a) with
// x = u + G phi + Gx a.  w->r[0] is left modified.
template <class T>
gp_status precond(const Hierarchy &H, VWork<T> &w, hipStream_t s, int nu_pre, int nu_post);

}  // namespace gp

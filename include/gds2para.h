/*
 * gds2para.h -- C ABI of the MI355X (gfx950) simulate-mode hot path.
 *
 * Plain pointers, fixed-width integers, no C++ or torch types.  Every entry
 * point returns a gp_status; nothing calls exit().  A gp_layout owns the
 * device-resident copies of one mesh, its stiffness matrix S, its V0 basis and
 * its multigrid hierarchy, so repeated solves on the same layout re-use them.
 *
 * Reference interfaces replaced (paths relative to the gds2Para tree):
 *   gp_assemble_stiffness   int generateStiff(fdtdMesh*)            src/fdtd.hpp:4448, src/generateStiff.cpp:6
 *                           + mklMatrixMulti_nt / COO2CSR_malloc    src/generateStiff.cpp:871-980
 *   gp_build_v0             fdtdMesh::merge_v0d1 / generateAd /
 *                           merge_v0c / generateAc                  src/fdtd.hpp:1823,2778,2866,3481
 *   gp_solve_fullwave       int reference(fdtdMesh*, int, myint*, myint*, double*)  src/generateStiff.cpp:568
 *                           void fdtdMesh::reference1(int, int, complex<double>*)   src/fdtd.hpp:3963
 *   gp_apply_fullwave       reference()'s A(w), applied             src/generateStiff.cpp:583-626
 *   gp_solve_real           int hypreSolve(fdtdMesh*, myint*, myint*, double*, myint,
 *                                          double*, myint, double*)  src/hypreSolver.h:14
 *   gp_solve_real_ex        hypreSolve with HYPRE_BoomerAMGSetRelaxType's choice  src/hypreSolve.cpp:230
 *   gp_solve_v0             paraGenerator's V0 section + Construct_Z_V0   src/matrixCon.cpp:366-560,
 *                                                                   src/fdtd.hpp:4178
 *   gp_v0_set_smoother      HYPRE_BoomerAMGSetRelaxType for the V0 solves  src/hypreSolve.cpp:230
 *   gp_z_v0_sweep           fdtdMesh::print_z_V0 (its store)        src/fdtd.hpp:4249
 *   gp_extract_z            fdtdMesh::Construct_Z_V0_Vh             src/fdtd.hpp:4133
 *   gp_find_vh              fdtdMesh::find_Vh(int k)                src/fdtd.hpp:3577
 *   gp_hessenberg_eigenvalues / _eigenvector
 *                           LAPACKE_dhseqr / LAPACKE_dhsein in find_Vh  src/fdtd.hpp:3754,3797
 *   gp_v0_vh_inputs         paraGenerator's u0 / u0a / yd           src/matrixCon.cpp:404-427,547-587
 *   gp_solve_vh             paraGenerator's per-frequency Vh solve  src/matrixCon.cpp:608-801
 *   gp_convert_param        Parasitics::convertParam                src/solnoutclass.hpp:1456
 *   gp_write_citi           SolverDataBase::dumpCITI                src/solnoutclass.hpp:6040
 *   gp_write_touchstone     SolverDataBase::dumpTouchstone          src/solnoutclass.hpp:6104
 *   gp_write_xyce           SolverDataBase::dumpXyce                src/solnoutclass.hpp:6032
 *   gp_write_spef           SolverDataBase::dumpSPEF, Parasitics::toSPEF  src/solnoutclass.hpp:6008,1972
 *   gp_freq_of              fdtdMesh::freqNo2freq                   src/fdtd.hpp:4319
 *   gp_edge_map             fdtdMesh::setMapEdge                    src/fdtd.hpp:1508
 */
#ifndef GDS2PARA_H
#define GDS2PARA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* status                                                                    */
/* ------------------------------------------------------------------------ */
typedef enum gp_status {
  GP_OK = 0,
  GP_ERR_ARG = 1,      /* malformed mesh or argument                         */
  GP_ERR_HIP = 2,      /* a HIP runtime call failed                          */
  GP_ERR_NOMEM = 3,    /* device or host allocation failed / over budget     */
  GP_ERR_NOCONV = 4,   /* solve stopped at max_iters above the tolerance     */
  GP_ERR_STATE = 5,    /* stage called before the stage it depends on        */
  GP_ERR_NODEVICE = 6  /* no gfx950 device visible                           */
} gp_status;

/* Human-readable text of the last error on this thread. */
const char *gp_last_error(void);

/* ------------------------------------------------------------------------ */
/* mesh contract (what meshAndMark / portSet leave in fdtdMesh)              */
/* ------------------------------------------------------------------------ */
/*
 * Numbering follows the reference (src/fdtd.hpp:4107-4130, src/mesh.cpp:395-410):
 *   Es = Ny(Nx+1) + Nx(Ny+1) surface edges per plane, Ev = (Nx+1)(Ny+1) vertical
 *   edges per layer, plane iz holds Es surface edges then (iz<Nz) Ev z-edges.
 *   y-edge (ix,iy):  iz(Es+Ev) + ix*Ny + iy
 *   x-edge (ix,iy):  iz(Es+Ev) + Ny(Nx+1) + ix(Ny+1) + iy
 *   z-edge (ix,iy):  iz(Es+Ev) + Es + ix(Ny+1) + iy
 *   node:            iz*Ev + ix(Ny+1) + iy
 * with Nx = nx-1 cells etc.  All arrays are host pointers owned by the caller.
 */
typedef struct gp_mesh {
  int64_t nx, ny, nz;             /* node counts per axis (N_cell_* + 1)          */
  const double *xn, *yn, *zn;     /* node coordinates (m), strictly increasing    */
  const double *stack_epsn;       /* [nz] relative permittivity by z-plane        */
  const double *stack_sign;       /* [nz] conductivity (S/m) by z-plane           */
  const int32_t *mark_edge;       /* [N_edge] conductor id (1-based) or 0         */
  const int32_t *mark_node;       /* [N_node] conductor id (1-based) or 0         */
  int32_t num_cdt;                /* number of isolated conductors                */
  const int32_t *cdt_mark_port;   /* [num_cdt] -1 reference, 0 floating, p+1 port */
  int64_t num_bd_edges;           /* |lbde u ubde|                                */
  const int64_t *bd_edges;        /* sorted ascending                             */
  int64_t num_bd_nodes;           /* |lbdn u ubdn|                                */
  const int64_t *bd_nodes;        /* sorted ascending                             */
  int32_t num_ports;
  const int32_t *port_side_ptr;   /* [num_ports+1] offsets into sides             */
  const int64_t *side_edge_ptr;   /* [num_sides+1] offsets into side_edges        */
  const int64_t *side_edges;      /* portEdge lists (original edge ids)           */
  const double *side_area;        /* [num_sides] portArea (m^2)                   */
  const int32_t *side_dir;        /* [num_sides] portDirection (+1/-1)            */
  const int32_t *side_cnd;        /* [num_sides] portCnd (1-based conductor id)   */
  double freq_unit, freq_start, freq_end;
  int32_t nfreq, freq_scale;      /* freq_scale 1 = linear, otherwise log         */
} gp_mesh;

typedef struct gp_mesh_sizes {
  int64_t n_edge, n_edge_s, n_edge_v, n_node, n_node_s, n_patch;
  int64_t n_unknowns;             /* N_edge - bden                                */
  int64_t nnz_s;                  /* nnz of S after PEC removal (closed form)     */
  int32_t num_sides;
} gp_mesh_sizes;

/* Closed-form sizes; pure host arithmetic, valid without a GPU. */
gp_status gp_mesh_sizes_of(const gp_mesh *mesh, gp_mesh_sizes *out);

/* Structural validation of a mesh (bounds, sortedness, boundary sets).  Host only. */
gp_status gp_mesh_validate(const gp_mesh *mesh);

/* mapEdge[N_edge] (new id or -1) and mapEdgeR[n] exactly as setMapEdge builds them. Host only. */
gp_status gp_edge_map(const gp_mesh *mesh, int64_t *map_edge, int64_t *map_edge_r);

/* freqNo2freq for the mesh's frequency grid. Host only. */
double gp_freq_of(const gp_mesh *mesh, int32_t freq_no);

/* ------------------------------------------------------------------------ */
/* synthetic layouts (stand-in for the GDSII/sim_input front end)            */
/* ------------------------------------------------------------------------ */
typedef enum gp_gen_kind {
  GP_GEN_MICROSTRIP = 0,          /* one strip over a ground plane, one z-port    */
  GP_GEN_MANHATTAN = 1,           /* random Manhattan metal layers with vias      */
  GP_GEN_PLATE = 2                /* plate over the ground plane, two dielectrics */
} gp_gen_kind;

typedef struct gp_gen_params {
  int32_t kind;                   /* gp_gen_kind                                  */
  int32_t nx_cells, ny_cells, nz_cells;
  double dx, dy, dz;              /* nominal cell sizes (m)                       */
  double grade;                   /* 0 = uniform; >0 = sinusoidal spacing ripple  */
  int32_t n_metal;                /* metal layers (MANHATTAN)                     */
  double fill;                    /* target area fill per metal layer             */
  int32_t wmin, wmax;             /* rectangle widths in cells                    */
  double via_density;             /* fraction of cells holding a via              */
  double eps_lo, eps_hi;          /* per-layer relative permittivity range        */
  double sigma;                   /* conductor conductivity (S/m)                 */
  int32_t n_ports;
  int32_t nfreq, freq_scale;
  double f_start, f_end;          /* Hz                                           */
  uint64_t seed;
} gp_gen_params;

typedef struct gp_owned_mesh gp_owned_mesh;   /* opaque; owns every array of .mesh */

gp_status gp_generate(const gp_gen_params *params, gp_owned_mesh **out);
const gp_mesh *gp_owned_mesh_view(const gp_owned_mesh *m);
void gp_owned_mesh_free(gp_owned_mesh *m);

/* ------------------------------------------------------------------------ */
/* device runtime                                                            */
/* ------------------------------------------------------------------------ */
typedef struct gp_device_info {
  int32_t device_count;
  char arch[64];                  /* gcnArchName of device 0                      */
  int64_t total_mem;              /* bytes                                        */
  int32_t cu_count;
} gp_device_info;

gp_status gp_device_query(gp_device_info *out);

/* ------------------------------------------------------------------------ */
/* layout handle: device-resident mesh, S, V0, hierarchy                     */
/* ------------------------------------------------------------------------ */
typedef struct gp_layout gp_layout;

/* Validates and uploads the mesh; builds mapEdge on the device. */
gp_status gp_layout_create(const gp_mesh *mesh, gp_layout **out);
void gp_layout_destroy(gp_layout *lay);

/* Device bytes currently held by the layout (matrices, hierarchy, the cached
 * solve workspace and coarse-level cache). */
int64_t gp_layout_device_bytes(const gp_layout *lay);

/* Releases the solve workspace and the per-frequency coarse-level cache that
 * solves keep with the layout (the matrices and hierarchy stay).  No
 * counterpart in the reference, which frees its PARDISO factors per call. */
gp_status gp_layout_trim(gp_layout *lay);

/* ---- stiffness (generateStiff + mklMatrixMulti_nt) ---------------------- */
/* Assembles S on the device: CSR, int64 row pointers, int32 columns in the
 * renumbered (mapEdge) index space, FP64 values v/MU.  Rows are the kept edges
 * in ascending order, columns ascending within a row. */
gp_status gp_assemble_stiffness(gp_layout *lay);
/* Host copy in CSR form; arrays sized n+1, nnz, nnz. */
gp_status gp_stiffness_copy_csr(const gp_layout *lay, int64_t *row_ptr, int32_t *col, double *val);
/* Host copy in the reference's COO layout (SRowId, SColId, Sval). */
gp_status gp_stiffness_copy_coo(const gp_layout *lay, int64_t *row, int64_t *col, double *val);

/* ---- V0 basis and nodal Laplacians ------------------------------------- */
typedef struct gp_v0_sizes {
  int64_t v0d_nnz, v0d_cols;      /* v0d1num, leng_v0d1                           */
  int64_t ad_nnz, ad_len;         /* kept entries, leng_Ad (incl. trailing zeros) */
  int64_t v0c_nnz, v0c_cols;      /* v0cnum, leng_v0c                             */
  int64_t ac_nnz;                 /* leng_Ac                                      */
} gp_v0_sizes;

gp_status gp_build_v0(gp_layout *lay);
gp_status gp_v0_sizes_get(const gp_layout *lay, gp_v0_sizes *out);
/* V0d in the reference's COO entry order: row = original edge id, col, val, aval
 * (before the sqrt(eps) scaling that paraGenerator applies). */
gp_status gp_v0d_copy(const gp_layout *lay, int64_t *row, int64_t *col, double *val, double *aval);
/* Ad COO sorted (row, col), |v| > 1e-8 entries only. */
gp_status gp_ad_copy(const gp_layout *lay, int64_t *row, int64_t *col, double *val);
gp_status gp_v0c_copy(const gp_layout *lay, int64_t *row, int64_t *col, double *val, double *aval);
gp_status gp_ac_copy(const gp_layout *lay, int64_t *row, int64_t *col, double *val);

/* ---- full-wave solve (reference / reference1) -------------------------- */
typedef enum gp_precision {
  GP_PREC_FP64 = 0,               /* FP64 preconditioner, single RHS per solve    */
  GP_PREC_MIXED = 1               /* FP32 hierarchy + V-cycle, FP64 outer FGMRES  */
} gp_precision;

typedef struct gp_solve_opts {
  double rtol;                    /* relative residual ||b-Ax||/||b||             */
  int32_t max_iters;              /* total Krylov iterations per column           */
  int32_t restart;                /* FGMRES restart length m                      */
  int32_t precision;              /* gp_precision                                 */
  int32_t batch;                  /* device slots (columns solved side by side);
                                     0 = autotuned with MIXED (gp_autotune, run on
                                     the layout's first such solve), 1 with FP64;
                                     1 = single RHS */
  int32_t verbose;
  int32_t smooth_edge;            /* edge/nodal smoothing sweeps per level (0: 1) */
  int32_t smooth_aux;             /* auxiliary AMG sweeps per level (0: 1)       */
  int64_t sell_min_rows;          /* operators with at least this many rows use the
                                     sliced-ELL kernels, smaller ones the CSR kernels
                                     (0: every operator, the default; a value above
                                     every operator's rows: CSR everywhere) */
  int32_t aux_post;               /* 1: auxiliary correction after the edge V-cycle
                                     as well (symmetric preconditioner); 0: before
                                     it only (the default, faster) */
  int32_t basis;                  /* Krylov basis (and Arnoldi vector) storage
                                     with MIXED: 0 = FP32 (the default;
                                     residuals, products, dot sums and the
                                     iterate stay FP64, every restart starts
                                     from the FP64 residual), 1 = FP64.  FP64
                                     precision always keeps an FP64 basis */
  int32_t graphs;                 /* 1 (default): the preconditioner of each
                                     Krylov step is captured once in a HIP graph
                                     and replayed; 0: launched directly.  Off
                                     while gp_kernel_timer records */
  int32_t edge_lds;               /* 1: with MIXED and a slot count that is a
                                     multiple of 8, the level-0 edge smoother's
                                     K u products read u from planes staged in
                                     LDS (same bits; its entry codes are built on
                                     the first such solve, 4 bytes per stored
                                     entry of level-0 K); 0 (default): u is
                                     gathered through the caches, faster at 16
                                     and 32 slots on MI355X */
} gp_solve_opts;

typedef struct gp_solve_stats {
  int32_t columns;                /* port x frequency columns solved              */
  int32_t converged;              /* columns that reached rtol                    */
  int32_t max_iters_used;
  int32_t batches;
  double max_rel_residual;        /* FP64 true residual, max over columns         */
  double setup_ms, solve_ms;      /* device wall time                             */
} gp_solve_stats;

/* Fills opts with the defaults (rtol 1e-8, restart 16, MIXED, batch 0). */
void gp_solve_opts_default(gp_solve_opts *opts);

/* Builds (or re-uses) the multigrid hierarchy for the layout. */
gp_status gp_prepare_solver(gp_layout *lay, int32_t precision);

/* Shape of the prepared hierarchy (no counterpart in the reference, whose
 * BoomerAMG setup is opaque; reported for diagnostics and benchmarks). */
#define GP_MAX_LEVELS 16
typedef struct {
  int32_t edge_levels;            /* geometric edge levels incl. the coarsest     */
  int64_t ne[GP_MAX_LEVELS], nn[GP_MAX_LEVELS], nnz_k[GP_MAX_LEVELS];
  int32_t aux_levels;             /* auxiliary [V0d|V0c] AMG levels               */
  int64_t na[GP_MAX_LEVELS], nnz_n[GP_MAX_LEVELS], nlong[GP_MAX_LEVELS];
  int64_t naux;                   /* columns of [V0d|V0c]                          */
  double setup_ms;
  int64_t coarse_entry_bytes;     /* device bytes of one frequency's coarse inverses */
  int32_t coarse_entries;         /* frequencies held in the coarse-level cache    */
  int64_t coarse_evictions;       /* entries evicted so far (least recently used)  */
  int64_t workspace_bytes;        /* cached solve workspace (0 if none)            */
  int32_t aux_coarse_blocks;      /* diagonal blocks of the auxiliary coarsest level */
  int32_t aux_coarse_max_block;   /* rows of its largest block                     */
  double coarse_form_ms;          /* time spent forming and inverting coarse entries */
  int32_t aux_coarse_eigen;       /* 1: the auxiliary coarsest level is solved through
                                     its pencil's eigenvectors (no per-frequency
                                     inverse); 0: per-frequency Gauss-Jordan       */
  int64_t graph_captures;         /* preconditioner graphs captured so far         */
  int64_t graph_replays;          /* preconditioner applications replayed from them */
  int32_t graph_failed;           /* 1: a capture failed and the solves launch directly */
  int32_t march_block_rows;       /* level-0 plane march (edge_lds): most rows in one
                                     (tile, plane, half) block; 0 = not built (no
                                     solve with edge_lds yet) or not applicable */
  int64_t march_entries;          /* stored level-0 K entries of real rows          */
  int64_t march_outside;          /* of those, read from global memory (outside the
                                     three staged planes of the row's tile half)   */
} gp_hierarchy_info;
gp_status gp_hierarchy_info_get(gp_layout *lay, gp_hierarchy_info *out);

/* Diagnostics (no counterpart in the reference): the auxiliary coarsest-level
 * matrix at freq (Hz) and the inverse the solve uses, assembled from its
 * per-block inverses or from its eigenvector form (O(n^3) host work: for
 * checks on small layouts); both n x n complex interleaved, row-major.  *n
 * receives the size; with mat and inv null nothing else is done. */
gp_status gp_coarse_aux_inverse(gp_layout *lay, double freq, int64_t *n, double *mat, double *inv);

/* Diagnostics: x = N(w)^-1 r at the auxiliary coarsest level through its
 * eigenvector form (the V-cycle's kernels; GP_ERR_STATE when the level has no
 * eigenvector form), for k <= 64 columns at one frequency; r and x are n x k
 * complex interleaved, row i column j at i*k + j.  precision GP_PREC_MIXED
 * passes r and x through FP32 as the mixed V-cycle does. */
gp_status gp_coarse_aux_apply(gp_layout *lay, double freq, int32_t k, int32_t precision, const double *r, double *x);

/* Autotuner (no counterpart in the reference): picks the mixed path's block
 * width among 8, 16 and 32 device slots and the lanes-per-row cap (16, 4 or
 * 2) of the V-cycle's sliced-ELL kernels by timing one FGMRES cycle of the
 * real solve on synthetic columns (restart length `restart`, 0 = 16) and
 * keeping the least time per column-iteration; widths whose workspace would
 * exceed half of the device memory are skipped.  Solves with batch 0 use the
 * choice.  slot_us / lanes_us: measured microseconds per column-iteration (-1
 * where skipped). */
typedef struct gp_autotune_info {
  int32_t slots;
  int32_t lanes_cap;
  double slot_us[3];              /* at 8, 16, 32 slots, lanes cap 16            */
  double lanes_us[3];             /* at the chosen width, caps 16, 4, 2          */
  double tune_ms;
} gp_autotune_info;
gp_status gp_autotune(gp_layout *lay, int32_t restart, gp_autotune_info *out);

/* Solves (S + diag(-w^2 eps + i w sigma)) x = -i w J for every (port, freq)
 * pair, exactly the system reference() hands to PARDISO.  freqs in Hz.
 * X (optional): complex interleaved (re,im), column c = f*np + p, n rows each.
 * Z (optional): nf*np*np complex, index k*np*np + resp + np*src, the layout
 * Construct_Z_V0_Vh writes into fdtdMesh::x. */
gp_status gp_solve_fullwave(gp_layout *lay, const double *freqs, int32_t nf,
                            const int32_t *ports, int32_t np,
                            const gp_solve_opts *opts,
                            double *X, double *Z, gp_solve_stats *stats);

/* Per-column outcome of a solve (no counterpart in the reference, which solves
 * every column with the same direct factorization). */
typedef struct gp_column_report {
  int32_t iters;                  /* Krylov iterations spent on the column       */
  int32_t converged;              /* 1 if the true residual reached rtol         */
  double rel_residual;            /* final FP64 ||b - A x|| / ||b||               */
  double start_ms;                /* when the column took a device slot          */
  double done_ms;                 /* completion time; both since the call start  */
} gp_column_report;

/* gp_solve_fullwave plus per-column reports (nf*np entries, column f*np + p,
 * may be NULL).  Columns are scheduled onto opts->batch device slots; a slot
 * whose column converges is refilled with the next pending column at the next
 * FGMRES restart (continuous batching), so every slot stays busy. */
gp_status gp_solve_fullwave_ex(gp_layout *lay, const double *freqs, int32_t nf,
                               const int32_t *ports, int32_t np,
                               const gp_solve_opts *opts,
                               double *X, double *Z, gp_solve_stats *stats,
                               gp_column_report *reports);

/* ---- continuous batching (serving) ---------------------------------------
 * No counterpart in the reference, which solves one frequency at a time with
 * a direct factorization.  gp_solve_stream keeps `slots` columns on the GPU
 * and, at every FGMRES restart, replaces each finished column with a new one
 * from the caller:
 *   pull(user, &col, block)  1: col filled; 0: nothing waiting; -1: closed.
 *                            block = 1 only for the first column while every
 *                            slot is idle (may wait); the rest of that refill
 *                            asks with block = 0.
 *   waiting(user)            columns ready to pull (steers early restarts).
 *   push(user, &col, z, rep) a finished column: z = its np complex Z entries
 *                            (Z[resp + np*col.port], resp = 0..np-1), or
 *                            null with rep->rel_residual < 0 for a rejected
 *                            column (port out of range, frequency <= 0).
 * Returns when pull() has returned -1 and every slot is idle.  The callbacks
 * run on the calling thread. */
/* Device bytes a solve with `slots` slots and restart length `restart` adds
 * on top of gp_layout_device_bytes of a layout that holds no solve workspace
 * yet: the Krylov bases, iterates and V-cycle vectors, plus the bound of the
 * per-frequency coarse-level cache (64 frequencies or 2 GiB, whichever is
 * less; entries in use by a running solve are never evicted).  A solve keeps
 * both with the layout for the next call (gp_layout_trim releases them);
 * prepares the solver first if needed.  For admission under a memory budget. */
gp_status gp_solve_workspace_bytes(gp_layout *lay, int32_t precision, int32_t slots, int32_t restart, int64_t *bytes);

typedef struct gp_stream_column {
  int32_t port;                   /* source port (0-based)                        */
  double freq;                    /* Hz                                           */
  int64_t tag;                    /* caller's id, echoed back                     */
} gp_stream_column;
typedef int32_t (*gp_pull_fn)(void *user, gp_stream_column *col, int32_t block);
typedef int64_t (*gp_waiting_fn)(void *user);
typedef void (*gp_push_fn)(void *user, const gp_stream_column *col, const double *z, const gp_column_report *rep);
gp_status gp_solve_stream(gp_layout *lay, const gp_solve_opts *opts, int32_t slots, gp_pull_fn pull,
                          gp_waiting_fn waiting, gp_push_fn push, void *user, gp_solve_stats *stats);

/* ---- request serving over resident layouts (gp_server) -------------------
 * No counterpart in the reference.  A server keeps layouts resident while
 * their bytes fit `budget_bytes` and runs one gp_solve_stream per resident
 * layout on a thread of its own, so columns of new requests join running
 * solves at the next FGMRES restart and layouts run side by side.  Admission
 * (opening a layout, evicting idle layouts least recently used first) runs on
 * a thread of its own; a layout larger than the budget opens only when nothing
 * else is resident.  Requests of one layout are queued in arrival order.
 *
 * The backend says how a layout key becomes resident.  Null: the library's
 * own -- key k is the k-th mesh given to gp_server_add_mesh, opened with
 * gp_layout_create, gp_assemble_stiffness and gp_prepare_solver(opts->precision),
 * sized by gp_layout_device_bytes + gp_solve_workspace_bytes, solved with
 * gp_solve_stream(opts, slots).  A caller backend supplies all six callbacks
 * (they run on the server's threads; `stream` must honour gp_solve_stream's
 * pull / waiting / push contract and pass `ctx` to them). */
typedef struct gp_server gp_server;
typedef struct gp_server_backend {
  void *user;
  int64_t (*estimate)(void *user, int32_t key);               /* bytes before opening */
  int32_t (*open)(void *user, int32_t key, void **handle);    /* gp_status */
  int64_t (*bytes)(void *user, void *handle);                 /* bytes once open, incl. workspace */
  int32_t (*num_ports)(void *user, int32_t key, void *handle);
  void (*close)(void *user, void *handle);
  int32_t (*stream)(void *user, void *handle, gp_pull_fn pull, gp_waiting_fn waiting, gp_push_fn push,
                    void *ctx);                                /* gp_status */
} gp_server_backend;
typedef struct gp_request_report {
  int32_t npt;                    /* ports of the layout (responses per column)    */
  int32_t columns;                /* ports x frequencies of the request            */
  int32_t converged;              /* columns that met rtol                         */
  int32_t rejected;               /* columns gp_solve_stream rejected (bad port / frequency) */
  int32_t max_iters;
  double max_rel_residual;
  double arrival_s, done_s;       /* seconds on the server's clock (since create)  */
} gp_request_report;
typedef struct gp_server_info {
  int64_t resident_bytes;         /* resident layouts plus the one being opened    */
  int32_t resident_layouts;
  int64_t evictions;
  int64_t backlog;                /* requests waiting for their layout             */
} gp_server_info;
/* opts: the library backend's solve options (null: gp_solve_opts_default). */
gp_status gp_server_create(const gp_server_backend *backend, const gp_solve_opts *opts, int32_t slots,
                           int64_t budget_bytes, gp_server **out);
/* Library backend only; the mesh must outlive the server. */
gp_status gp_server_add_mesh(gp_server *srv, const gp_mesh *mesh, int32_t *key);
/* Queues a request; returns at once with its id. */
gp_status gp_server_submit(gp_server *srv, int32_t key, const int32_t *ports, int32_t np, const double *freqs,
                           int32_t nf, int64_t *request);
/* Waits for one request (timeout_s < 0: no limit) and collects it: Z receives
 * nf*np columns of the layout's npt port responses, column (f, p) at
 * Z[2*((f*np + p)*npt + resp)] (complex interleaved; zeros for a rejected
 * column).  With Z null the request is waited for and reported (rep->npt
 * sizes Z) but not collected.  GP_ERR_STATE on a time-out, when a layout
 * failed to open or solve (the message names it), or when another thread
 * collected the request first. */
gp_status gp_server_wait(gp_server *srv, int64_t request, double timeout_s, double *Z, gp_request_report *rep);
gp_status gp_server_info_get(gp_server *srv, gp_server_info *out);
/* Stops admission, lets every resident layout finish its queued columns, and
 * closes them. */
void gp_server_destroy(gp_server *srv);

/* Construct_Z_V0_Vh for one solution vector x (n complex, renumbered space):
 * accumulates column src of the np*np block Zk (complex interleaved). Host only. */
gp_status gp_extract_z(const gp_mesh *mesh, const double *x, int32_t source_port, double *Zk);

/* ---- real nodal solve (hypreSolve) ------------------------------------- */
/* Solves A x = b for a real sparse A given in the reference's row-sorted COO,
 * FGMRES + multigrid on the device.  rtol 0 selects the reference's 1e-4. */
gp_status gp_solve_real(const int64_t *row, const int64_t *col, const double *val, int64_t nnz,
                        int64_t n, const double *b, double *x, double rtol,
                        int32_t *iters_out, double *relres_out);

/* Smoother of the real solve's multigrid, chosen per solve as the reference
 * chooses BoomerAMG's with HYPRE_BoomerAMGSetRelaxType (src/hypreSolve.cpp:230,
 * hybrid symmetric Gauss-Seidel there).  Both do two matrix products per side
 * of each level. */
typedef enum gp_real_smoother {
  GP_SMOOTH_L1_JACOBI = 0,        /* l1-Jacobi sweeps (gp_solve_real's)           */
  GP_SMOOTH_CHEBYSHEV = 1         /* Chebyshev steps on D^-1 A, interval [0.3 b, b],
                                     b = 1.1 x its power-iteration lambda_max    */
} gp_real_smoother;

/* gp_solve_real with the smoother given (gp_real_smoother; anything else is
 * GP_ERR_ARG).  Same interface as hypreSolve (src/hypreSolve.cpp:9) plus the
 * relaxation choice. */
gp_status gp_solve_real_ex(const int64_t *row, const int64_t *col, const double *val, int64_t nnz,
                           int64_t n, const double *b, double *x, double rtol, int32_t smoother,
                           int32_t *iters_out, double *relres_out);

/* ---- V0 quasi-static path (paraGenerator's V0 section) ------------------ */
typedef struct gp_v0_stats {
  int32_t solves;                 /* real solves run (3 per port, 2 without conductors) */
  int32_t max_iters;              /* largest FGMRES iteration count of those      */
  double max_rel_residual;        /* largest final relative residual             */
  double setup_ms;                /* V0 operators + Ad / Ac multigrid setup (once per layout) */
  double solve_ms;                /* this call's solves and Z extraction          */
  int32_t ad_levels, ac_levels;   /* multigrid levels of Ad and Ac                */
  int64_t ad_rows, ac_rows;       /* leng_v0d1, leng_v0c                          */
} gp_v0_stats;

/* paraGenerator's V0 section (src/matrixCon.cpp:366-560) for every source
 * port: three real solves (Ad, Ac, Ad; hypreSolve at :392, :511, :544) and
 * Construct_Z_V0 (src/fdtd.hpp:4178-4215) at freqStart.  Z0 receives the
 * np*np block (complex interleaved, index resp + np*src).  X, if non-null,
 * receives sys->x as print_z_V0 leaves it (nfreq*np*np complex, see
 * gp_z_v0_sweep).  rtol is the real solves' relative residual; 0 selects the
 * reference's 1e-4.  Builds the V0 basis if the layout has none; the V0
 * operators and the Ad / Ac hierarchies stay with the layout for later calls. */
gp_status gp_solve_v0(gp_layout *lay, double rtol, double *Z0, double *X, gp_v0_stats *stats);

/* The smoother (gp_real_smoother) of the V0 path's Ad and Ac multigrid, the
 * relaxation the reference sets with HYPRE_BoomerAMGSetRelaxType for those
 * hypreSolve calls (src/hypreSolve.cpp:230).  GP_SMOOTH_CHEBYSHEV until set
 * (fewer iterations than l1-Jacobi at the same products per level); a
 * hierarchy built with the other smoother is dropped and rebuilt by the next
 * gp_solve_v0 / gp_v0_vh_inputs. */
gp_status gp_v0_set_smoother(gp_layout *lay, int32_t smoother);

/* print_z_V0's store (src/fdtd.hpp:4249-4316), host only: X[0 .. np*np) holds
 * Z0 on entry; slot k >= 1 receives Re Z0[j + i np] + i Im Z0[j + i np] f0 / f_k
 * at index k np^2 + i + j np (the reference stores it transposed), with f_k
 * = freqEnd * unit for the last slot and the linear / log grid between. */
gp_status gp_z_v0_sweep(const gp_mesh *mesh, double *X);

/* ---- Vh high-frequency basis (fdtdMesh::find_Vh(k)) --------------------- */
typedef struct gp_vh_info {
  int32_t k;                      /* Arnoldi steps run                             */
  int32_t leng_vh;                /* Ritz values with |Im| > eps (= Vh's columns)  */
  double h_next;                  /* |w| after the last step (H's missing h_{k+1,k}) */
  double arnoldi_ms, eig_ms, vh_ms;
} gp_vh_info;

/* fdtdMesh::find_Vh(int k) (src/fdtd.hpp:3577-3859): k Arnoldi steps on the
 * 2n x 2n companion operator of S u + lambda D_sig u + lambda^2 D_eps u = 0,
 * scaled by `scale` (the reference uses 1e14), from the start vector
 * 1 / sqrt(2n); the eigenvalues of the k x k Hessenberg matrix (dhseqr, :3754);
 * eigenvectors of those with |Im| > eps (dhsein, :3797; the reference uses
 * eps = 1e-6); Vh = V_top * eigenvectors (:3836-3850).  Arnoldi runs on the
 * device (classical Gram-Schmidt, one re-orthogonalization pass); the k x k
 * eigenproblem on the host (Francis double-shift QR, inverse iteration).
 * Outputs, each optional (null skips it):
 *   hess  k*k, column-major (H[j*k + i] = h_ij, zero below the subdiagonal)
 *   ritz  2k doubles, the eigenvalues times scale (lambda of the pencil) as
 *         (re, im) pairs, conjugate pairs positive imaginary part first
 *   vh    n * max_cols complex (interleaved), column c at vh[2 n c ..], rows in
 *         the renumbered edge space; a conjugate pair's columns are y, conj(y)
 *         as the reference stores them, each y scaled so its largest entry is 1.
 * GP_ERR_ARG if Vh has more than max_cols columns (info->leng_vh still set);
 * max_cols 0 computes no vectors. */
gp_status gp_find_vh(gp_layout *lay, int32_t k, double scale, double eps, double *hess, double *ritz,
                     int64_t max_cols, double *vh, gp_vh_info *info);

/* The two host LAPACK calls of find_Vh, host only.  h is an n x n upper
 * Hessenberg matrix, column-major (entries below the subdiagonal ignored).
 * gp_hessenberg_eigenvalues: LAPACKE_dhseqr(job 'E', compz 'N')
 * (src/fdtd.hpp:3754): all eigenvalues as (re, im) pairs in w[2n], conjugate
 * pairs adjacent, positive imaginary part first.
 * gp_hessenberg_eigenvector: LAPACKE_dhsein(side 'R') (src/fdtd.hpp:3797) for
 * one eigenvalue re + i im: the right eigenvector in v[2n] (complex
 * interleaved), scaled so its largest entry is 1. */
gp_status gp_hessenberg_eigenvalues(int32_t n, const double *h, double *w);
gp_status gp_hessenberg_eigenvector(int32_t n, const double *h, double re, double im, double *v);

/* The V0 vectors paraGenerator's Vh section reads for one source port
 * (src/matrixCon.cpp:404-427, :547-587), from the same three real solves as
 * gp_solve_v0 (rtol 0 = the reference's 1e-4):
 *   u0, u0a  n x 2 real, column-major (u0[c*n + i]), rows by mapEdgeR:
 *            u0 = [V0d y0d, V0d y0d2 + V0c y0c], u0a = [V0da y0d, V0da y0d2 +
 *            V0ca y0c], each column divided by its norm over all N_edge entries
 *   yd       N_edge complex (interleaved), original edge order: the V0 solution
 *            (V0d y0d2 + V0c y0c) - i V0d y0d */
gp_status gp_v0_vh_inputs(gp_layout *lay, int32_t source_port, double rtol, double *u0, double *u0a, double *yd);

/* paraGenerator's per-frequency Vh solve (src/matrixCon.cpp:608-801) for one
 * source port: at each frequency, deflate u0 from Vh, solve the reduced
 * system Vh'^H (S + D) Vh' y = Vh'^H (J - D y_w), x = y_w + Vh' y with
 * y_w = Re yd + i Im yd f0 / f, then Construct_Z_V0_Vh(x) into Z.  D uses this
 * loop's own coefficients (eps from layer (e + N_edge_v) / (N_edge_s + N_edge_v),
 * sigma = 5.8e7 on conductor edges) and J = -i w on the port's edges.
 *   vh     n x leng_vh complex, column-major (as gp_find_vh returns it)
 *   u0/u0a/yd as gp_v0_vh_inputs returns them
 *   X      n x nf complex, or null (the reference's final_x per frequency)
 *   Z      nf blocks of np*np complex; column source_port of block k receives
 *          Construct_Z_V0_Vh as gp_extract_z computes it (zero Z first)
 * GP_ERR_STATE if a 2 x 2 deflation system or the reduced system is singular. */
gp_status gp_solve_vh(gp_layout *lay, const double *vh, int64_t leng_vh, const double *u0, const double *u0a,
                      const double *yd, int32_t source_port, const double *freqs, int32_t nf, double *X, double *Z);

/* ---- network-parameter output (host only) ------------------------------- */
/* Matrices: nf blocks of np x np complex (interleaved), row-major,
 * mat[(f*np + i)*np + j] = entry (i, j) at frequency f -- the layout
 * Parasitics::saveNetworkParam reads sys->x with (src/solnoutclass.hpp:1394),
 * so a Z array from the solves is passed unchanged. */

/* Parasitics::convertParam (src/solnoutclass.hpp:1456-1600): from / to are
 * 'Z', 'Y' or 'S'; zsource[np] are the ports' reference impedances (ohm),
 * needed for conversions to or from S.  in and out may not overlap. */
gp_status gp_convert_param(int32_t np, int32_t nf, char from, char to, const double *zsource,
                           const double *in, double *out);

/* SolverDataBase::dumpCITI (src/solnoutclass.hpp:6040-6100).  timestamp
 * replaces the reference's time stamp (formatted there from an uninitialised
 * time_t); null writes the current local time in the same format. */
gp_status gp_write_citi(const char *path, const char *design, const char *timestamp, char param,
                        int32_t np, int32_t nf, const double *freqs, const double *mat);

/* SolverDataBase::dumpTouchstone (src/solnoutclass.hpp:6104-6215), Touchstone
 * 2.0 with zsource[np] as the [Reference] line; timestamp as for gp_write_citi. */
gp_status gp_write_touchstone(const char *path, const char *design, const char *timestamp, char param,
                              int32_t np, int32_t nf, const double *freqs, const double *mat,
                              const double *zsource);

/* SolverDataBase::dumpXyce (src/solnoutclass.hpp:6032-6036): the SPICE
 * subcircuit that Parasitics::computeYBusFromParam(0) and toXyce
 * (:1680-1766, :2094-2248) write from the lowest frequency's block.  mat is
 * that np x np block (the layout of gp_convert_param, so Z[0..2 np^2) of a
 * solve), f0 its frequency (Hz), port_names[np] the node names, save_thresh
 * the reference's WRITE_THRESH (1e-5).  As in the reference: param 'Z' is
 * used as Y without inversion, 'Y' is inverted, C = -1 / (Im{Y} 2 pi f0)
 * entrywise, G = Re{Y}, both symmetrized unless already symmetric to 1e-12.
 * 'S' is rejected (the reference's branch leaves Y empty).  timestamp as for
 * gp_write_citi. */
gp_status gp_write_xyce(const char *path, const char *design, const char *timestamp, char param, int32_t np,
                        const char *const *port_names, double f0, const double *mat, double save_thresh);

/* Parasitics::toSPEF + SolverDataBase::dumpSPEF (src/solnoutclass.hpp:1972,
 * :6008): the elements of gp_write_xyce's reduction (same param, f0, mat and
 * save_thresh rules; negative capacitances stay capacitances) as one IEEE
 * 1481-1998 net "all" with *NAME_MAP, *PORTS, *CONN, *CAP and *RES sections.
 * port_dirs[np] holds 'I', 'O' or 'B' per port (null: all 'B').  The
 * reference prints through Parser-SPEF, which is absent here, so only the
 * record's content, not its spacing, follows the reference.  timestamp as for
 * gp_write_citi. */
gp_status gp_write_spef(const char *path, const char *design, const char *timestamp, char param, int32_t np,
                        const char *const *port_names, const char *port_dirs, double f0, const double *mat,
                        double save_thresh);

/* y = A(w) x for k columns (1 <= k <= 64) at freqs[k] (Hz), with A(w) the
 * complex matrix reference() hands PARDISO (src/generateStiff.cpp:583-626):
 * S + diag(w(-w eps_r EPSILON0) + i w sigma), the diagonal summed in FP64 as
 * the reference sums it.  x and y are host arrays of n*k complex
 * (interleaved), column j at offset j*n, in the renumbered (mapEdge) space.
 * Needs the stiffness; builds the solver hierarchy if the layout has none. */
gp_status gp_apply_fullwave(gp_layout *lay, const double *freqs, int32_t k, const double *x, double *y);

/* gp_apply_fullwave without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results (the same bits) as gp_apply_fullwave, but the
 * layout must already hold the stiffness and a prepared solver
 * (gp_assemble_stiffness, gp_prepare_solver; GP_ERR_STATE otherwise), and x,
 * y and the frequencies are staged in one static 96 MiB device buffer per GPU,
 * so n*k <= 2,097,146 (GP_ERR_ARG beyond).  Every call on one GPU uses that
 * GPU's buffer: calls that overlap on one GPU race and must be serialised by
 * the caller. */
gp_status gp_apply_fullwave_static(gp_layout *lay, const double *freqs, int32_t k, const double *x, double *y);

/* ---- measurement hooks -------------------------------------------------- */
/* Waits for all work the layout queued on its stream. */
gp_status gp_layout_sync(gp_layout *lay);

/* Times every launch of the level-0 edge smoother (the solve's dominant
 * kernel) with HIP events on the layout stream while enabled; enable = 1
 * also clears the record.  read returns the launch count, the summed device
 * time and the summed algorithmic bytes (each matrix and vector byte the
 * launch must move, counted once). */
gp_status gp_kernel_timer(gp_layout *lay, int32_t enable);
gp_status gp_kernel_timer_read(gp_layout *lay, int64_t *launches, double *total_ms, double *total_bytes);

/* Average device time (ms, HIP events on the library stream) of one complex
 * product y = A(w) x with k columns, the kernel of gp_apply_fullwave on
 * device-resident vectors.  precision is ignored (FP64). */
gp_status gp_bench_spmv(gp_layout *lay, int32_t k, int32_t precision, int32_t iters, double *avg_ms);

/* Average device time of a plain copy of `bytes` bytes (one read and one write
 * of each byte) by the kernel gp_copy_probe: the reference point that scales
 * the FETCH_SIZE / WRITE_SIZE counters into HBM bytes. */
gp_status gp_bench_copy(int64_t bytes, int32_t iters, double *avg_ms);

#ifdef __cplusplus
}
#endif

#endif /* GDS2PARA_H */

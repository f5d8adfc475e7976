"""ctypes mirror of include/gds2para.h.

Loads the in-tree ``libgds2para.so`` (built by ``gds2para_amd.build``).  The
product path never falls back to anything else: if the library is missing the
import of a compute function raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgds2para.so")

i32, i64, f64 = C.c_int32, C.c_int64, C.c_double
P = C.POINTER

GP_OK, GP_ERR_ARG, GP_ERR_HIP, GP_ERR_NOMEM, GP_ERR_NOCONV, GP_ERR_STATE, GP_ERR_NODEVICE = range(7)
GP_GEN_MICROSTRIP, GP_GEN_MANHATTAN, GP_GEN_PLATE = 0, 1, 2
GP_PREC_FP64, GP_PREC_MIXED = 0, 1
GP_SMOOTH_L1_JACOBI, GP_SMOOTH_CHEBYSHEV = 0, 1


class GpMesh(C.Structure):
    _fields_ = [
        ("nx", i64), ("ny", i64), ("nz", i64),
        ("xn", P(f64)), ("yn", P(f64)), ("zn", P(f64)),
        ("stack_epsn", P(f64)), ("stack_sign", P(f64)),
        ("mark_edge", P(i32)), ("mark_node", P(i32)),
        ("num_cdt", i32), ("cdt_mark_port", P(i32)),
        ("num_bd_edges", i64), ("bd_edges", P(i64)),
        ("num_bd_nodes", i64), ("bd_nodes", P(i64)),
        ("num_ports", i32), ("port_side_ptr", P(i32)),
        ("side_edge_ptr", P(i64)), ("side_edges", P(i64)),
        ("side_area", P(f64)), ("side_dir", P(i32)), ("side_cnd", P(i32)),
        ("freq_unit", f64), ("freq_start", f64), ("freq_end", f64),
        ("nfreq", i32), ("freq_scale", i32),
    ]


class GpMeshSizes(C.Structure):
    _fields_ = [
        ("n_edge", i64), ("n_edge_s", i64), ("n_edge_v", i64), ("n_node", i64),
        ("n_node_s", i64), ("n_patch", i64), ("n_unknowns", i64), ("nnz_s", i64),
        ("num_sides", i32),
    ]


class GpGenParams(C.Structure):
    _fields_ = [
        ("kind", i32), ("nx_cells", i32), ("ny_cells", i32), ("nz_cells", i32),
        ("dx", f64), ("dy", f64), ("dz", f64), ("grade", f64),
        ("n_metal", i32), ("fill", f64), ("wmin", i32), ("wmax", i32),
        ("via_density", f64), ("eps_lo", f64), ("eps_hi", f64), ("sigma", f64),
        ("n_ports", i32), ("nfreq", i32), ("freq_scale", i32),
        ("f_start", f64), ("f_end", f64), ("seed", C.c_uint64),
    ]


class GpDeviceInfo(C.Structure):
    _fields_ = [("device_count", i32), ("arch", C.c_char * 64), ("total_mem", i64), ("cu_count", i32)]


class GpV0Sizes(C.Structure):
    _fields_ = [
        ("v0d_nnz", i64), ("v0d_cols", i64), ("ad_nnz", i64), ("ad_len", i64),
        ("v0c_nnz", i64), ("v0c_cols", i64), ("ac_nnz", i64),
    ]


class GpSolveOpts(C.Structure):
    _fields_ = [
        ("rtol", f64), ("max_iters", i32), ("restart", i32), ("precision", i32),
        ("batch", i32), ("verbose", i32), ("smooth_edge", i32), ("smooth_aux", i32),
        ("sell_min_rows", i64), ("aux_post", i32), ("basis", i32), ("graphs", i32),
        ("edge_lds", i32),
    ]


class GpHierarchyInfo(C.Structure):
    _fields_ = [
        ("edge_levels", i32), ("ne", i64 * 16), ("nn", i64 * 16), ("nnz_k", i64 * 16),
        ("aux_levels", i32), ("na", i64 * 16), ("nnz_n", i64 * 16), ("nlong", i64 * 16),
        ("naux", i64), ("setup_ms", f64), ("coarse_entry_bytes", i64), ("coarse_entries", i32),
        ("coarse_evictions", i64), ("workspace_bytes", i64), ("aux_coarse_blocks", i32),
        ("aux_coarse_max_block", i32), ("coarse_form_ms", f64), ("aux_coarse_eigen", i32),
        ("graph_captures", i64), ("graph_replays", i64), ("graph_failed", i32),
        ("march_block_rows", i32), ("march_entries", i64), ("march_outside", i64),
    ]


class GpAutotuneInfo(C.Structure):
    _fields_ = [("slots", i32), ("lanes_cap", i32), ("slot_us", f64 * 3), ("lanes_us", f64 * 3), ("tune_ms", f64)]


class GpVhInfo(C.Structure):
    _fields_ = [("k", i32), ("leng_vh", i32), ("h_next", f64), ("arnoldi_ms", f64), ("eig_ms", f64), ("vh_ms", f64)]


class GpColumnReport(C.Structure):
    _fields_ = [("iters", i32), ("converged", i32), ("rel_residual", f64), ("start_ms", f64), ("done_ms", f64)]


class GpStreamColumn(C.Structure):
    _fields_ = [("port", i32), ("freq", f64), ("tag", i64)]


PULL_FN = C.CFUNCTYPE(i32, C.c_void_p, C.POINTER(GpStreamColumn), i32)
WAITING_FN = C.CFUNCTYPE(i64, C.c_void_p)
PUSH_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(GpStreamColumn), C.POINTER(C.c_double), C.POINTER(GpColumnReport))


class GpRequestReport(C.Structure):
    _fields_ = [("npt", i32), ("columns", i32), ("converged", i32), ("rejected", i32), ("max_iters", i32),
                ("max_rel_residual", f64), ("arrival_s", f64), ("done_s", f64)]


class GpServerInfo(C.Structure):
    _fields_ = [("resident_bytes", i64), ("resident_layouts", i32), ("evictions", i64), ("backlog", i64)]


SB_ESTIMATE = C.CFUNCTYPE(i64, C.c_void_p, i32)
SB_OPEN = C.CFUNCTYPE(i32, C.c_void_p, i32, C.POINTER(C.c_void_p))
SB_BYTES = C.CFUNCTYPE(i64, C.c_void_p, C.c_void_p)
SB_NUM_PORTS = C.CFUNCTYPE(i32, C.c_void_p, i32, C.c_void_p)
SB_CLOSE = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p)
SB_STREAM = C.CFUNCTYPE(i32, C.c_void_p, C.c_void_p, PULL_FN, WAITING_FN, PUSH_FN, C.c_void_p)


class GpServerBackend(C.Structure):
    _fields_ = [("user", C.c_void_p), ("estimate", SB_ESTIMATE), ("open", SB_OPEN), ("bytes", SB_BYTES),
                ("num_ports", SB_NUM_PORTS), ("close", SB_CLOSE), ("stream", SB_STREAM)]


class GpSolveStats(C.Structure):
    _fields_ = [
        ("columns", i32), ("converged", i32), ("max_iters_used", i32), ("batches", i32),
        ("max_rel_residual", f64), ("setup_ms", f64), ("solve_ms", f64),
    ]


class GpV0Stats(C.Structure):
    _fields_ = [
        ("solves", i32), ("max_iters", i32), ("max_rel_residual", f64), ("setup_ms", f64), ("solve_ms", f64),
        ("ad_levels", i32), ("ac_levels", i32), ("ad_rows", i64), ("ac_rows", i64),
    ]


# (name, restype, argtypes) for every entry point declared in include/gds2para.h
SIGNATURES = [
    ("gp_last_error", C.c_char_p, []),
    ("gp_mesh_sizes_of", i32, [P(GpMesh), P(GpMeshSizes)]),
    ("gp_mesh_validate", i32, [P(GpMesh)]),
    ("gp_edge_map", i32, [P(GpMesh), P(i64), P(i64)]),
    ("gp_freq_of", f64, [P(GpMesh), i32]),
    ("gp_generate", i32, [P(GpGenParams), P(C.c_void_p)]),
    ("gp_owned_mesh_view", P(GpMesh), [C.c_void_p]),
    ("gp_owned_mesh_free", None, [C.c_void_p]),
    ("gp_device_query", i32, [P(GpDeviceInfo)]),
    ("gp_layout_create", i32, [P(GpMesh), P(C.c_void_p)]),
    ("gp_layout_destroy", None, [C.c_void_p]),
    ("gp_layout_device_bytes", i64, [C.c_void_p]),
    ("gp_layout_trim", i32, [C.c_void_p]),
    ("gp_assemble_stiffness", i32, [C.c_void_p]),
    ("gp_stiffness_copy_csr", i32, [C.c_void_p, P(i64), P(i32), P(f64)]),
    ("gp_stiffness_copy_coo", i32, [C.c_void_p, P(i64), P(i64), P(f64)]),
    ("gp_build_v0", i32, [C.c_void_p]),
    ("gp_v0_sizes_get", i32, [C.c_void_p, P(GpV0Sizes)]),
    ("gp_v0d_copy", i32, [C.c_void_p, P(i64), P(i64), P(f64), P(f64)]),
    ("gp_ad_copy", i32, [C.c_void_p, P(i64), P(i64), P(f64)]),
    ("gp_v0c_copy", i32, [C.c_void_p, P(i64), P(i64), P(f64), P(f64)]),
    ("gp_ac_copy", i32, [C.c_void_p, P(i64), P(i64), P(f64)]),
    ("gp_solve_opts_default", None, [P(GpSolveOpts)]),
    ("gp_prepare_solver", i32, [C.c_void_p, i32]),
    ("gp_hierarchy_info_get", i32, [C.c_void_p, P(GpHierarchyInfo)]),
    ("gp_autotune", i32, [C.c_void_p, i32, P(GpAutotuneInfo)]),
    ("gp_coarse_aux_inverse", i32, [C.c_void_p, f64, P(i64), P(f64), P(f64)]),
    ("gp_coarse_aux_apply", i32, [C.c_void_p, f64, i32, i32, P(f64), P(f64)]),
    ("gp_solve_fullwave", i32, [C.c_void_p, P(f64), i32, P(i32), i32, P(GpSolveOpts), P(f64), P(f64), P(GpSolveStats)]),
    ("gp_solve_fullwave_ex", i32, [C.c_void_p, P(f64), i32, P(i32), i32, P(GpSolveOpts), P(f64), P(f64),
                                   P(GpSolveStats), P(GpColumnReport)]),
    ("gp_solve_workspace_bytes", i32, [C.c_void_p, i32, i32, i32, P(i64)]),
    ("gp_solve_stream", i32, [C.c_void_p, P(GpSolveOpts), i32, PULL_FN, WAITING_FN, PUSH_FN, C.c_void_p,
                              P(GpSolveStats)]),
    ("gp_server_create", i32, [P(GpServerBackend), P(GpSolveOpts), i32, i64, P(C.c_void_p)]),
    ("gp_server_add_mesh", i32, [C.c_void_p, P(GpMesh), P(i32)]),
    ("gp_server_submit", i32, [C.c_void_p, i32, P(i32), i32, P(f64), i32, P(i64)]),
    ("gp_server_wait", i32, [C.c_void_p, i64, f64, P(f64), P(GpRequestReport)]),
    ("gp_server_info_get", i32, [C.c_void_p, P(GpServerInfo)]),
    ("gp_server_destroy", None, [C.c_void_p]),
    ("gp_extract_z", i32, [P(GpMesh), P(f64), i32, P(f64)]),
    ("gp_solve_real", i32, [P(i64), P(i64), P(f64), i64, i64, P(f64), P(f64), f64, P(i32), P(f64)]),
    ("gp_solve_real_ex", i32, [P(i64), P(i64), P(f64), i64, i64, P(f64), P(f64), f64, i32, P(i32), P(f64)]),
    ("gp_solve_v0", i32, [C.c_void_p, f64, P(f64), P(f64), P(GpV0Stats)]),
    ("gp_v0_set_smoother", i32, [C.c_void_p, i32]),
    ("gp_z_v0_sweep", i32, [P(GpMesh), P(f64)]),
    ("gp_find_vh", i32, [C.c_void_p, i32, f64, f64, P(f64), P(f64), i64, P(f64), P(GpVhInfo)]),
    ("gp_hessenberg_eigenvalues", i32, [i32, P(f64), P(f64)]),
    ("gp_hessenberg_eigenvector", i32, [i32, P(f64), f64, f64, P(f64)]),
    ("gp_v0_vh_inputs", i32, [C.c_void_p, i32, f64, P(f64), P(f64), P(f64)]),
    ("gp_solve_vh", i32, [C.c_void_p, P(f64), i64, P(f64), P(f64), P(f64), i32, P(f64), i32, P(f64), P(f64)]),
    ("gp_convert_param", i32, [i32, i32, C.c_char, C.c_char, P(f64), P(f64), P(f64)]),
    ("gp_write_citi", i32, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char, i32, i32, P(f64), P(f64)]),
    ("gp_write_touchstone", i32, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char, i32, i32, P(f64), P(f64), P(f64)]),
    ("gp_write_xyce", i32, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char, i32, P(C.c_char_p), f64, P(f64), f64]),
    ("gp_write_spef", i32, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char, i32, P(C.c_char_p), C.c_char_p, f64,
                            P(f64), f64]),
    ("gp_apply_fullwave", i32, [C.c_void_p, P(f64), i32, P(f64), P(f64)]),
    ("gp_apply_fullwave_static", i32, [C.c_void_p, P(f64), i32, P(f64), P(f64)]),
    ("gp_bench_spmv", i32, [C.c_void_p, i32, i32, i32, P(f64)]),
    ("gp_layout_sync", i32, [C.c_void_p]),
    ("gp_bench_copy", i32, [i64, i32, P(f64)]),
    ("gp_kernel_timer", i32, [C.c_void_p, i32]),
    ("gp_kernel_timer_read", i32, [C.c_void_p, P(i64), P(f64), P(f64)]),
]

_lib = None


def load(path: str = LIB_PATH):
    """Load the HIP library once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libgds2para.so not built at {path}; run gds2para_amd.build.build()")
    lib = C.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class GpError(RuntimeError):
    def __init__(self, code, where):
        msg = load().gp_last_error().decode(errors="replace")
        super().__init__(f"{where} failed with status {code}: {msg}")
        self.code = code


def check(code, where):
    if code != GP_OK:
        raise GpError(code, where)

"""Device-resident layout: the Python face of gp_layout.

Mirrors the reference's call order on fdtdMesh (src/TestMain.cpp:343-370):
generateStiff -> V0 basis -> per-port / per-frequency solves -> Z.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi

i64p, i32p, f64p = C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_double)


def _p(a, t):
    return a.ctypes.data_as(t)


class Layout:
    def __init__(self, mesh):
        self.lib = capi.load()
        self.mesh = mesh
        h = C.c_void_p()
        capi.check(self.lib.gp_layout_create(mesh.ptr, C.byref(h)), "gp_layout_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.gp_layout_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_bytes(self):
        return int(self.lib.gp_layout_device_bytes(self.h))

    def trim(self):
        """Releases the cached solve workspace and coarse-level cache."""
        capi.check(self.lib.gp_layout_trim(self.h), "gp_layout_trim")

    # generateStiff --------------------------------------------------------
    def assemble(self):
        capi.check(self.lib.gp_assemble_stiffness(self.h), "gp_assemble_stiffness")

    def stiffness(self):
        """(row_ptr, col, val) of S on the host."""
        n, nnz = self.mesh.n, int(self.mesh.sizes.nnz_s)
        rp = np.empty(n + 1, np.int64)
        ci = np.empty(nnz, np.int32)
        v = np.empty(nnz)
        capi.check(self.lib.gp_stiffness_copy_csr(self.h, _p(rp, i64p), _p(ci, i32p), _p(v, f64p)),
                   "gp_stiffness_copy_csr")
        return rp, ci, v

    def stiffness_coo(self):
        nnz = int(self.mesh.sizes.nnz_s)
        r, c, v = np.empty(nnz, np.int64), np.empty(nnz, np.int64), np.empty(nnz)
        capi.check(self.lib.gp_stiffness_copy_coo(self.h, _p(r, i64p), _p(c, i64p), _p(v, f64p)),
                   "gp_stiffness_copy_coo")
        return r, c, v

    # V0 basis -------------------------------------------------------------
    def build_v0(self):
        capi.check(self.lib.gp_build_v0(self.h), "gp_build_v0")

    def v0_sizes(self):
        s = capi.GpV0Sizes()
        capi.check(self.lib.gp_v0_sizes_get(self.h, C.byref(s)), "gp_v0_sizes_get")
        return dict(v0d_nnz=s.v0d_nnz, v0d_cols=s.v0d_cols, ad_nnz=s.ad_nnz, ad_len=s.ad_len,
                    v0c_nnz=s.v0c_nnz, v0c_cols=s.v0c_cols, ac_nnz=s.ac_nnz)

    def v0d(self):
        s = self.v0_sizes()
        n = s["v0d_nnz"]
        r, c, v, a = np.empty(n, np.int64), np.empty(n, np.int64), np.empty(n), np.empty(n)
        capi.check(self.lib.gp_v0d_copy(self.h, _p(r, i64p), _p(c, i64p), _p(v, f64p), _p(a, f64p)), "gp_v0d_copy")
        m = s["ad_nnz"]
        ar, ac, av = np.empty(m, np.int64), np.empty(m, np.int64), np.empty(m)
        capi.check(self.lib.gp_ad_copy(self.h, _p(ar, i64p), _p(ac, i64p), _p(av, f64p)), "gp_ad_copy")
        return dict(row=r, col=c, val=v, aval=a, cols=s["v0d_cols"], ad_row=ar, ad_col=ac, ad_val=av,
                    ad_len=s["ad_len"])

    def v0c(self):
        s = self.v0_sizes()
        n = s["v0c_nnz"]
        r, c, v, a = np.empty(n, np.int64), np.empty(n, np.int64), np.empty(n), np.empty(n)
        capi.check(self.lib.gp_v0c_copy(self.h, _p(r, i64p), _p(c, i64p), _p(v, f64p), _p(a, f64p)), "gp_v0c_copy")
        m = s["ac_nnz"]
        ar, ac, av = np.empty(m, np.int64), np.empty(m, np.int64), np.empty(m)
        capi.check(self.lib.gp_ac_copy(self.h, _p(ar, i64p), _p(ac, i64p), _p(av, f64p)), "gp_ac_copy")
        return dict(row=r, col=c, val=v, aval=a, cols=s["v0c_cols"], ac_row=ar, ac_col=ac, ac_val=av)

    # full-wave solve --------------------------------------------------------
    def prepare(self, precision=capi.GP_PREC_MIXED):
        capi.check(self.lib.gp_prepare_solver(self.h, int(precision)), "gp_prepare_solver")

    def hierarchy_info(self):
        h = capi.GpHierarchyInfo()
        capi.check(self.lib.gp_hierarchy_info_get(self.h, C.byref(h)), "gp_hierarchy_info_get")
        el, al = h.edge_levels, h.aux_levels
        return dict(edge=[(h.ne[i], h.nn[i], h.nnz_k[i]) for i in range(el)],
                    aux=[(h.na[i], h.nnz_n[i], h.nlong[i]) for i in range(al)], naux=h.naux, setup_ms=h.setup_ms,
                    coarse_entry_bytes=h.coarse_entry_bytes, coarse_entries=h.coarse_entries,
                    coarse_evictions=h.coarse_evictions, workspace_bytes=h.workspace_bytes,
                    aux_coarse_blocks=h.aux_coarse_blocks, aux_coarse_max_block=h.aux_coarse_max_block,
                    coarse_form_ms=h.coarse_form_ms, aux_coarse_eigen=h.aux_coarse_eigen,
                    graph_captures=h.graph_captures, graph_replays=h.graph_replays, graph_failed=h.graph_failed,
                    march_block_rows=h.march_block_rows, march_entries=h.march_entries,
                    march_outside=h.march_outside)

    def autotune(self, restart=16):
        """gp_autotune: the mixed path's block width and lanes-per-row cap."""
        info = capi.GpAutotuneInfo()
        capi.check(self.lib.gp_autotune(self.h, int(restart), C.byref(info)), "gp_autotune")
        return dict(slots=info.slots, lanes_cap=info.lanes_cap, slot_us=list(info.slot_us),
                    lanes_us=list(info.lanes_us), tune_ms=info.tune_ms)

    def apply_fullwave(self, freqs, X):
        """Y = A(w) X (gp_apply_fullwave): X [k, n] complex, one column per
        frequency in freqs, in the renumbered space; returns Y [k, n]."""
        freqs = np.ascontiguousarray(np.atleast_1d(np.asarray(freqs, np.float64)))
        n = self.mesh.n
        X = np.asarray(X, np.complex128).reshape(len(freqs), n)
        xi = np.ascontiguousarray(np.stack([X.real, X.imag], -1).ravel())
        yi = np.zeros_like(xi)
        capi.check(self.lib.gp_apply_fullwave(self.h, _p(freqs, f64p), len(freqs), _p(xi, f64p), _p(yi, f64p)),
                   "gp_apply_fullwave")
        return (yi[0::2] + 1j * yi[1::2]).reshape(len(freqs), n)

    def set_v0_smoother(self, smoother):
        """gp_v0_set_smoother: GP_SMOOTH_CHEBYSHEV (the default) or
        GP_SMOOTH_L1_JACOBI for the V0 path's Ad / Ac multigrid; a hierarchy
        built with the other one is rebuilt by the next V0 call."""
        capi.check(self.lib.gp_v0_set_smoother(self.h, int(smoother)), "gp_v0_set_smoother")

    def solve_v0(self, rtol=0.0):
        """paraGenerator's V0 quasi-static solve (gp_solve_v0).

        Returns (Z0 [np*np] complex at freqStart, index resp + np*src;
        X [nfreq*np*np] complex as print_z_V0 leaves sys->x; stats).  rtol is
        the real solves' relative residual, 0 = the reference's 1e-4."""
        npt, nf = self.mesh.num_ports, max(int(self.mesh.c.nfreq), 1)
        z = np.zeros(2 * npt * npt)
        x = np.zeros(2 * nf * npt * npt)
        st = capi.GpV0Stats()
        capi.check(self.lib.gp_solve_v0(self.h, float(rtol), _p(z, f64p), _p(x, f64p), C.byref(st)), "gp_solve_v0")
        stats = {k: getattr(st, k) for k, _ in capi.GpV0Stats._fields_}
        return z[0::2] + 1j * z[1::2], x[0::2] + 1j * x[1::2], stats

    def find_vh(self, k, scale=1e14, eps=1e-6, max_cols=None):
        """fdtdMesh::find_Vh(k) (src/fdtd.hpp:3577-3859) through gp_find_vh.

        Returns (Vh [n, leng_Vh] complex in the renumbered space, Ritz values
        lambda [k] complex (eigenvalues of H times scale), H [k, k], info).
        max_cols defaults to k (every Ritz value could have |Im| > eps)."""
        n = self.mesh.n
        mc = int(k if max_cols is None else max_cols)
        hess = np.zeros(k * k)
        ritz = np.zeros(2 * k)
        vh = np.zeros(2 * n * max(mc, 1))
        info = capi.GpVhInfo()
        capi.check(self.lib.gp_find_vh(self.h, int(k), float(scale), float(eps), _p(hess, f64p), _p(ritz, f64p), mc,
                                       _p(vh, f64p), C.byref(info)), "gp_find_vh")
        cols = info.leng_vh if mc > 0 else 0
        V = (vh[0::2] + 1j * vh[1::2])[:n * cols].reshape(cols, n).T
        stats = {f: getattr(info, f) for f, _ in capi.GpVhInfo._fields_}
        return V, ritz[0::2] + 1j * ritz[1::2], hess.reshape(k, k).T, stats

    def v0_vh_inputs(self, port, rtol=0.0):
        """(u0 [n, 2], u0a [n, 2], yd [N_edge] complex) of paraGenerator's Vh
        section for one source port (gp_v0_vh_inputs)."""
        n, ne = self.mesh.n, self.mesh.n_edge
        u0, u0a, yd = np.zeros(2 * n), np.zeros(2 * n), np.zeros(2 * ne)
        capi.check(self.lib.gp_v0_vh_inputs(self.h, int(port), float(rtol), _p(u0, f64p), _p(u0a, f64p),
                                            _p(yd, f64p)), "gp_v0_vh_inputs")
        return u0.reshape(2, n).T, u0a.reshape(2, n).T, yd[0::2] + 1j * yd[1::2]

    def solve_vh(self, Vh, u0, u0a, yd, port, freqs):
        """paraGenerator's per-frequency Vh solve (gp_solve_vh): returns
        (X [nf, n] complex, Z [nf, np*np] complex with column `port` filled)."""
        n, npt = self.mesh.n, self.mesh.num_ports
        freqs = np.ascontiguousarray(np.atleast_1d(np.asarray(freqs, np.float64)))
        nf, L = len(freqs), Vh.shape[1]
        vh = np.ascontiguousarray(np.stack([Vh.T.real, Vh.T.imag], -1).ravel())
        u = np.ascontiguousarray(np.asarray(u0, np.float64).T.ravel())
        ua = np.ascontiguousarray(np.asarray(u0a, np.float64).T.ravel())
        y = np.ascontiguousarray(np.stack([yd.real, yd.imag], -1).ravel())
        X = np.zeros(2 * n * nf)
        Z = np.zeros(2 * nf * npt * npt)
        capi.check(self.lib.gp_solve_vh(self.h, _p(vh, f64p), L, _p(u, f64p), _p(ua, f64p), _p(y, f64p), int(port),
                                        _p(freqs, f64p), nf, _p(X, f64p), _p(Z, f64p)), "gp_solve_vh")
        return (X[0::2] + 1j * X[1::2]).reshape(nf, n), (Z[0::2] + 1j * Z[1::2]).reshape(nf, npt * npt)

    def solve(self, freqs, ports=None, precision=capi.GP_PREC_MIXED, rtol=1e-8, batch=0, restart=16,
              max_iters=2000, want_x=False, verbose=0, smooth_edge=0, smooth_aux=0, sell_min_rows=0,
              aux_post=0, basis=0, graphs=1, edge_lds=0):
        """Returns (Z[nf, np*np] complex, stats, X or None)."""
        ports = list(range(self.mesh.num_ports)) if ports is None else list(ports)
        freqs = np.ascontiguousarray(np.asarray(freqs, np.float64))
        pr = np.ascontiguousarray(np.asarray(ports, np.int32))
        nf, npp, npt = len(freqs), len(pr), self.mesh.num_ports
        o = capi.GpSolveOpts()
        self.lib.gp_solve_opts_default(C.byref(o))
        o.rtol, o.precision, o.batch, o.restart, o.max_iters, o.verbose = rtol, precision, batch, restart, max_iters, verbose
        o.smooth_edge, o.smooth_aux, o.sell_min_rows = smooth_edge, smooth_aux, sell_min_rows
        o.aux_post, o.basis, o.graphs, o.edge_lds = aux_post, basis, graphs, edge_lds
        Z = np.zeros(2 * nf * npt * npt)
        X = np.zeros(2 * self.mesh.n * nf * npp) if want_x else None
        st = capi.GpSolveStats()
        reps = (capi.GpColumnReport * (nf * npp))()
        code = self.lib.gp_solve_fullwave_ex(self.h, _p(freqs, f64p), nf, _p(pr, i32p), npp, C.byref(o),
                                             _p(X, f64p) if want_x else None, _p(Z, f64p), C.byref(st), reps)
        if code not in (capi.GP_OK, capi.GP_ERR_NOCONV):
            capi.check(code, "gp_solve_fullwave")
        Zc = (Z[0::2] + 1j * Z[1::2]).reshape(nf, npt * npt)
        Xc = None
        if want_x:
            Xc = (X[0::2] + 1j * X[1::2]).reshape(nf * npp, self.mesh.n)
        stats = dict(columns=st.columns, converged=st.converged, max_iters=st.max_iters_used, batches=st.batches,
                     max_rel_residual=st.max_rel_residual, setup_ms=st.setup_ms, solve_ms=st.solve_ms,
                     status=code,
                     iters=[r.iters for r in reps], rel_residual=[r.rel_residual for r in reps],
                     start_ms=[r.start_ms for r in reps], done_ms=[r.done_ms for r in reps])
        return Zc, stats, Xc

    def sync(self):
        capi.check(self.lib.gp_layout_sync(self.h), "gp_layout_sync")

    def kernel_timer(self, enable=True):
        capi.check(self.lib.gp_kernel_timer(self.h, 1 if enable else 0), "gp_kernel_timer")

    def kernel_timer_read(self):
        n, ms, b = C.c_int64(), C.c_double(), C.c_double()
        capi.check(self.lib.gp_kernel_timer_read(self.h, C.byref(n), C.byref(ms), C.byref(b)), "gp_kernel_timer_read")
        return n.value, ms.value, b.value

    def bench_spmv(self, k=1, precision=capi.GP_PREC_FP64, iters=20):
        ms = C.c_double()
        capi.check(self.lib.gp_bench_spmv(self.h, int(k), int(precision), int(iters), C.byref(ms)), "gp_bench_spmv")
        return ms.value


def z_v0_sweep(mesh, Z0):
    """print_z_V0's store on the host (gp_z_v0_sweep): nfreq*np*np complex."""
    lib = capi.load()
    npt, nf = mesh.num_ports, max(int(mesh.c.nfreq), 1)
    x = np.zeros(2 * nf * npt * npt)
    x[:2 * npt * npt] = np.stack([np.real(Z0), np.imag(Z0)], 1).ravel()
    capi.check(lib.gp_z_v0_sweep(mesh.ptr, _p(x, f64p)), "gp_z_v0_sweep")
    return x[0::2] + 1j * x[1::2]


def solve_real(row, col, val, n, b, rtol=0.0, smoother=None):
    """hypreSolve on the device (gp_solve_real): real A x = b, A as COO triplets.

    rtol = 0 selects the reference's 1e-4.  smoother (GP_SMOOTH_L1_JACOBI or
    GP_SMOOTH_CHEBYSHEV) picks the multigrid's relaxation through
    gp_solve_real_ex; None keeps gp_solve_real's l1-Jacobi.  Returns (x,
    iterations, relative residual); raises GpError unless the tolerance is
    met."""
    lib = capi.load()
    row = np.ascontiguousarray(row, np.int64)
    col = np.ascontiguousarray(col, np.int64)
    val = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(int(n))
    it, rr = C.c_int32(), C.c_double()
    if smoother is None:
        capi.check(lib.gp_solve_real(_p(row, i64p), _p(col, i64p), _p(val, f64p), len(val), int(n), _p(b, f64p),
                                     _p(x, f64p), float(rtol), C.byref(it), C.byref(rr)), "gp_solve_real")
    else:
        capi.check(lib.gp_solve_real_ex(_p(row, i64p), _p(col, i64p), _p(val, f64p), len(val), int(n),
                                        _p(b, f64p), _p(x, f64p), float(rtol), int(smoother), C.byref(it),
                                        C.byref(rr)), "gp_solve_real_ex")
    return x, it.value, rr.value


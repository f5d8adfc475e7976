"""Python face of the CPU oracle (test infrastructure only).

Wraps oracle/_build/libgp_oracle.so (the C restatement in gp_oracle.c) and adds
the FP64 direct solve that stands in for reference()'s MKL PARDISO call
(src/generateStiff.cpp:638-672): SciPy's SuperLU on the same complex CSR
matrix S + diag(-w^2 eps + i w sigma) and the same right-hand side J.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "libgp_oracle.so")
i64p, i32p, f64p = C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_double)
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            import subprocess
            subprocess.run(["make", "-C", _HERE], check=True, capture_output=True)
        lib = C.CDLL(LIB)
        lib.or_se_sh_count.restype = C.c_int64
        lib.or_se_sh_count.argtypes = [C.c_void_p]
        lib.or_se_sh.argtypes = [C.c_void_p, i64p, i64p, f64p, f64p]
        lib.or_stiffness_nnz.restype = C.c_int64
        lib.or_stiffness_nnz.argtypes = [C.c_void_p]
        lib.or_stiffness.argtypes = [C.c_void_p, i64p, i32p, f64p]
        lib.or_v0_sizes.argtypes = [C.c_void_p, i64p]
        lib.or_v0d.argtypes = [C.c_void_p, i64p, i64p, f64p, f64p, i64p, i64p, f64p]
        lib.or_v0c.argtypes = [C.c_void_p, i64p, i64p, f64p, f64p, i64p, i64p, f64p]
        lib.or_fullwave_diag.argtypes = [C.c_void_p, C.c_double, f64p]
        lib.or_rhs.argtypes = [C.c_void_p, C.c_double, C.c_int32, f64p]
        lib.or_extract_z.argtypes = [C.c_void_p, f64p, C.c_int32, f64p]
        lib.or_freq.restype = C.c_double
        lib.or_freq.argtypes = [C.c_void_p, C.c_int32]
        lib.or_port_j.argtypes = [C.c_void_p, C.c_int32, f64p]
        lib.or_edge_eps.argtypes = [C.c_void_p, f64p]
        lib.or_vh_coeffs.argtypes = [C.c_void_p, f64p, f64p]
        lib.or_edge_map.argtypes = [C.c_void_p, i64p, i64p]
        lib.or_vh_solve_coeffs.argtypes = [C.c_void_p, f64p, f64p]
        lib.or_extract_z_v0.argtypes = [C.c_void_p, f64p, C.c_int32, f64p]
        lib.or_print_z_v0.argtypes = [C.c_void_p, f64p]
        _lib = lib
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def se_sh(mesh):
    """Raw Se/Sh incidence triplets in the reference's emission order."""
    lib = _load()
    n = lib.or_se_sh_count(mesh.ptr)
    row, col = np.empty(n, np.int64), np.empty(n, np.int64)
    se, sh = np.empty(n), np.empty(n)
    lib.or_se_sh(mesh.ptr, _p(row, i64p), _p(col, i64p), _p(se, f64p), _p(sh, f64p))
    return row, col, se, sh


def stiffness(mesh):
    """S (renumbered, PEC removed) as (row_ptr int64, col int32, val float64)."""
    lib = _load()
    nnz = lib.or_stiffness_nnz(mesh.ptr)
    rp = np.empty(mesh.n + 1, np.int64)
    ci = np.empty(nnz, np.int32)
    v = np.empty(nnz)
    lib.or_stiffness(mesh.ptr, _p(rp, i64p), _p(ci, i32p), _p(v, f64p))
    return rp, ci, v


def stiffness_csr(mesh):
    rp, ci, v = stiffness(mesh)
    return sp.csr_matrix((v, ci, rp), shape=(mesh.n, mesh.n))


def v0_sizes(mesh):
    sz = np.zeros(7, np.int64)
    _load().or_v0_sizes(mesh.ptr, _p(sz, i64p))
    return dict(v0d_nnz=int(sz[0]), v0d_cols=int(sz[1]), ad_nnz=int(sz[2]), ad_len=int(sz[3]),
                v0c_nnz=int(sz[4]), v0c_cols=int(sz[5]), ac_nnz=int(sz[6]))


def v0d(mesh):
    s = v0_sizes(mesh)
    r, c = np.empty(s["v0d_nnz"], np.int64), np.empty(s["v0d_nnz"], np.int64)
    v, a = np.empty(s["v0d_nnz"]), np.empty(s["v0d_nnz"])
    ar, ac = np.empty(s["ad_nnz"], np.int64), np.empty(s["ad_nnz"], np.int64)
    av = np.empty(s["ad_nnz"])
    _load().or_v0d(mesh.ptr, _p(r, i64p), _p(c, i64p), _p(v, f64p), _p(a, f64p),
                   _p(ar, i64p), _p(ac, i64p), _p(av, f64p))
    return dict(row=r, col=c, val=v, aval=a, cols=s["v0d_cols"], ad_row=ar, ad_col=ac, ad_val=av,
                ad_len=s["ad_len"])


def v0c(mesh):
    s = v0_sizes(mesh)
    r, c = np.empty(s["v0c_nnz"], np.int64), np.empty(s["v0c_nnz"], np.int64)
    v, a = np.empty(s["v0c_nnz"]), np.empty(s["v0c_nnz"])
    ar, ac = np.empty(s["ac_nnz"], np.int64), np.empty(s["ac_nnz"], np.int64)
    av = np.empty(s["ac_nnz"])
    _load().or_v0c(mesh.ptr, _p(r, i64p), _p(c, i64p), _p(v, f64p), _p(a, f64p),
                   _p(ar, i64p), _p(ac, i64p), _p(av, f64p))
    return dict(row=r, col=c, val=v, aval=a, cols=s["v0c_cols"], ac_row=ar, ac_col=ac, ac_val=av)


def fullwave_diag(mesh, freq):
    d = np.zeros(2 * mesh.n)
    _load().or_fullwave_diag(mesh.ptr, float(freq), _p(d, f64p))
    return d[0::2] + 1j * d[1::2]


def rhs(mesh, freq, port):
    b = np.zeros(2 * mesh.n)
    _load().or_rhs(mesh.ptr, float(freq), int(port), _p(b, f64p))
    return b[0::2] + 1j * b[1::2]


def extract_z(mesh, x, src, Zk=None):
    np_ = mesh.num_ports
    buf = np.zeros(2 * np_ * np_) if Zk is None else np.ascontiguousarray(
        np.stack([Zk.real.ravel(), Zk.imag.ravel()], 1).ravel())
    xi = np.ascontiguousarray(np.stack([x.real, x.imag], 1).ravel())
    _load().or_extract_z(mesh.ptr, _p(xi, f64p), int(src), _p(buf, f64p))
    return buf[0::2] + 1j * buf[1::2]


def freq(mesh, k):
    return _load().or_freq(mesh.ptr, int(k))


def system(mesh, f, S=None):
    """A(w) = S + diag(...) as complex CSR, exactly as reference() assembles valc."""
    S = stiffness_csr(mesh) if S is None else S
    A = S.astype(np.complex128)
    A = A + sp.diags(fullwave_diag(mesh, f))
    return A.tocsc()


def _matvec_ld(rp, ci, vr, vi, xr, xi):
    """(vr + i vi) @ (xr + i xi) with products and sums in long double."""
    a = vr * xr[ci] - vi * xi[ci]
    b = vr * xi[ci] + vi * xr[ci]
    nz = np.diff(rp) > 0
    yr = np.zeros(len(rp) - 1, np.longdouble)
    yi = np.zeros(len(rp) - 1, np.longdouble)
    yr[nz] = np.add.reduceat(a, rp[:-1][nz])
    yi[nz] = np.add.reduceat(b, rp[:-1][nz])
    return yr, yi


def refined_solve(lu, A, b, steps=4):
    """FP64 LU solve plus iterative refinement with long-double residuals.

    The FP64 direct solution has a forward error ~ cond * eps (about 1e-6 at
    100 MHz on micron meshes, where the gradient modes sit ~1e11 below the
    curl modes).  Keeping x and the residual in 80-bit long double and
    correcting with the same LU removes that error down to ~cond * 5e-20."""
    A = A.tocsr()
    rp = A.indptr.astype(np.int64)
    ci = A.indices
    vr = A.data.real.astype(np.longdouble)
    vi = A.data.imag.astype(np.longdouble)
    br, bi = b.real.astype(np.longdouble), b.imag.astype(np.longdouble)
    x = lu.solve(b)
    xr, xi = x.real.astype(np.longdouble), x.imag.astype(np.longdouble)
    for _ in range(steps):
        yr, yi = _matvec_ld(rp, ci, vr, vi, xr, xi)
        rr, ri = br - yr, bi - yi
        d = lu.solve((rr.astype(np.float64) + 1j * ri.astype(np.float64)))
        xr += d.real.astype(np.longdouble)
        xi += d.imag.astype(np.longdouble)
    return xr, xi


# ---- double-double residuals (error-free transformations) -----------------
# refined_solve keeps x and the residual in 80-bit long double, which leaves a
# floor of about cond * 5e-20: at 10 MHz on a 48x48x12 mesh (cond ~ 1e13-1e15)
# the refinement stalls near 1e-7 of |Z| and then drifts.  refined_solve_dd
# carries x as an unevaluated sum of two doubles and forms b - A x exactly up
# to ~1e-32 relative (Dekker's product split, Knuth's TwoSum), so each step
# contracts the error by cond * 1e-16 until the FP64 system is solved exactly.

def _split(a):
    c = 134217729.0 * a                      # 2^27 + 1
    hi = c - (c - a)
    return hi, a - hi


def _two_prod(a, b):
    p = a * b
    ah, al = _split(a)
    bh, bl = _split(b)
    return p, ((ah * bh - p) + ah * bl + al * bh) + al * bl


def _two_sum(a, b):
    s = a + b
    bb = s - a
    return s, (a - (s - bb)) + (b - bb)


def _dd_add(ah, al, bh, bl):
    s, e = _two_sum(ah, bh)
    e = e + (al + bl)
    h = s + e
    return h, e - (h - s)


def _dd_mul(xh, xl, a):
    """(xh + xl) * a for a double a."""
    p, e = _two_prod(xh, a)
    e = e + xl * a
    h = p + e
    return h, e - (h - p)


def _padded(A):
    """CSR as dense (n, w) column / value arrays, padded with zero values."""
    A = A.tocsr()
    n = A.shape[0]
    cnt = np.diff(A.indptr)
    w = int(cnt.max()) if n else 0
    col = np.zeros((n, w), np.int64)
    vr = np.zeros((n, w))
    vi = np.zeros((n, w))
    pos = np.arange(A.nnz) - np.repeat(A.indptr[:-1], cnt)
    rows = np.repeat(np.arange(n), cnt)
    col[rows, pos] = A.indices
    vr[rows, pos] = A.data.real
    vi[rows, pos] = A.data.imag
    return col, vr, vi


def residual_dd(P, b, xr, xi):
    """b - A x in double-double; x = (xr_hi, xr_lo) + i (xi_hi, xi_lo)."""
    col, vr, vi = P
    n, w = col.shape
    rh, rl = b.real.copy(), np.zeros(n)
    ih, il = b.imag.copy(), np.zeros(n)
    for t in range(w):
        c = col[:, t]
        ar, ai = vr[:, t], vi[:, t]
        xrh, xrl, xih, xil = xr[0][c], xr[1][c], xi[0][c], xi[1][c]
        for (yh, yl, a, sg, re) in ((xrh, xrl, ar, -1.0, True), (xih, xil, ai, 1.0, True),
                                    (xih, xil, ar, -1.0, False), (xrh, xrl, ai, -1.0, False)):
            ph, pl = _dd_mul(yh, yl, a)
            if re:
                rh, rl = _dd_add(rh, rl, sg * ph, sg * pl)
            else:
                ih, il = _dd_add(ih, il, sg * ph, sg * pl)
    return rh + 1j * ih, (rh, rl), (ih, il)


def refined_solve_dd(lu, A, b, steps=60, tol=1e-18):
    """FP64 factorization (anything with .solve) plus iterative refinement
    with double-double iterate and residual.  Returns (xr, xi) as
    (hi, lo) pairs; stops once a correction is below tol of |x|."""
    P = _padded(A)
    n = A.shape[0]
    x = lu.solve(b)
    xr, xi = (x.real.copy(), np.zeros(n)), (x.imag.copy(), np.zeros(n))
    for _ in range(steps):
        r, _, _ = residual_dd(P, b, xr, xi)
        d = lu.solve(r)
        xr = _dd_add(xr[0], xr[1], d.real, np.zeros(n))
        xi = _dd_add(xi[0], xi[1], d.imag, np.zeros(n))
        if np.abs(d).max() <= tol * np.abs(xr[0] + 1j * xi[0]).max():
            break
    return xr, xi


def direct_solve(mesh, freqs, ports=None, S=None, refine=False, factor=None):
    """FP64 direct solve of every (port, freq) column; returns (X, Z).

    refine=True adds long-double iterative refinement (refined_solve),
    refine="dd" double-double refinement (refined_solve_dd), which solves the
    FP64 system exactly even where cond * 5e-20 is not small.  factor(A), if
    given, makes the factorization (default SuperLU with COLAMD ordering).

    X[k][p] is the solution vector, Z[k] the np x np block laid out as
    Construct_Z_V0_Vh writes it (index resp + np*src)."""
    ports = list(range(mesh.num_ports)) if ports is None else list(ports)
    S = stiffness_csr(mesh) if S is None else S
    np_ = mesh.num_ports
    X, Z = [], []
    for f in freqs:
        A = system(mesh, f, S)
        lu = factor(A) if factor is not None else spla.splu(A, permc_spec="COLAMD")
        xs = []
        Zk = np.zeros(np_ * np_, np.complex128)
        for p in ports:
            b = rhs(mesh, f, p)
            if refine == "dd":
                xr, xi = refined_solve_dd(lu, A, b)
                x = (xr[0] + xr[1]) + 1j * (xi[0] + xi[1])
            elif refine:
                xr, xi = refined_solve(lu, A, b)
                x = xr.astype(np.float64) + 1j * xi.astype(np.float64)
            else:
                x = lu.solve(b)
            xs.append(x)
            Zk = extract_z(mesh, x, p, Zk)
        X.append(xs)
        Z.append(Zk)
    return X, np.array(Z)


# ---- V0 quasi-static path (paraGenerator, src/matrixCon.cpp:48-603) ---------

def port_j(mesh, port):
    """J over original edges for one source port (src/matrixCon.cpp:367-374)."""
    J = np.zeros(mesh.n_edge)
    _load().or_port_j(mesh.ptr, int(port), _p(J, f64p))
    return J


def edge_eps(mesh):
    """stackEpsn[e / (N_edge_s + N_edge_v)] * EPSILON0 per original edge."""
    e = np.zeros(mesh.n_edge)
    _load().or_edge_eps(mesh.ptr, _p(e, f64p))
    return e


def extract_z_v0(mesh, yd, src, Z0):
    """Construct_Z_V0 (src/fdtd.hpp:4178-4215) into Z0[resp + np*src]."""
    buf = np.ascontiguousarray(np.stack([Z0.real, Z0.imag], 1).ravel())
    xi = np.ascontiguousarray(np.stack([yd.real, yd.imag], 1).ravel())
    _load().or_extract_z_v0(mesh.ptr, _p(xi, f64p), int(src), _p(buf, f64p))
    return buf[0::2] + 1j * buf[1::2]


def print_z_v0(mesh, Z0):
    """print_z_V0's store (src/fdtd.hpp:4249-4316): the nf*np*np sweep array
    sys->x, slot 0 = Z0, later slots Re + i Im f0/f stored transposed."""
    np_, nf = mesh.num_ports, mesh.c.nfreq
    X = np.zeros(2 * nf * np_ * np_)
    X[:2 * np_ * np_] = np.stack([Z0.real, Z0.imag], 1).ravel()
    _load().or_print_z_v0(mesh.ptr, _p(X, f64p))
    return X[0::2] + 1j * X[1::2]


def v0_operators(mesh):
    """V0d, V0da, V0c, V0ca (N_edge x cols, original edge rows) and Ad, Ac as
    SciPy CSR, from the restated basis (values before the sqrt(eps) /
    sqrt(sigma) scaling, which is what the MKL handles of the V0 section hold,
    src/matrixCon.cpp:104-108, 316-322)."""
    d, c = v0d(mesh), v0c(mesh)
    ne = mesh.n_edge
    nd, ncl = int(d["cols"]), int(c["cols"])
    Vd = sp.csr_matrix((d["val"], (d["row"], d["col"])), shape=(ne, nd))
    Vda = sp.csr_matrix((d["aval"], (d["row"], d["col"])), shape=(ne, nd))
    Vc = sp.csr_matrix((c["val"], (c["row"], c["col"])), shape=(ne, ncl))
    Vca = sp.csr_matrix((c["aval"], (c["row"], c["col"])), shape=(ne, ncl))
    k = v0_sizes(mesh)["ad_nnz"]          # kept entries; the trailing zero slots are never read
    Ad = sp.csr_matrix((d["ad_val"][:k], (d["ad_row"][:k], d["ad_col"][:k])), shape=(nd, nd))
    Ac = sp.csr_matrix((c["ac_val"], (c["ac_row"], c["ac_col"])), shape=(ncl, ncl))
    return dict(Vd=Vd, Vda=Vda, Vc=Vc, Vca=Vca, Ad=Ad, Ac=Ac)


def v0_path(mesh, ops=None, ports=None):
    """Z at freqStart from the V0 quasi-static solve of every source port,
    restating src/matrixCon.cpp:366-560 with exact (SuperLU) solves in place
    of hypreSolve.  Returns (Z0 [np*np], index resp + np*src; yd per port)."""
    ops = v0_operators(mesh) if ops is None else ops
    Vd, Vda, Vc, Vca = ops["Vd"], ops["Vda"], ops["Vc"], ops["Vca"]
    lud = spla.splu(ops["Ad"].tocsc())
    luc = spla.splu(ops["Ac"].tocsc())
    eps = edge_eps(mesh)
    w0 = 2 * np.pi * mesh.c.freq_start * mesh.c.freq_unit
    np_ = mesh.num_ports
    Z0 = np.zeros(np_ * np_, complex)
    yds = {}
    for src in (range(np_) if ports is None else ports):
        J = port_j(mesh, src)
        y0d = lud.solve(-(Vda.T @ J))                 # :377-392
        y0d = y0d / w0                                # :401-403
        yd1 = Vd @ y0d                                # :411 (ydt), :441 (yd1)
        ydt = yd1 * (-1.0 * w0 * eps)                 # :442
        rhs_c = -(Vca.T @ J) + Vca.T @ ydt            # :434-460
        y0c = luc.solve(rhs_c)                        # :511
        yc = Vc @ y0c                                 # :531
        yccp = -yc * eps                              # :536-538
        y0d2 = lud.solve(Vda.T @ yccp)                # :543-544
        yd2 = Vd @ y0d2                               # :552
        yd = (yd2 - 1j * yd1) + yc                    # :574-587
        Z0 = extract_z_v0(mesh, yd, src, Z0)
        yds[src] = yd
    return Z0, yds


# ---- Vh basis (fdtdMesh::find_Vh(k), src/fdtd.hpp:3577-3859) ----------------

def vh_coeffs(mesh):
    """(eps, sigma) per unknown as find_Vh reads them (src/fdtd.hpp:3601-3605)."""
    e, s = np.zeros(mesh.n), np.zeros(mesh.n)
    _load().or_vh_coeffs(mesh.ptr, _p(e, f64p), _p(s, f64p))
    return e, s


def vh_operator(mesh, scale=1e14, S=None):
    """The 2n x 2n operator find_Vh runs Arnoldi on (src/fdtd.hpp:3595-3606):
    [v1; v2] -> [v2; -S v1 / (eps s^2) - sigma v2 / (eps s)]."""
    S = stiffness_csr(mesh) if S is None else S
    eps, sig = vh_coeffs(mesh)
    n = mesh.n
    top = sp.hstack([sp.csr_matrix((n, n)), sp.identity(n, format="csr")])
    bot = sp.hstack([-sp.diags(1.0 / (eps * scale * scale)) @ S, sp.diags(-sig / (eps * scale))])
    return sp.vstack([top, bot]).tocsr()


def find_vh(mesh, k, scale=1e14, eps=1e-6, S=None):
    """find_Vh(k): k Arnoldi steps with modified Gram-Schmidt from the start
    vector 1 / sqrt(2n) (src/fdtd.hpp:3589-3636), the eigenvalues of the k x k
    Hessenberg H (dhseqr, :3754; numpy's LAPACK eigensolver here), the
    eigenvectors of those with |Im| > eps (dhsein, :3797), and
    Vh = V_top * eigenvectors (:3836-3850).  Conjugate pairs give the columns
    y, conj(y) (:3811-3822).  Returns (Vh [n, leng], ritz [k] (times scale),
    H [k, k], V [k + 1, 2n], h_next)."""
    M = vh_operator(mesh, scale, S)
    n2 = M.shape[0]
    V = np.zeros((k + 1, n2))
    H = np.zeros((k, k))
    V[0] = 1.0 / np.sqrt(n2)
    h_next = 0.0
    for j in range(1, k + 1):
        w = M @ V[j - 1]
        for i in range(j):
            H[i, j - 1] = V[i] @ w
            w = w - H[i, j - 1] * V[i]
        nrm = np.sqrt(w @ w)
        if j < k:
            H[j, j - 1] = nrm
        else:
            h_next = nrm
        V[j] = w / nrm
    lam, Y = np.linalg.eig(H)
    # dhseqr's order is not numpy's: order conjugate pairs positive part first
    sel = [i for i in range(k) if abs(lam[i].imag) > eps and lam[i].imag > 0]
    cols = []
    for i in sel:
        y = Y[:, i] / Y[np.argmax(np.abs(Y[:, i])), i]
        cols += [y, y.conj()]
    Yc = np.array(cols).T if cols else np.zeros((k, 0), complex)
    Vh = V[:k, :mesh.n].T @ Yc
    return Vh, lam * scale, H, V, h_next


def edge_map(mesh):
    """(mapEdge [N_edge], mapEdgeR [n]) as setMapEdge builds them."""
    m, r = np.empty(mesh.n_edge, np.int64), np.empty(mesh.n, np.int64)
    _load().or_edge_map(mesh.ptr, _p(m, i64p), _p(r, i64p))
    return m, r


def vh_inputs(mesh, src, ops=None):
    """u0, u0a (n x 2) and yd (N_edge complex) of paraGenerator's Vh section
    for one source port (src/matrixCon.cpp:377-427, :434-587), with exact
    (SuperLU) solves in place of hypreSolve."""
    ops = v0_operators(mesh) if ops is None else ops
    Vd, Vda, Vc, Vca = ops["Vd"], ops["Vda"], ops["Vc"], ops["Vca"]
    lud = spla.splu(ops["Ad"].tocsc())
    eps = edge_eps(mesh)
    w0 = 2 * np.pi * mesh.c.freq_start * mesh.c.freq_unit
    _, mr = edge_map(mesh)
    J = port_j(mesh, src)
    y0d = lud.solve(-(Vda.T @ J)) / w0                # :392, :401-403
    ydt, ydat = Vd @ y0d, Vda @ y0d                   # :411-412
    n = mesh.n
    u0, u0a = np.zeros((n, 2)), np.zeros((n, 2))
    u0[:, 0] = ydt[mr] / np.sqrt(np.sum(ydt * ydt))   # :416-427
    u0a[:, 0] = ydat[mr] / np.sqrt(np.sum(ydat * ydat))
    if ops["Ac"].shape[0] > 0:
        luc = spla.splu(ops["Ac"].tocsc())
        y0c = luc.solve(-(Vca.T @ J) + Vca.T @ (ydt * (-1.0 * w0 * eps)))   # :434-511
        yc, yca = Vc @ y0c, Vca @ y0c                                      # :528-529
    else:
        yc = yca = np.zeros(mesh.n_edge)
    y0d2 = lud.solve(Vda.T @ (-yc * eps))             # :536-544
    yd2, yd2a = Vd @ y0d2, Vda @ y0d2                 # :552-553
    u0[:, 1] = (yd2 + yc)[mr] / np.sqrt(np.sum((yd2 + yc) * (yd2 + yc)))   # :556-568
    u0a[:, 1] = (yd2a + yca)[mr] / np.sqrt(np.sum((yd2a + yca) * (yd2a + yca)))
    yd = (yd2 - 1j * ydt) + yc                        # :570-587
    return u0, u0a, yd


def vh_solve_coeffs(mesh):
    e, s = np.zeros(mesh.n), np.zeros(mesh.n)
    _load().or_vh_solve_coeffs(mesh.ptr, _p(e, f64p), _p(s, f64p))
    return e, s


def vh_port_rows(mesh, src):
    """Rows where the per-frequency Vh solve puts J = -i w: the unknowns of
    the source port's edges with J[e] != 0 (src/matrixCon.cpp:727-731)."""
    J = port_j(mesh, src)
    m, _ = edge_map(mesh)
    rows = m[np.nonzero(J)[0]]
    return rows[rows >= 0]


def vh_deflate(Vh, u0, u0a, d):
    """Vh - u0 (u0a^H D u0)^-1 (u0a^H D Vh) (src/matrixCon.cpp:611-690; u0 and
    u0a are real, so ^H is ^T)."""
    y_re = u0a.T @ (d[:, None] * Vh)                           # :612-641
    t4 = u0a.T @ (d[:, None] * u0)                             # :642-661
    return Vh - u0 @ np.linalg.solve(t4, y_re)                 # :671-690


def solve_vh(mesh, Vh, u0, u0a, yd, src, freqs, S=None):
    """paraGenerator's per-frequency Vh solve (src/matrixCon.cpp:608-801);
    matrix_multi 'T' is the conjugate transpose (src/findVh.cpp:1205-1218).
    Returns (X [nf, n], Z [nf, np*np] with column src filled)."""
    S = stiffness_csr(mesh) if S is None else S
    eps, sig = vh_solve_coeffs(mesh)
    _, mr = edge_map(mesh)
    rows = vh_port_rows(mesh, src)
    f0 = mesh.c.freq_start * mesh.c.freq_unit
    ydr = yd[mr]
    npt = mesh.num_ports
    Xs, Zs = [], []
    for f in freqs:
        w = 2 * np.pi * f
        d = -(w * w) * eps + 1j * w * sig
        Vp = vh_deflate(Vh, u0, u0a, d)                        # :611-690
        T = S @ Vp + d[:, None] * Vp                           # :692-722
        m_h = Vp.conj().T @ T                                  # :723-724
        Jc = np.zeros(mesh.n, complex)
        Jc[rows] = -1j * w                                     # :726-731
        yw = ydr.real + 1j * ydr.imag * f0 / f
        rhs = Vp.conj().T @ Jc - Vp.conj().T @ (d * yw)         # :732-776
        y = np.linalg.solve(m_h, rhs)                          # :777-781
        x = yw + Vp @ y                                        # :782-786
        Xs.append(x)
        Zs.append(extract_z(mesh, x, src))
    return np.array(Xs), np.array(Zs)

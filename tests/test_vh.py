"""Vh basis (fdtdMesh::find_Vh(k), src/fdtd.hpp:3577-3859), CPU side.

The Hessenberg matrix of find_Vh's Arnoldi process is not reproducible entry
by entry: the companion operator's spectrum spans about four decades (the
conductor damping -sigma / eps next to the resonances), and a relative
perturbation of 1e-16 in one step changes H by O(1) within a few steps, in
any implementation, the reference's included.  What is reproducible, and
pinned here, is the converged part: Ritz values whose residual estimate
|h_{k+1,k} y_k| is small are eigenvalues of the operator, and those match
the eigenvalues of find_reference_Vh's dense pencil (src/fdtd.hpp:3862-3960,
LAPACK dggev there, scipy.linalg.eig here).

Host routines of the product (the dhseqr / dhsein stand-ins, host-only C-ABI
calls) are checked against numpy's LAPACK eigensolver on random and
structured Hessenberg matrices.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg as sl

import oracle
from gds2para_amd import _capi as capi
from gds2para_amd import mesh as M

f64p = C.POINTER(C.c_double)


def _hess_eigs(H):
    n = H.shape[0]
    h = np.ascontiguousarray(H.T.ravel())          # column-major
    w = np.zeros(2 * n)
    capi.check(capi.load().gp_hessenberg_eigenvalues(n, h.ctypes.data_as(f64p), w.ctypes.data_as(f64p)),
               "gp_hessenberg_eigenvalues")
    return w[0::2] + 1j * w[1::2]


def _hess_vec(H, mu):
    n = H.shape[0]
    h = np.ascontiguousarray(H.T.ravel())
    v = np.zeros(2 * n)
    capi.check(capi.load().gp_hessenberg_eigenvector(n, h.ctypes.data_as(f64p), float(mu.real), float(mu.imag),
                                                     v.ctypes.data_as(f64p)), "gp_hessenberg_eigenvector")
    return v[0::2] + 1j * v[1::2]


def _match(a, b, tol):
    """Every value of a within tol of a distinct value of b (greedy, |a| = |b|)."""
    b = list(b)
    worst = 0.0
    for x in a:
        d = [abs(x - y) for y in b]
        i = int(np.argmin(d))
        worst = max(worst, d[i])
        b.pop(i)
    assert worst <= tol, worst
    return worst


def match_spectrum(w, ref, scale, sep=1e-3, tight=1e-9, loose=1e-6):
    """w against ref: eigenvalues of ref farther than sep * scale from every
    other one to tight * scale, the rest (clusters, whose eigenvalues the data
    fix only to about eps^(1/m)) to loose * scale."""
    ref = np.asarray(ref)
    d = np.abs(ref[:, None] - ref[None, :])
    np.fill_diagonal(d, np.inf)
    iso = d.min(axis=1) > sep * scale
    for t in ref[iso]:
        assert np.abs(w - t).min() <= tight * scale, t
    _match(w, ref, loose * scale)


def _hessenberg_cases():
    rng = np.random.default_rng(7)
    out = []
    for n in (1, 2, 3, 4, 5, 8, 17, 40, 100):
        out.append(("random", np.triu(rng.standard_normal((n, n)), -1)))
    # a split into independent blocks (zero subdiagonal entries)
    H = np.triu(rng.standard_normal((30, 30)), -1)
    H[10, 9] = 0.0
    H[20, 19] = 0.0
    out.append(("split", H))
    # companion matrix of a polynomial with roots on the unit circle (conjugate pairs)
    roots = np.exp(2j * np.pi * (np.arange(12) + 0.5) / 12) * 1.5
    c = np.real(np.poly(roots))
    Cm = np.zeros((12, 12))
    Cm[0, :] = -c[1:]
    Cm[1:, :-1] = np.eye(11)
    out.append(("companion", Cm))
    # symmetric tridiagonal (all eigenvalues real)
    d, e = rng.standard_normal(25), rng.standard_normal(24)
    out.append(("tridiagonal", np.diag(d) + np.diag(e, 1) + np.diag(e, -1)))
    # rotation blocks: eigenvalues a +- i b exactly
    R = np.zeros((8, 8))
    for q, (a, b) in enumerate([(1.0, 2.0), (-0.5, 0.25), (3.0, 1e-3), (0.0, 1.0)]):
        R[2 * q:2 * q + 2, 2 * q:2 * q + 2] = [[a, b], [-b, a]]
    R += np.triu(rng.standard_normal((8, 8)), 2) * 0.1
    out.append(("rotations", R))
    # widely spread scales, as in the Arnoldi matrices of find_Vh
    H = np.triu(rng.standard_normal((20, 20)), -1)
    H[0, 0] = -1.5e4
    out.append(("spread", H))
    return out


@pytest.mark.parametrize("name,H", _hessenberg_cases(), ids=lambda v: v if isinstance(v, str) else "")
def test_hessenberg_eigenvalues_match_lapack(name, H):
    w = _hess_eigs(H)
    ref = np.linalg.eigvals(H)
    scale = max(np.abs(H).max(), 1.0)
    _match(w, ref, 1e-9 * scale * (10 if name == "companion" else 1))
    # conjugate pairs adjacent, positive imaginary part first
    i = 0
    while i < len(w):
        if w[i].imag != 0:
            assert w[i].imag > 0 and w[i + 1] == np.conj(w[i]), (name, i)
            i += 2
        else:
            i += 1
    if name == "rotations":
        _match(w[np.abs(w.imag) > 0.5], [1 + 2j, 1 - 2j, 1j, -1j], 1e-12)


@pytest.mark.parametrize("name,H", [c for c in _hessenberg_cases() if c[1].shape[0] >= 3],
                         ids=lambda v: v if isinstance(v, str) else "")
def test_hessenberg_eigenvector_residual(name, H):
    nrm = np.abs(H).max()
    for mu in _hess_eigs(H):
        v = _hess_vec(H, mu)
        assert abs(np.abs(v).max() - 1.0) < 1e-12
        r = np.linalg.norm(H @ v - mu * v) / (nrm * np.linalg.norm(v))
        assert r < 1e-11, (name, mu, r)


def test_hessenberg_eigenvalues_arnoldi_cluster():
    """find_Vh's Hessenberg matrix on a Manhattan mesh at k = 200: a cluster
    of about 130 nearly equal real values (the conductor damping) on which
    the plain deflation test stalls; the QR still ends, isolated eigenvalues
    agree with LAPACK's to 1e-9 |H|, the cluster to 1e-6 |H|."""
    m = M.Mesh(M.manhattan_spec((16, 16, 10), n_ports=2, nfreq=2, seed=3))
    H = oracle.find_vh(m, 200)[2]
    w = _hess_eigs(H)
    match_spectrum(w, np.linalg.eigvals(H), np.abs(H).max())


def test_hessenberg_argument_errors():
    lib = capi.load()
    w = np.zeros(4)
    assert lib.gp_hessenberg_eigenvalues(0, w.ctypes.data_as(f64p), w.ctypes.data_as(f64p)) == capi.GP_ERR_ARG
    assert lib.gp_hessenberg_eigenvalues(2, None, w.ctypes.data_as(f64p)) == capi.GP_ERR_ARG


# ---- oracle pins ----------------------------------------------------------------

def _pencil_eigs(m, scale=1e14):
    """find_reference_Vh's pencil (src/fdtd.hpp:3868-3889): A = [-S 0; 0 s^2 eps],
    B = [s sigma  s^2 eps; s^2 eps  0]; eigenvalues mu = lambda / s."""
    S = oracle.stiffness_csr(m).toarray()
    eps, sig = oracle.vh_coeffs(m)
    n = m.n
    A = np.zeros((2 * n, 2 * n))
    B = np.zeros((2 * n, 2 * n))
    A[:n, :n] = -S
    A[n:, n:] = scale * scale * np.diag(eps)
    B[:n, :n] = scale * np.diag(sig)
    B[n:, :n] = scale * scale * np.diag(eps)
    B[:n, n:] = scale * scale * np.diag(eps)
    mu = sl.eigvals(A, B)
    return mu[np.isfinite(mu)]


def converged(H, h_next, tol):
    """Ritz values of H whose residual estimate |h_next y_k| / |theta| <= tol."""
    th, Y = np.linalg.eig(H)
    Y = Y / np.linalg.norm(Y, axis=0)
    res = np.abs(h_next * Y[-1, :]) / np.maximum(np.abs(th), 1e-300)
    return th[res <= tol]


def quad_residual(m, u, lam, S=None):
    """|S u + lam sigma u + lam^2 eps u| relative to the size of its terms."""
    S = oracle.stiffness_csr(m) if S is None else S
    eps, sig = oracle.vh_coeffs(m)
    r = S @ u + lam * sig * u + lam * lam * eps * u
    size = np.linalg.norm(S @ u) + abs(lam) * np.linalg.norm(sig * u) + abs(lam) ** 2 * np.linalg.norm(eps * u)
    return np.linalg.norm(r) / size


@pytest.mark.parametrize("cells,k", [((3, 3, 6), 60), ((4, 4, 6), 80)])
def test_oracle_vh_against_dense_pencil(cells, k):
    m = M.Mesh(M.plate_spec(cells))
    Vh, lam, H, V, h_next = oracle.find_vh(m, k)
    # Arnoldi relation of the oracle's H with the operator built apart from it
    Mop = oracle.vh_operator(m)
    Hbar = np.vstack([H, np.zeros((1, k))])
    Hbar[k, k - 1] = h_next
    assert np.abs(Mop @ V[:k].T - V.T @ Hbar).max() <= 1e-10 * np.abs(H).max()
    mu = _pencil_eigs(m)
    conv = converged(H, h_next, 1e-10)
    cplx = conv[np.abs(conv.imag) > 1e-6]
    assert len(cplx) >= 10
    for t in cplx:
        assert np.abs(mu - t).min() <= 1e-9 * abs(t), t
    # the real ones belong to the conductor-damping cluster (-sigma / eps, tightly
    # clustered and ill-conditioned): a small residual bounds them less tightly
    for t in conv[np.abs(conv.imag) <= 1e-6]:
        assert np.abs(mu - t).min() <= 1e-8 * abs(t), t
    # converged Vh columns solve the quadratic eigenproblem
    S = oracle.stiffness_csr(m)
    lam_sel = [l for l in lam if abs(l.imag) > 1e-6 * 1e14 and l.imag > 0]
    assert Vh.shape[1] == 2 * len(lam_sel)
    checked = 0
    for c, l in enumerate(lam_sel):
        if np.abs(cplx - l / 1e14).min() <= 1e-12 * abs(l / 1e14):
            u = Vh[:, 2 * c]
            assert quad_residual(m, u, l, S) < 1e-8
            assert np.array_equal(Vh[:, 2 * c + 1], u.conj())
            checked += 1
    assert checked >= 5


def test_oracle_vh_planted_sigma_mistake_is_caught():
    """The pencil check fails for an operator that drops the conductor term:
    its converged Ritz values are not eigenvalues of the pencil."""
    m = M.Mesh(M.plate_spec((3, 3, 6)))
    mu = _pencil_eigs(m)
    eps, sig = oracle.vh_coeffs(m)
    S = oracle.stiffness_csr(m)
    n, s = m.n, 1e14
    import scipy.sparse as sp
    top = sp.hstack([sp.csr_matrix((n, n)), sp.identity(n, format="csr")])
    bot = sp.hstack([-sp.diags(1.0 / (eps * s * s)) @ S, sp.csr_matrix((n, n))])   # sigma dropped
    Mbad = sp.vstack([top, bot]).tocsr()
    k = 60
    V = np.zeros((k + 1, 2 * n))
    H = np.zeros((k, k))
    V[0] = 1 / np.sqrt(2 * n)
    for j in range(1, k + 1):
        w = Mbad @ V[j - 1]
        for i in range(j):
            H[i, j - 1] = V[i] @ w
            w = w - H[i, j - 1] * V[i]
        if j < k:
            H[j, j - 1] = np.linalg.norm(w)
        V[j] = w / np.linalg.norm(w)
    conv = converged(H, np.linalg.norm(w), 1e-10)
    assert len(conv) > 0
    assert max(np.abs(mu - t).min() / abs(t) for t in conv) > 1e-6


# ---- per-frequency Vh solve (src/matrixCon.cpp:608-801) --------------------------

VH_SOLVE_MESHES = {
    "plate": (lambda: M.plate_spec((4, 4, 6)), 80),
    "manhattan": (lambda: M.manhattan_spec((16, 16, 10), n_ports=2, nfreq=2, seed=3), 120),
}


def exact_in_span_case(m, f, u0, yd, seed=1):
    """Inputs for which the Vh solve's answer is known: x* solves
    (S + D(w)) x = J with the loop's own coefficients (SuperLU with
    double-double refinement), Vh holds
    v = x* - y_w and two random columns, and u0a is random but orthogonal to
    Re(D v) and Im(D v), so the deflation leaves v in the span.  The Galerkin
    solution on y_w + span(Vh') is then x* (S y_w = 0: yd is a gradient field)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    S = oracle.stiffness_csr(m)
    eps, sig = oracle.vh_solve_coeffs(m)
    w = 2 * np.pi * f
    d = -(w * w) * eps + 1j * w * sig
    Jc = np.zeros(m.n, complex)
    Jc[oracle.vh_port_rows(m, 0)] = -1j * w
    A = (S + sp.diags(d)).tocsc()
    xr, xi = oracle.refined_solve_dd(spla.splu(A), A, Jc)   # low-frequency breakdown: refine
    xs = (xr[0] + xr[1]) + 1j * (xi[0] + xi[1])
    _, mr = oracle.edge_map(m)
    f0 = m.c.freq_start * m.c.freq_unit
    yw = yd[mr].real + 1j * yd[mr].imag * f0 / f
    v = xs - yw
    rng = np.random.default_rng(seed)
    Q = np.linalg.qr(np.stack([(d * v).real, (d * v).imag], 1))[0]
    ua = rng.standard_normal((m.n, 2))
    ua -= Q @ (Q.T @ ua)
    Vh = np.stack([v, rng.standard_normal(m.n) + 1j * rng.standard_normal(m.n), rng.standard_normal(m.n) + 0j], 1)
    return Vh, ua, xs


@pytest.fixture(scope="module", params=sorted(VH_SOLVE_MESHES))
def vh_case(request):
    spec, k = VH_SOLVE_MESHES[request.param]
    m = M.Mesh(spec())
    u0, u0a, yd = oracle.vh_inputs(m, 0)
    return request.param, m, k, u0, u0a, yd


def test_oracle_vh_inputs_are_v0_fields(vh_case):
    """u0's columns and yd lie in the null space of S (gradient fields, S V0 = 0)
    to rounding, each u0 column has norm <= 1 over the kept rows (normalised
    over all N_edge entries), and yd's Construct_Z_V0 is the V0 path's Z0."""
    name, m, k, u0, u0a, yd = vh_case
    S = oracle.stiffness_csr(m)
    smax = abs(S).max()
    _, mr = oracle.edge_map(m)
    for c in range(2):
        assert np.linalg.norm(S @ u0[:, c]) <= 1e-14 * smax * np.linalg.norm(u0[:, c])
        assert np.linalg.norm(u0[:, c]) <= 1 + 1e-12 and np.linalg.norm(u0a[:, c]) <= 1 + 1e-12
    assert np.linalg.norm(S @ yd[mr]) <= 1e-14 * smax * np.linalg.norm(yd[mr])
    Z0, _ = oracle.v0_path(m, ports=[0])
    Zy = oracle.extract_z_v0(m, yd, 0, np.zeros(m.num_ports ** 2, complex))
    np.testing.assert_allclose(Zy, Z0, rtol=0, atol=1e-12 * np.abs(Z0).max())


def test_oracle_vh_deflation_properties(vh_case):
    """Vh' = Vh - u0 Y is D-orthogonal to u0a (u0a^T D Vh' = 0) and differs
    from Vh only inside span(u0)."""
    name, m, k, u0, u0a, yd = vh_case
    Vh = oracle.find_vh(m, k)[0]
    eps, sig = oracle.vh_solve_coeffs(m)
    for f in (1e8, 1e10):
        w = 2 * np.pi * f
        d = -(w * w) * eps + 1j * w * sig
        Vp = oracle.vh_deflate(Vh, u0, u0a, d)
        lhs = u0a.T @ (d[:, None] * Vp)
        assert np.abs(lhs).max() <= 1e-6 * np.abs(u0a.T @ (d[:, None] * Vh)).max()
        Y = np.linalg.lstsq(u0, Vh - Vp, rcond=None)[0]
        assert np.linalg.norm(Vh - Vp - u0 @ Y) <= 1e-12 * np.linalg.norm(Vh - Vp)


def span_tol(f):
    """Accuracy of the Vh solve in FP64 when the answer is in the span: forming
    (S + D) Vh' cancels |S v| down to |D v|, so rounding grows like
    |S| / (w^2 eps), about 1e-5 relative at 100 MHz on these meshes."""
    return 5e-5 * (1e8 / f) ** 2 + 1e-8


@pytest.mark.parametrize("f", [1e8, 1e9, 1e10])
def test_oracle_solve_vh_exact_in_span(vh_case, f):
    name, m, k, u0, u0a, yd = vh_case
    Vh, ua, xs = exact_in_span_case(m, f, u0, yd)
    X, Z = oracle.solve_vh(m, Vh, u0, ua, yd, 0, [f])
    assert np.linalg.norm(X[0] - xs) <= span_tol(f) * np.linalg.norm(xs)
    Zs = oracle.extract_z(m, xs, 0)
    assert np.abs(Z[0] - Zs).max() <= span_tol(f) * np.abs(Zs).max()


@pytest.mark.parametrize("basis", ["find_vh", "random"])
def test_oracle_solve_vh_galerkin_residual(vh_case, basis):
    """The residual J - (S + D) x of the Vh solution is conjugate-orthogonal
    to the deflated basis, Vh'^H r = -Vh'^H S y_w ~ 0, within the rounding of
    the products (|S| |x| bounds that of S x).  find_Vh's basis is closed
    under conjugation (pairs y, conj y), where ^T and ^H give the same
    solution; the random complex basis is not, so it tells them apart."""
    name, m, k, u0, u0a, yd = vh_case
    if basis == "find_vh":
        Vh = oracle.find_vh(m, k)[0]
    else:
        rng = np.random.default_rng(4)
        Vh = rng.standard_normal((m.n, 6)) + 1j * rng.standard_normal((m.n, 6))
    S = oracle.stiffness_csr(m)
    eps, sig = oracle.vh_solve_coeffs(m)
    rows = oracle.vh_port_rows(m, 0)
    for f in (1e8, 1e10):
        X, _ = oracle.solve_vh(m, Vh, u0, u0a, yd, 0, [f], S)
        w = 2 * np.pi * f
        d = -(w * w) * eps + 1j * w * sig
        Jc = np.zeros(m.n, complex)
        Jc[rows] = -1j * w
        r = Jc - (S @ X[0] + d * X[0])
        Vp = oracle.vh_deflate(Vh, u0, u0a, d)
        g = Vp.conj().T @ r
        rnd = 64 * 2.2e-16 * np.linalg.norm(abs(S) @ np.abs(X[0]))
        scale = np.linalg.norm(Vp, axis=0) * (1e-8 * (np.linalg.norm(Jc) + np.linalg.norm(d * X[0])) + rnd)
        assert np.all(np.abs(g) <= scale), np.max(np.abs(g) / scale)

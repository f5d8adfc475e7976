"""Device memory held by a layout across solves: one hierarchy whichever
precision comes first, a solve workspace reused (not reallocated) by calls of
the same shape, and a per-frequency coarse-level cache that stays within its
bound while the results stay those of the unbounded solve."""
import numpy as np
import pytest

from gds2para_amd import _capi as capi
from gds2para_amd import mesh as M
from gds2para_amd.layout import Layout

pytestmark = pytest.mark.gpu


def _spec():
    return M.manhattan_spec((24, 20, 10), n_ports=3, nfreq=3, f_range=(1e8, 1e10), seed=11)


def _relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_fp64_then_mixed_keeps_one_hierarchy():
    m = M.Mesh(_spec())
    fs = [1e9]
    print("[mem] mixed-only layout", flush=True)
    ref = Layout(m)
    ref.prepare(capi.GP_PREC_MIXED)
    base = ref.device_bytes
    Zr, _, _ = ref.solve(fs, precision=capi.GP_PREC_MIXED, batch=4)
    ref.close()
    print("[mem] FP64 solve, then mixed solve, on one layout", flush=True)
    lay = Layout(m)
    Zd, std, _ = lay.solve(fs, precision=capi.GP_PREC_FP64, batch=1)
    Zm, stm, _ = lay.solve(fs, precision=capi.GP_PREC_MIXED, batch=4)
    lay.trim()
    after = lay.device_bytes
    print(f"[mem] mixed-only {base} B, FP64 then mixed {after} B", flush=True)
    assert after == base
    assert std["status"] == stm["status"] == capi.GP_OK
    assert _relerr(Zm, Zr) < 1e-6 and _relerr(Zd, Zr) < 1e-6
    lay.close()


def test_workspace_reused_across_calls():
    m = M.Mesh(_spec())
    lay = Layout(m)
    fs = m.freqs()
    print("[mem] two solves of one shape", flush=True)
    Z1, s1, _ = lay.solve(fs, precision=capi.GP_PREC_MIXED, batch=8)
    b1 = lay.device_bytes
    ws1 = lay.hierarchy_info()["workspace_bytes"]
    Z2, s2, _ = lay.solve(fs, precision=capi.GP_PREC_MIXED, batch=8)
    b2 = lay.device_bytes
    assert ws1 > 0 and b1 == b2
    assert s1["status"] == s2["status"] == capi.GP_OK
    assert np.array_equal(Z1, Z2)           # same buffers, same schedule: the same bits
    lay.trim()
    info = lay.hierarchy_info()
    assert info["workspace_bytes"] == 0 and info["coarse_entries"] == 0
    assert lay.device_bytes < b2
    Z3, s3, _ = lay.solve(fs, precision=capi.GP_PREC_MIXED, batch=8)
    assert np.array_equal(Z1, Z3)
    lay.close()


def test_coarse_cache_bounded():
    """70 distinct frequencies through a cache that holds 64: the oldest
    unpinned entries are evicted, and every column still matches the same
    frequency solved on a fresh layout."""
    m = M.Mesh(_spec())
    fs = np.logspace(8, 10, 70)
    lay = Layout(m)
    print("[mem] 70 frequencies, 1 port, 4 slots", flush=True)
    Z, st, _ = lay.solve(fs, ports=[0], precision=capi.GP_PREC_MIXED, batch=4)
    info = lay.hierarchy_info()
    print(f"[mem] entries {info['coarse_entries']} evictions {info['coarse_evictions']} "
          f"entry {info['coarse_entry_bytes']} B", flush=True)
    assert st["status"] == capi.GP_OK and st["converged"] == 70
    assert info["coarse_entries"] <= 64 and info["coarse_evictions"] >= 6
    lay.close()
    pick = [0, 5, 33, 69]
    fresh = Layout(m)
    Zf, sf, _ = fresh.solve(fs[pick], ports=[0], precision=capi.GP_PREC_MIXED, batch=4)
    fresh.close()
    npt = m.num_ports
    for a, k in enumerate(pick):
        col, ref = Z[k].reshape(npt, npt)[0], Zf[a].reshape(npt, npt)[0]   # source port 0: index resp + np*0
        assert _relerr(col, ref) < 1e-6, (k, _relerr(col, ref))


@pytest.mark.parametrize("f", [1e7, 1e9, 5e10])
def test_aux_coarse_block_inverse(f):
    """The auxiliary coarsest level is inverted block by block (its connected
    components); assembled, the inverse must be the inverse of the whole
    matrix: A A^-1 = I to 1e-10 and the same as numpy's dense inverse."""
    import ctypes as C
    m = M.Mesh(M.WORKLOADS["manhattan_48x48x12"]())
    lay = Layout(m)
    lay.prepare(capi.GP_PREC_MIXED)
    info = lay.hierarchy_info()
    n = C.c_int64()
    capi.check(lay.lib.gp_coarse_aux_inverse(lay.h, f, C.byref(n), None, None), "gp_coarse_aux_inverse")
    n = n.value
    mat, inv = np.zeros(2 * n * n), np.zeros(2 * n * n)
    f64p = C.POINTER(C.c_double)
    capi.check(lay.lib.gp_coarse_aux_inverse(lay.h, f, C.byref(C.c_int64()), mat.ctypes.data_as(f64p),
                                             inv.ctypes.data_as(f64p)), "gp_coarse_aux_inverse")
    lay.close()
    A = (mat[0::2] + 1j * mat[1::2]).reshape(n, n)
    Ai = (inv[0::2] + 1j * inv[1::2]).reshape(n, n)
    print(f"[mem] aux coarsest {n} rows, {info['aux_coarse_blocks']} blocks, largest "
          f"{info['aux_coarse_max_block']}, eigenvector form {info['aux_coarse_eigen']}", flush=True)
    assert info["aux_coarse_blocks"] >= 1 and 1 <= info["aux_coarse_max_block"] <= n
    assert np.abs(A @ Ai - np.eye(n)).max() < 1e-10
    ref = np.linalg.inv(A)
    assert np.abs(Ai - ref).max() / np.abs(ref).max() < 1e-10


@pytest.mark.parametrize("prec,k", [(capi.GP_PREC_FP64, 16), (capi.GP_PREC_FP64, 1), (capi.GP_PREC_MIXED, 8)])
def test_aux_coarse_eigen_apply(prec, k):
    """The auxiliary coarsest solve through the pencil's eigenvectors, on the
    FP64 matrix cores, against numpy's dense solve of the same matrix at
    1 GHz: 1e-10 relative in FP64, 1e-6 with the FP32 input and output of the
    mixed V-cycle.  k = 1 and 16 columns cover partial and full MFMA column
    tiles; the columns carry unrelated random data."""
    import ctypes as C
    m = M.Mesh(M.WORKLOADS["manhattan_48x48x12"]())
    lay = Layout(m)
    lay.prepare(capi.GP_PREC_MIXED)
    assert lay.hierarchy_info()["aux_coarse_eigen"] == 1
    f64p = C.POINTER(C.c_double)
    n = C.c_int64()
    capi.check(lay.lib.gp_coarse_aux_inverse(lay.h, 1e9, C.byref(n), None, None), "gp_coarse_aux_inverse")
    n = n.value
    mat = np.zeros(2 * n * n)
    capi.check(lay.lib.gp_coarse_aux_inverse(lay.h, 1e9, C.byref(C.c_int64()), mat.ctypes.data_as(f64p), None),
               "gp_coarse_aux_inverse")
    A = (mat[0::2] + 1j * mat[1::2]).reshape(n, n)
    rng = np.random.default_rng(7)
    R = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    r = np.ascontiguousarray(np.stack([R.real, R.imag], -1).ravel())
    x = np.zeros_like(r)
    print(f"[mem] eigen-form coarse apply n={n} k={k} prec={prec}", flush=True)
    capi.check(lay.lib.gp_coarse_aux_apply(lay.h, 1e9, k, prec, r.ctypes.data_as(f64p), x.ctypes.data_as(f64p)),
               "gp_coarse_aux_apply")
    lay.close()
    X = (x[0::2] + 1j * x[1::2]).reshape(n, k)
    if prec == capi.GP_PREC_MIXED:               # the kernel sees R rounded to FP32
        R = R.astype(np.complex64).astype(np.complex128)
    ref = np.linalg.solve(A, R)
    err = np.abs(X - ref).max() / np.abs(ref).max()
    print(f"[mem] relative error {err:.2e}", flush=True)
    assert err < (1e-10 if prec == capi.GP_PREC_FP64 else 1e-6)

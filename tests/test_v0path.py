"""The V0 quasi-static path: paraGenerator's V0 section (src/matrixCon.cpp:366-560),
Construct_Z_V0 (src/fdtd.hpp:4178-4215) and print_z_V0's store
(src/fdtd.hpp:4249-4316).

Pins of the oracle (CPU):
  * physics: at a low frequency f the quasi-static Z, extrapolated as
    Re Z0 + i Im Z0 f0/f, agrees with the independent full-wave direct solve
    (S assembly + SuperLU), to 1e-4 of |Z| overall and 1e-3 on the real part.
    Dropping the conductor term yc, the second dielectric solve y0d2 or the
    coupling right-hand side, or flipping the sign of the coupling, moves
    the real part by 35% to 280% on both layouts, so each of those fails here;
  * known answers for print_z_V0's store on an asymmetric block, worked by
    hand from the reference's loop (transposed store, f0/f scaling, the last
    slot at freqEnd), on a log and a linear grid.
GPU: gp_solve_v0 against the oracle to 1e-6 (real solves at rtol 1e-10), the
sweep array, repeat calls (setup kept, identical Z), and the reference's
default tolerance (rtol 0 -> 1e-4)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from gds2para_amd import _capi as capi
from gds2para_amd import mesh as M
from gds2para_amd.layout import z_v0_sweep

SPECS = {
    "microstrip": lambda: M.microstrip_spec(),
    "manhattan_small": lambda: M.manhattan_spec((24, 20, 10), n_ports=3, nfreq=3, f_range=(1e8, 1e10), seed=11),
    "manhattan_graded": lambda: M.manhattan_spec((64, 48, 16), n_ports=4, nfreq=2, seed=5),
}


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("name", ["microstrip", "manhattan_small"])
def test_oracle_v0_matches_fullwave_at_low_frequency(name):
    m = M.Mesh(SPECS[name]())
    Z0, _ = oracle.v0_path(m)
    f0 = m.c.freq_start * m.c.freq_unit
    f = 1e8
    _, Zf = oracle.direct_solve(m, [f], refine=True)
    Ze = Z0.real + 1j * Z0.imag * f0 / f
    assert _rel(Ze, Zf[0]) < 1e-4
    assert _rel(Ze.real, Zf[0].real) < 1e-3
    # reciprocity of the quasi-static block
    Zm = Z0.reshape(m.num_ports, m.num_ports)
    assert _rel(Zm, Zm.T) < 1e-6


def _kat_mesh(scale):
    m = M.Mesh(M.manhattan_spec((24, 20, 10), n_ports=3, nfreq=4, f_range=(1e8, 1e10), seed=11))
    c = capi.GpMesh.from_buffer_copy(m.c)
    c.freq_scale = scale
    return m, c


@pytest.mark.parametrize("scale", [0, 1])
def test_print_z_v0_known_answers(scale, lib):
    m, c = _kat_mesh(scale)
    Z0 = np.arange(9) + 1j * 10.0 * np.arange(9)          # Z0[k] = k + 10k i, asymmetric
    # grid: freqStart 1e8, freqEnd 1e10, 4 points; slot 3 is freqEnd exactly
    f = [1e8, 1e8 * 100 ** (1 / 3), 1e8 * 100 ** (2 / 3), 1e10] if scale != 1 else [1e8, 3.4e9, 6.7e9, 1e10]
    want = np.zeros(36, complex)
    want[:9] = Z0
    for k in range(1, 4):
        for i in range(3):
            for j in range(3):
                src = Z0[j + 3 * i]                      # read (resp j, src i) ...
                want[9 * k + i + 3 * j] = src.real + 1j * src.imag * 1e8 / f[k]   # ... stored at (resp i, src j)
    # e.g. slot 1, position 1 holds Z0[3] = 3 + 30i with Im scaled by f0/f1
    assert want[10].real == 3.0
    np_ = 3
    x = np.zeros(2 * 4 * np_ * np_)
    x[:18] = np.stack([Z0.real, Z0.imag], 1).ravel()
    capi.check(lib.gp_z_v0_sweep(C.pointer(c), x.ctypes.data_as(C.POINTER(C.c_double))), "gp_z_v0_sweep")
    got = x[0::2] + 1j * x[1::2]
    assert np.allclose(got, want, rtol=1e-14, atol=0)
    xo = np.zeros(2 * 36)
    xo[:18] = x[:18]
    oracle._load().or_print_z_v0(C.pointer(c), xo.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.array_equal(xo, x)


def test_sweep_wrapper_matches_oracle():
    m = M.Mesh(SPECS["manhattan_small"]())
    Z0 = np.random.default_rng(3).standard_normal(9) + 1j * np.random.default_rng(4).standard_normal(9)
    assert np.array_equal(z_v0_sweep(m, Z0), oracle.print_z_v0(m, Z0))


def test_v0_set_smoother_rejects_bad_arguments(lib):
    # argument checks come before any device call, so this runs without a GPU
    assert lib.gp_v0_set_smoother(None, capi.GP_SMOOTH_CHEBYSHEV) == capi.GP_ERR_ARG


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(SPECS))
def test_gpu_v0_matches_oracle(name):
    from gds2para_amd.layout import Layout
    m = M.Mesh(SPECS[name]())
    print(f"[v0path] {name} np={m.num_ports}", flush=True)
    Zo, _ = oracle.v0_path(m)
    Xo = oracle.print_z_v0(m, Zo)
    lay = Layout(m)
    Z0, X, st = lay.solve_v0(rtol=1e-10)
    print(f"[v0path] {name} stats {st}", flush=True)
    assert st["max_rel_residual"] <= 1e-10
    has_c = st["ac_rows"] > 0
    assert st["solves"] == (3 if has_c else 2) * m.num_ports
    assert _rel(Z0, Zo) < 1e-6, _rel(Z0, Zo)
    assert _rel(X, Xo) < 1e-6
    # a second call keeps the setup and repeats the result exactly
    Z1, X1, st1 = lay.solve_v0(rtol=1e-10)
    assert st1["setup_ms"] == st["setup_ms"]
    assert np.array_equal(Z1, Z0) and np.array_equal(X1, X)
    # l1-Jacobi smoothing (the default is Chebyshev) rebuilds the Ad / Ac
    # multigrid and meets the same bar
    lay.set_v0_smoother(capi.GP_SMOOTH_L1_JACOBI)
    Z2, X2, st2 = lay.solve_v0(rtol=1e-10)
    print(f"[v0path] {name} l1-Jacobi: max {st2['max_iters']} iterations (Chebyshev {st['max_iters']})", flush=True)
    assert st2["max_rel_residual"] <= 1e-10
    assert _rel(Z2, Zo) < 1e-6, _rel(Z2, Zo)
    assert _rel(X2, Xo) < 1e-6
    lay.close()


@pytest.mark.gpu
def test_gpu_v0_reference_tolerance():
    from gds2para_amd.layout import Layout
    m = M.Mesh(SPECS["manhattan_small"]())
    print("[v0path] manhattan_small at rtol 0 (1e-4)", flush=True)
    Zo, _ = oracle.v0_path(m)
    lay = Layout(m)
    Z0, X, st = lay.solve_v0(rtol=0.0)
    assert st["max_rel_residual"] <= 1e-4
    assert _rel(Z0, Zo) < 1e-2
    lay.close()

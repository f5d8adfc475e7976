"""gp_apply_fullwave_static, the allocation-free y = A(w) x, against
gp_apply_fullwave on the same layout: the same bits on two calls in a row,
the second with a smaller input, with the layout's device bytes unchanged;
sizes beyond its static buffer and unprepared layouts are refused."""
import ctypes as C

import numpy as np
import pytest

from gds2para_amd import _capi as capi
from gds2para_amd import mesh as M
from gds2para_amd.layout import Layout

pytestmark = pytest.mark.gpu
f64p = C.POINTER(C.c_double)


def _static(lay, freqs, X):
    freqs = np.ascontiguousarray(np.asarray(freqs, np.float64))
    n = lay.mesh.n
    xi = np.ascontiguousarray(np.stack([X.real, X.imag], -1).ravel())
    yi = np.zeros_like(xi)
    code = lay.lib.gp_apply_fullwave_static(lay.h, freqs.ctypes.data_as(f64p), len(freqs),
                                            xi.ctypes.data_as(f64p), yi.ctypes.data_as(f64p))
    return code, (yi[0::2] + 1j * yi[1::2]).reshape(len(freqs), n)


def test_static_apply_matches_allocating_apply():
    m = M.Mesh(M.microstrip_spec())
    lay = Layout(m)
    lay.assemble()
    lay.prepare(capi.GP_PREC_FP64)
    rng = np.random.default_rng(3)
    for k in (8, 2):                                # second call smaller than the first
        fs = np.logspace(7, 10, k)
        X = rng.standard_normal((k, m.n)) + 1j * rng.standard_normal((k, m.n))
        print(f"[static] microstrip n={m.n} k={k}", flush=True)
        before = lay.device_bytes
        code, Ys = _static(lay, fs, X)
        assert code == capi.GP_OK, capi.load().gp_last_error()
        assert lay.device_bytes == before
        Yr = lay.apply_fullwave(fs, X)
        assert np.array_equal(Ys, Yr)
    lay.close()


def test_static_apply_refuses_what_it_cannot_do():
    m = M.Mesh(M.WORKLOADS["manhattan_48x48x12"]())
    lay = Layout(m)
    lay.assemble()
    X = np.zeros((32, m.n), complex)
    print("[static] unprepared layout, then n*k above the static buffer", flush=True)
    code, _ = _static(lay, np.full(32, 1e9), X)
    assert code == capi.GP_ERR_STATE
    lay.prepare(capi.GP_PREC_FP64)
    code, _ = _static(lay, np.full(32, 1e9), X)    # 85,260 x 32 > 2,097,146
    assert code == capi.GP_ERR_ARG
    assert b"static buffer" in capi.load().gp_last_error()
    code, Y = _static(lay, np.full(16, 1e9), X[:16] + 1.0)
    assert code == capi.GP_OK
    assert np.array_equal(Y, lay.apply_fullwave(np.full(16, 1e9), X[:16] + 1.0))
    lay.close()

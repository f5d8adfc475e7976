"""GPU parity of the full-wave solve (reference()/reference1() PARDISO path)
against the refined FP64 direct solve of the oracle: Z within 1e-6 relative,
solution vectors within 1e-6, true residual <= 1e-8, for the plain FP64
single-column path and the mixed-precision batched path."""
import numpy as np
import pytest

import oracle
from gds2para_amd import _capi as capi
from gds2para_amd import mesh as M
from gds2para_amd.layout import Layout

pytestmark = pytest.mark.gpu

ZTOL = 1e-6


def _relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture(scope="module")
def strip():
    m = M.Mesh(M.microstrip_spec())
    _, Zr = oracle.direct_solve(m, [1e9], refine=True)
    return m, Zr


@pytest.fixture(scope="module")
def manh():
    m = M.Mesh(M.manhattan_spec((24, 20, 10), n_ports=3, nfreq=3, f_range=(1e8, 1e10), seed=11))
    fs = m.freqs()
    X, Zr = oracle.direct_solve(m, fs, refine=True)
    return m, fs, X, Zr


@pytest.mark.parametrize("prec", [capi.GP_PREC_FP64, capi.GP_PREC_MIXED])
def test_microstrip_z(strip, prec):
    m, Zr = strip
    print(f"[solve] microstrip prec={prec}", flush=True)
    lay = Layout(m)
    Z, st, _ = lay.solve([1e9], precision=prec, rtol=1e-8)
    print(st, Z, Zr, flush=True)
    assert st["status"] == capi.GP_OK and st["max_rel_residual"] <= 1e-8
    assert _relerr(Z, Zr) < ZTOL
    lay.close()


@pytest.mark.parametrize("prec,batch", [(capi.GP_PREC_FP64, 1), (capi.GP_PREC_MIXED, 8), (capi.GP_PREC_MIXED, 4)])
def test_manhattan_z_and_x(manh, prec, batch):
    m, fs, Xr, Zr = manh
    print(f"[solve] manhattan prec={prec} batch={batch}", flush=True)
    lay = Layout(m)
    Z, st, X = lay.solve(fs, precision=prec, rtol=1e-8, batch=batch, want_x=True)
    print(st, flush=True)
    assert st["status"] == capi.GP_OK and st["max_rel_residual"] <= 1e-8
    for k in range(len(fs)):
        assert _relerr(Z[k], Zr[k]) < ZTOL, (k, _relerr(Z[k], Zr[k]))
    npt = m.num_ports
    for k in range(len(fs)):
        for p in range(npt):
            x, xr = X[k * npt + p], Xr[k][p]
            assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-6
    lay.close()


@pytest.mark.parametrize("prec,sweeps", [(capi.GP_PREC_MIXED, (2, 1)), (capi.GP_PREC_MIXED, (3, 2)),
                                         (capi.GP_PREC_FP64, (2, 3))])
def test_manhattan_smoothing_sweeps(manh, prec, sweeps):
    """Edge/nodal and auxiliary sweep counts other than the default one: the
    V-cycle picks its starting buffers by the parity of the sweeps so that the
    last sweep writes each level's own vector, which even and odd counts both
    exercise here."""
    m, fs, Xr, Zr = manh
    se, sa = sweeps
    print(f"[solve] manhattan prec={prec} smooth_edge={se} smooth_aux={sa}", flush=True)
    lay = Layout(m)
    Z, st, _ = lay.solve(fs, precision=prec, rtol=1e-8, batch=8 if prec == capi.GP_PREC_MIXED else 1,
                         smooth_edge=se, smooth_aux=sa)
    assert st["status"] == capi.GP_OK and st["max_rel_residual"] <= 1e-8
    for k in range(len(fs)):
        assert _relerr(Z[k], Zr[k]) < ZTOL, (k, _relerr(Z[k], Zr[k]))
    lay.close()


def test_iteration_cap_reports_nonconvergence(manh):
    """A column cap far below what the solve needs: the call returns
    GP_ERR_NOCONV, every column's report says not converged with its true
    residual above the tolerance, and the same layout then solves normally."""
    m, fs, Xr, Zr = manh
    print("[solve] manhattan with max_iters 3", flush=True)
    lay = Layout(m)
    Z, st, _ = lay.solve(fs, precision=capi.GP_PREC_MIXED, rtol=1e-8, batch=4, max_iters=3)
    assert st["status"] == capi.GP_ERR_NOCONV
    assert st["converged"] < st["columns"] == len(fs) * m.num_ports
    assert st["max_iters"] <= 3 + 16          # stops at the first restart past the cap
    assert max(st["rel_residual"]) > 1e-8
    Z, st, _ = lay.solve(fs, precision=capi.GP_PREC_MIXED, rtol=1e-8, batch=4)
    assert st["status"] == capi.GP_OK
    for k in range(len(fs)):
        assert _relerr(Z[k], Zr[k]) < ZTOL
    lay.close()


@pytest.mark.parametrize("prec", [capi.GP_PREC_FP64, capi.GP_PREC_MIXED])
def test_parallel_plate_closed_form_gpu(prec):
    """The plate over two dielectric layers on the device: Z at 1 GHz is
    1 / (i w C) with C = eps0 area / (d1 / eps1 + d2 / eps2) to 2e-6 (the
    area from the reference's dual widths), in both precisions."""
    from test_oracle import _plate_capacitance
    cells = (24, 20, 12)
    m = M.Mesh(M.plate_spec(cells))
    C = _plate_capacitance(m, cells)
    print(f"[solve] parallel plate {cells} prec={prec}", flush=True)
    lay = Layout(m)
    Z, st, _ = lay.solve([1e9], precision=prec, rtol=1e-8)
    lay.close()
    zc = 1 / (1j * 2 * np.pi * 1e9 * C)
    err = abs(Z[0][0].imag - zc.imag) / abs(zc)
    print(f"[solve] Z {Z[0][0]:.8g} closed form {zc:.8g} rel {err:.2e}", flush=True)
    assert st["status"] == capi.GP_OK and err < 2e-6


def test_autotune_then_solve(manh):
    """gp_autotune measures the mixed path at 8, 16 and 32 slots and at three
    lanes-per-row caps, keeps the fastest per column-iteration, and solves
    with batch 0 then use that choice and still meet the fixtures."""
    m, fs, Xr, Zr = manh
    lay = Layout(m)
    print("[solve] autotune on the 24x20x10 Manhattan", flush=True)
    info = lay.autotune()
    print(f"[solve] autotune {info}", flush=True)
    assert info["slots"] in (8, 16, 32) and info["lanes_cap"] in (16, 4, 2)
    q = (8, 16, 32).index(info["slots"])
    assert info["slot_us"][q] > 0 and min(u for u in info["slot_us"] if u > 0) == info["slot_us"][q]
    assert min(info["lanes_us"]) > 0
    Z, st, _ = lay.solve(fs, precision=capi.GP_PREC_MIXED, rtol=1e-8, batch=0)
    assert st["status"] == capi.GP_OK and st["max_rel_residual"] <= 1e-8
    for k in range(len(fs)):
        assert _relerr(Z[k], Zr[k]) < ZTOL
    lay.close()


def test_graphs_replay_same_bits(manh):
    """The preconditioner replayed from captured HIP graphs gives the same
    bits as the same launches issued directly: on a first call (capture), on
    a second call (replay), and across a coarse-table re-allocation in the
    middle of a solve (20 distinct frequencies outgrow the table's first 16
    entries, so the graphs are captured again)."""
    m, fs, Xr, Zr = manh
    f20 = np.logspace(8, 10, 20)
    print("[solve] preconditioner graphs vs direct launches", flush=True)
    a = Layout(m)
    Za, sa, _ = a.solve(f20, ports=[0], precision=capi.GP_PREC_MIXED, batch=4, graphs=0)
    a.close()
    b = Layout(m)
    Zb, sb, _ = b.solve(f20, ports=[0], precision=capi.GP_PREC_MIXED, batch=4, graphs=1)
    i1 = b.hierarchy_info()
    Zc, sc, _ = b.solve(f20, ports=[0], precision=capi.GP_PREC_MIXED, batch=4, graphs=1)
    i2 = b.hierarchy_info()
    b.close()
    print(f"[solve] graphs captured {i2['graph_captures']} replayed {i2['graph_replays']}", flush=True)
    # the graphs were used: captured on the first call and replayed on both
    assert i1["graph_failed"] == 0 and i1["graph_captures"] > 0 and i1["graph_replays"] > 0
    assert i2["graph_replays"] > i1["graph_replays"]
    assert sa["status"] == sb["status"] == sc["status"] == capi.GP_OK
    assert np.array_equal(Za, Zb) and np.array_equal(Zb, Zc)
    assert sa["iters"] == sb["iters"] == sc["iters"]


@pytest.mark.parametrize("batch,graphs,sweeps", [(8, 0, 0), (16, 0, 0), (32, 0, 0), (16, 1, 0), (16, 0, 2)])
def test_edge_march_same_bits(manh, batch, graphs, sweeps):
    """The level-0 edge smoother with u staged plane by plane in LDS
    (edge_lds=1) gives the same bits as the kernel that gathers u through the
    caches (edge_lds=0): it sums each row's entries in the same order with the
    same operations.  8 slots take 8-slot workgroups, 16 one 16-slot group, 32
    two groups; graphs=1 replays the march from captured graphs; two sweeps
    add the Jacobi sweeps that gather K u."""
    m, fs, Xr, Zr = manh
    print(f"[solve] edge march vs gather, batch {batch} graphs {graphs} sweeps {sweeps}", flush=True)
    lay = Layout(m)
    kw = dict(precision=capi.GP_PREC_MIXED, batch=batch, graphs=graphs, smooth_edge=sweeps)
    Za, sa, _ = lay.solve(fs, edge_lds=0, **kw)
    Zb, sb, _ = lay.solve(fs, edge_lds=1, **kw)
    info = lay.hierarchy_info()
    lay.close()
    print(f"[solve] iterations {sa['iters']} / {sb['iters']}, max |dZ| {np.abs(Za - Zb).max():.3e}, "
          f"block rows {info['march_block_rows']}, outside {info['march_outside']} of {info['march_entries']}",
          flush=True)
    assert info["march_block_rows"] > 0
    assert 0 < info["march_outside"] < 0.3 * info["march_entries"]
    assert sa["status"] == sb["status"] == capi.GP_OK
    assert sa["iters"] == sb["iters"]
    assert np.array_equal(Za, Zb)
    for k in range(len(fs)):
        assert _relerr(Zb[k], Zr[k]) < ZTOL


def test_edge_march_same_bits_larger_mesh():
    """The same on a mesh of 5 x 5 tiles and 13 planes (two runs of planes
    per tile half), all 8 ports at two frequencies."""
    m = M.Mesh(M.manhattan_spec((64, 64, 12), 8, 10, (1e7, 1e10), 2001))
    fs = m.freqs()[[0, 9]]
    print("[solve] edge march vs gather, 64x64x12", flush=True)
    lay = Layout(m)
    Za, sa, _ = lay.solve(fs, precision=capi.GP_PREC_MIXED, batch=16, edge_lds=0)
    Zb, sb, _ = lay.solve(fs, precision=capi.GP_PREC_MIXED, batch=16, edge_lds=1)
    info = lay.hierarchy_info()
    lay.close()
    print(f"[solve] iterations {sa['iters']} / {sb['iters']}, max |dZ| {np.abs(Za - Zb).max():.3e}, "
          f"block rows {info['march_block_rows']}, outside {info['march_outside']} of {info['march_entries']}",
          flush=True)
    assert info["march_block_rows"] > 0
    assert sa["iters"] == sb["iters"]
    assert np.array_equal(Za, Zb)

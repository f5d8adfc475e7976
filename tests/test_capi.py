"""C-ABI checks that need no GPU: every declared symbol is exported, and the
host-side services agree with the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from gds2para_amd import _capi as capi
from gds2para_amd import mesh as M

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "gds2para.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gp_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported(lib):
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    bound = {name for name, _, _ in capi.SIGNATURES}
    assert set(syms) == bound, set(syms) ^ bound


def test_sizes_and_validation(microstrip, lib):
    assert lib.gp_mesh_validate(microstrip.ptr) == capi.GP_OK
    s = microstrip.sizes
    assert (s.n_edge, s.n_unknowns, s.nnz_s, s.n_node) == (32568, 30088, 368488, 11529)
    bad = capi.GpMesh.from_buffer_copy(microstrip.c)
    bad.nx = 1
    assert lib.gp_mesh_validate(C.byref(bad)) == capi.GP_ERR_ARG
    assert b"2 nodes" in lib.gp_last_error()


def test_edge_map_and_freqs(small_manhattan):
    m = small_manhattan
    mp, mr = m.edge_map()
    bd = np.ctypeslib.as_array(m.c.bd_edges, shape=(m.c.num_bd_edges,))
    assert np.all(mp[bd] == -1)
    kept = np.setdiff1d(np.arange(m.n_edge), bd)
    assert np.array_equal(mr, kept) and np.array_equal(mp[kept], np.arange(len(kept)))
    fs = m.freqs()
    assert fs[0] == 1e8 and abs(fs[-1] - 1e10) < 1e-3 and np.all(np.diff(fs) > 0)
    assert all(fs[k] == oracle.freq(m, k) for k in range(len(fs)))


def test_extract_z_matches_oracle(small_manhattan, lib):
    m = small_manhattan
    rng = np.random.default_rng(0)
    x = rng.standard_normal(m.n) + 1j * rng.standard_normal(m.n)
    npt = m.num_ports
    for src in range(npt):
        zo = oracle.extract_z(m, x, src)
        buf = np.zeros(2 * npt * npt)
        xi = np.ascontiguousarray(np.stack([x.real, x.imag], 1).ravel())
        assert lib.gp_extract_z(m.ptr, xi.ctypes.data_as(C.POINTER(C.c_double)), src,
                                buf.ctypes.data_as(C.POINTER(C.c_double))) == 0
        assert np.array_equal(buf[0::2] + 1j * buf[1::2], zo)


def test_generator_is_deterministic():
    a = M.Mesh(M.manhattan_spec((40, 40, 12), n_ports=4, seed=7))
    b = M.Mesh(M.manhattan_spec((40, 40, 12), n_ports=4, seed=7))
    c = M.Mesh(M.manhattan_spec((40, 40, 12), n_ports=4, seed=8))
    assert np.array_equal(a.mark_edge(), b.mark_edge())
    assert not np.array_equal(a.mark_edge(), c.mark_edge())
    assert a.c.num_cdt > 4
    # every port runs from the ground plane up through dielectric to a non-ground conductor
    me = a.mark_edge()
    for k in range(a.c.side_edge_ptr[a.c.port_side_ptr[a.num_ports]]):
        assert me[a.c.side_edges[k]] == 0


def test_layout_create_rejects_bad_mesh(microstrip, lib):
    bad = capi.GpMesh.from_buffer_copy(microstrip.c)
    bad.num_bd_edges = 0  # port edges stay valid; validation passes the boundary check
    h = C.c_void_p()
    bad2 = capi.GpMesh.from_buffer_copy(microstrip.c)
    bad2.ny = 1
    assert lib.gp_layout_create(C.byref(bad2), C.byref(h)) == capi.GP_ERR_ARG
    assert not h.value

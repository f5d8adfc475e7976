"""Host-side mesh objects: synthetic layouts and numpy views of a gp_mesh.

The reference builds its mesh in meshAndMark/portSet (src/mesh.cpp:5-1594) from
GDSII + sim_input.  Here the seeded generator of the C library produces the
same contract (grid, conductor marks, PEC sets, port sides) for the structured
workloads the benchmarks use.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _capi as capi

UM = 1e-6


@dataclass
class GenSpec:
    """Parameters of the seeded layout generator (mirrors gp_gen_params)."""
    kind: int = capi.GP_GEN_MANHATTAN
    cells: tuple = (200, 200, 24)
    h: tuple = (1.0 * UM, 1.0 * UM, 0.5 * UM)
    grade: float = 0.2
    n_metal: int = 4
    fill: float = 0.30
    wmin: int = 1
    wmax: int = 8
    via_density: float = 0.02
    eps: tuple = (3.0, 4.5)
    sigma: float = 5.8e7
    n_ports: int = 8
    nfreq: int = 10
    freq_scale: int = 0
    f_range: tuple = (1e7, 1e10)
    seed: int = 2000

    def to_c(self) -> capi.GpGenParams:
        p = capi.GpGenParams()
        p.kind = self.kind
        p.nx_cells, p.ny_cells, p.nz_cells = self.cells
        p.dx, p.dy, p.dz = self.h
        p.grade = self.grade
        p.n_metal = self.n_metal
        p.fill = self.fill
        p.wmin, p.wmax = self.wmin, self.wmax
        p.via_density = self.via_density
        p.eps_lo, p.eps_hi = self.eps
        p.sigma = self.sigma
        p.n_ports = self.n_ports
        p.nfreq, p.freq_scale = self.nfreq, self.freq_scale
        p.f_start, p.f_end = self.f_range
        p.seed = self.seed
        return p


# The workloads: the reference's CPU-runnable microstrip, then the structured
# Manhattan layouts at the three sizes the measurements use.
def microstrip_spec() -> GenSpec:
    """60x20x8 cells, 100x2 um copper strip over ground, 1 z-port, 1 GHz."""
    return GenSpec(kind=capi.GP_GEN_MICROSTRIP, cells=(60, 20, 8), h=(2 * UM, 1 * UM, 0.5 * UM),
                   grade=0.0, eps=(3.9, 4.4), n_ports=1, nfreq=1, f_range=(1e9, 1e9), seed=1000)


def plate_spec(cells=(24, 20, 12), eps=(3.9, 4.4), f=1e7) -> GenSpec:
    """A plate over the whole domain above two dielectric layers (eps[0] for
    the lower, eps[1] for the upper), one z-port; uniform 1 x 1 x 0.5 um cells."""
    return GenSpec(kind=capi.GP_GEN_PLATE, cells=cells, h=(1 * UM, 1 * UM, 0.5 * UM), grade=0.0, eps=eps,
                   n_ports=1, nfreq=1, f_range=(f, f), seed=1)


def manhattan_spec(cells=(200, 200, 24), n_ports=8, nfreq=10, f_range=(1e7, 1e10), seed=2000) -> GenSpec:
    return GenSpec(cells=cells, n_ports=n_ports, nfreq=nfreq, f_range=f_range, seed=seed)


WORKLOADS = {
    "microstrip_60x20x8": microstrip_spec,
    "manhattan_200x200x24": lambda: manhattan_spec((200, 200, 24), 8, 10, (1e7, 1e10), 2000),
    "manhattan_400x400x32": lambda: manhattan_spec((400, 400, 32), 16, 20, (1e7, 1e10), 3000),
    "manhattan_300x300x32": lambda: manhattan_spec((300, 300, 32), 64, 32, (1e8, 5e10), 4000),
    # the 200x200x24 and 300x300x32 generators scaled to 48x48x12 cells, 4 ports
    "manhattan_48x48x12": lambda: manhattan_spec((48, 48, 12), 4, 4, (1e7, 1e10), 2000),
    "manhattan_48x48x12_hi": lambda: manhattan_spec((48, 48, 12), 4, 4, (1e8, 5e10), 4000),
}


class Mesh:
    """A generated mesh owned by the C library, with numpy views of its arrays."""

    def __init__(self, spec: GenSpec):
        lib = capi.load()
        self.spec = spec
        self._params = spec.to_c()
        h = C.c_void_p()
        capi.check(lib.gp_generate(C.byref(self._params), C.byref(h)), "gp_generate")
        self._h = h
        self.c = lib.gp_owned_mesh_view(h).contents
        self.ptr = C.pointer(self.c)
        sz = capi.GpMeshSizes()
        capi.check(lib.gp_mesh_sizes_of(self.ptr, C.byref(sz)), "gp_mesh_sizes_of")
        self.sizes = sz

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                capi.load().gp_owned_mesh_free(self._h)
                self._h = None
        except Exception:
            pass

    # numpy views ---------------------------------------------------------
    def _arr(self, name, n, dt):
        return np.ctypeslib.as_array(getattr(self.c, name), shape=(n,)).astype(dt, copy=False)

    @property
    def n(self):
        return int(self.sizes.n_unknowns)

    @property
    def n_edge(self):
        return int(self.sizes.n_edge)

    @property
    def num_ports(self):
        return int(self.c.num_ports)

    def coords(self):
        c = self.c
        return (self._arr("xn", c.nx, np.float64), self._arr("yn", c.ny, np.float64),
                self._arr("zn", c.nz, np.float64))

    def mark_edge(self):
        return self._arr("mark_edge", self.n_edge, np.int32)

    def mark_node(self):
        return self._arr("mark_node", int(self.sizes.n_node), np.int32)

    def freqs(self):
        lib = capi.load()
        return np.array([lib.gp_freq_of(self.ptr, k) for k in range(self.c.nfreq)])

    def edge_map(self):
        lib = capi.load()
        m = np.empty(self.n_edge, np.int64)
        r = np.empty(self.n, np.int64)
        capi.check(lib.gp_edge_map(self.ptr, m.ctypes.data_as(C.POINTER(C.c_int64)),
                                   r.ctypes.data_as(C.POINTER(C.c_int64))), "gp_edge_map")
        return m, r

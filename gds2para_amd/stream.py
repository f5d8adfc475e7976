"""Seeded request stream of the serving workload.

24 distinct layouts of the seeded Manhattan generator, sizes log-uniform
between 5e4 and 2e7 edges, each reused by many requests.  A request asks for
the Z columns of 1-64 ports at 1-40 frequencies of one layout; both counts are
log-uniform integers (many small requests, a few large ones), the ports are
distinct ports of the layout and the frequencies distinct points of the
layout's 40-point log grid from 10 MHz to 10 GHz (so a layout's coarse-level
cache never needs more than 40 entries).  Arrivals are a Poisson stream.

The reference has no serving path; this generator only feeds serve.Server.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import mesh as M


@dataclass
class StreamSpec:
    n_layouts: int = 24
    edges: tuple = (5e4, 2e7)       # log-uniform layout sizes (edges)
    max_ports: int = 64
    max_freqs: int = 40
    f_range: tuple = (1e7, 1e10)
    n_requests: int = 1000
    rate_hz: float = 1.0             # Poisson arrival rate
    seed: int = 5000


def _cells_for(edges, nz):
    """Square xy grid with nz cells in z holding about `edges` edges
    (N_edge ~ 3 (nx+1)(ny+1)(nz+1))."""
    c = int(round(math.sqrt(edges / (3.0 * (nz + 1))))) - 1
    return max(c, 16)


def layouts(spec: StreamSpec):
    """key -> GenSpec of the stream's layouts (64 ports each, 4 metal layers)."""
    rng = np.random.default_rng(spec.seed)
    lo, hi = math.log(spec.edges[0]), math.log(spec.edges[1])
    out = {}
    for i in range(spec.n_layouts):
        e = math.exp(rng.uniform(lo, hi))
        nz = int(rng.integers(10, 33))
        c = _cells_for(e, nz)
        key = f"L{i:02d}_{c}x{c}x{nz}"
        out[key] = M.manhattan_spec((c, c, nz), n_ports=spec.max_ports, nfreq=spec.max_freqs,
                                    f_range=spec.f_range, seed=spec.seed + 1 + i)
    return out


def _log_uniform_int(rng, hi):
    return int(min(hi, max(1, math.floor(math.exp(rng.uniform(0.0, math.log(hi + 1)))))))


def requests(spec: StreamSpec, keys=None):
    """[(layout key, ports, freqs Hz)] and their arrival offsets (s)."""
    keys = list(layouts(spec)) if keys is None else list(keys)
    rng = np.random.default_rng(spec.seed + 100003)
    grid = np.logspace(math.log10(spec.f_range[0]), math.log10(spec.f_range[1]), spec.max_freqs)
    out = []
    for _ in range(spec.n_requests):
        key = keys[int(rng.integers(0, len(keys)))]
        npq = _log_uniform_int(rng, spec.max_ports)
        nfq = _log_uniform_int(rng, spec.max_freqs)
        ports = sorted(rng.choice(spec.max_ports, size=npq, replace=False).tolist())
        freqs = sorted(grid[rng.choice(spec.max_freqs, size=nfq, replace=False)].tolist())
        out.append((key, ports, freqs))
    offsets = np.cumsum(rng.exponential(1.0 / spec.rate_hz, spec.n_requests))
    return out, offsets


def edges_of(gen: M.GenSpec) -> int:
    nx, ny, nz = gen.cells
    es = ny * (nx + 1) + nx * (ny + 1)
    ev = (nx + 1) * (ny + 1)
    return es * (nz + 1) + ev * nz

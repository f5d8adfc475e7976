"""Benchmark: port x frequency full-wave solves per second on one MI355X.

Workload: the seeded Manhattan layout at 200x200x24 cells (~2.9M edges),
4 metal layers, 8 ports, 10 log-spaced frequencies 1e7..1e10 Hz (80 columns).
A step is one pass of the hot path over that batch: every (port, freq) column
solved to a 1e-8 FP64 relative residual by the mixed-precision batched path
(FP32 multigrid preconditioner, FP64 FGMRES with exact residuals), with the
stiffness, V0 basis and hierarchy resident in HBM (assembled before timing,
as the reference assembles once per layout before its frequency loop).

Besides the headline rate the line reports:
  fp64_single     the plain FP64 single-RHS path on a sample of the columns,
                  and the speedup of the mixed batched path over it;
  latency         p50/p99 of the per-column service time in the timed steps;
  serving         a seeded Poisson stream of requests over three layouts served
                  with continuous batching: per-request latency p50/p99;
  roofline        achieved algorithmic bandwidth of the dominant kernel (the
                  level-0 edge smoother), timed with HIP events on its stream
                  over the timed region, against 8 TB/s;
  cpu_baseline    the oracle's FP64 direct solve (the restatement of the
                  reference's PARDISO path) on a bounded sample of the same
                  generator, on the host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "port×freq solves/sec at 1e-8 resid, mixed+batched vs FP64; p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def cpu_baseline(cells=(32, 32, 12)):
    """Oracle direct FP64 solve (assembly + SuperLU) of one column on a small
    instance of the same generator; solves/s on one host core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from gds2para_amd import mesh as M
    m = M.Mesh(M.manhattan_spec(cells, n_ports=8, nfreq=10, f_range=(1e7, 1e10), seed=2000))
    t = time.perf_counter()
    S = oracle.stiffness_csr(m)
    oracle.direct_solve(m, [1e9], ports=[0], S=S, refine=False)
    dt = time.perf_counter() - t
    return {"value": 1.0 / dt, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"1 column (port 0, 1 GHz) of a {cells[0]}x{cells[1]}x{cells[2]} instance of the same "
                      f"generator ({m.n} edges): oracle assembly + SuperLU FP64 direct solve, {dt:.1f} s"}


def serving_leg(n_requests, rate_hz, seed=5):
    """Seeded Poisson stream of requests over three resident layouts of the
    same generator (150k, 700k and 2.9M edges), served with continuous
    batching under a 260 GB budget; per-request latency p50/p99."""
    from gds2para_amd import mesh as M
    from gds2para_amd.serve import GpuBackend, Request, Server, poisson_arrivals, run_stream
    specs = {"manhattan_64x64x12": M.manhattan_spec((64, 64, 12), 8, 10, (1e7, 1e10), 2001),
             "manhattan_120x120x16": M.manhattan_spec((120, 120, 16), 8, 10, (1e7, 1e10), 2002),
             "manhattan_200x200x24": M.WORKLOADS["manhattan_200x200x24"]()}
    keys = list(specs)
    rng = np.random.default_rng(seed)
    grid = np.logspace(7, 10, 10)
    reqs = []
    for _ in range(n_requests):
        key = keys[int(rng.choice(3, p=[0.4, 0.35, 0.25]))]
        ports = sorted(rng.choice(8, size=int(rng.integers(1, 9)), replace=False).tolist())
        freqs = sorted(rng.choice(grid, size=int(rng.integers(1, 5)), replace=False).tolist())
        reqs.append(Request(key, ports, freqs))
    srv = Server(GpuBackend(specs, slots=16), budget_bytes=260 << 30)
    for k in keys:                       # layouts resident before the stream starts
        srv.submit(Request(k, [0], [1e9]))
    srv.wait()
    t = time.perf_counter()
    run_stream(srv, reqs, poisson_arrivals(rate_hz, len(reqs), seed=seed))
    dt = time.perf_counter() - t
    srv.close()
    lat = [r.latency for r in reqs]
    cols = sum(len(r.ports) * len(r.freqs) for r in reqs)
    conv = sum(1 for r in reqs for rep in r.reports.values() if rep["converged"] == 1)
    return {"requests": len(reqs), "columns": cols, "converged": f"{conv}/{cols}", "layouts": keys,
            "arrival_rate_hz": rate_hz, "slots": 16, "latency_s": {"p50": float(np.percentile(lat, 50)),
                                                                  "p99": float(np.percentile(lat, 99))},
            "columns_per_s": cols / dt, "wall_s": dt}


def serving_leg_guarded(n_requests, rate_hz, limit_s=150):
    """serving_leg in a child process with a time limit: a stream that stalls
    is stopped and reported as such instead of holding up the benchmark line."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--serving-only", "--serve-requests", str(n_requests),
           "--serve-rate", str(rate_hz)]
    try:
        p = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=limit_s)
    except subprocess.TimeoutExpired:
        return {"error": f"the serving stream did not finish within {limit_s} s and was stopped"}
    if p.returncode != 0:
        return {"error": f"the serving process ended with status {p.returncode}"}
    return json.loads(p.stdout.strip().splitlines()[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="manhattan_200x200x24")
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--restart", type=int, default=16)
    ap.add_argument("--max-iters", type=int, default=1500)
    ap.add_argument("--fp64-sample", type=int, default=3, help="columns of the FP64 single-RHS sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=24, help="requests in the serving leg (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=2.0, help="Poisson arrival rate of the serving leg (1/s)")
    ap.add_argument("--smooth-edge", type=int, default=0, help="edge/nodal smoothing sweeps per level (0: library default)")
    ap.add_argument("--smooth-aux", type=int, default=0, help="auxiliary AMG sweeps per level (0: library default)")
    ap.add_argument("--sell-min-rows", type=int, default=0,
                    help="operators with at least this many rows use the sliced-ELL kernels (0: library default)")
    ap.add_argument("--basis64", action="store_true", help="FP64 Krylov basis on the mixed path (default FP32)")
    ap.add_argument("--serving-only", action="store_true", help=argparse.SUPPRESS)
    a = ap.parse_args()
    if a.serving_only:                   # the child process of serving_leg_guarded
        print(json.dumps(serving_leg(a.serve_requests, a.serve_rate)), flush=True)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {a.gpus} is not supported "
              "(the solve path is single-device)", file=sys.stderr)
        sys.exit(2)

    from gds2para_amd import _capi as capi
    from gds2para_amd import mesh as M
    from gds2para_amd.layout import Layout

    spec = M.WORKLOADS[a.workload]()
    t0 = time.perf_counter()
    m = M.Mesh(spec)
    lay = Layout(m)
    lay.assemble()
    lay.build_v0()
    lay.prepare(capi.GP_PREC_MIXED)
    lay.sync()
    setup_s = time.perf_counter() - t0
    freqs = m.freqs()
    ports = list(range(m.num_ports))
    ncol = len(freqs) * len(ports)
    kw = dict(ports=ports, precision=capi.GP_PREC_MIXED, rtol=1e-8, batch=a.batch, restart=a.restart,
              max_iters=a.max_iters, sell_min_rows=a.sell_min_rows, smooth_edge=a.smooth_edge,
              smooth_aux=a.smooth_aux, basis=1 if a.basis64 else 0)

    def step():
        return lay.solve(freqs, **kw)

    # plain-copy reference point (the kernel the HBM counters are scaled by)
    import ctypes as C
    cms = C.c_double()
    capi.check(lay.lib.gp_bench_copy(1 << 30, 5, C.byref(cms)), "gp_bench_copy")
    copy_gbs = 2.0 * (1 << 30) / (cms.value * 1e-3) / 1e9

    for w in range(a.warmup):
        print(f"[bench] warmup step {w + 1}/{a.warmup}", file=sys.stderr, flush=True)
        step()
    lay.sync()
    lay.kernel_timer(True)
    lat, iters, conv, res_max = [], [], 0, 0.0
    Zm = None
    t = time.perf_counter()
    for k in range(a.steps):
        Zm, st, _ = step()
        lat += [d - s for d, s in zip(st["done_ms"], st["start_ms"])]
        iters += st["iters"]
        conv += st["converged"]
        res_max = max(res_max, st["max_rel_residual"])
        print(f"[bench] step {k + 1}/{a.steps}: {st['solve_ms'] / 1e3:.2f} s, converged {st['converged']}/{ncol}, "
              f"max iters {st['max_iters']}", file=sys.stderr, flush=True)
    lay.sync()
    elapsed = time.perf_counter() - t
    launches, kms, kbytes = lay.kernel_timer_read()
    lay.kernel_timer(False)
    value = a.steps * ncol / elapsed

    # plain FP64 single-RHS path on a sample of the columns (port 0, spread frequencies)
    fp64 = None
    zdiff = None
    if a.fp64_sample > 0:
        fidx = np.linspace(0, len(freqs) - 1, a.fp64_sample).round().astype(int)
        fs = freqs[fidx]
        lay.prepare(capi.GP_PREC_FP64)
        t = time.perf_counter()
        Zd, std, _ = lay.solve(fs, ports=[0], precision=capi.GP_PREC_FP64, rtol=1e-8, batch=1, restart=a.restart,
                               max_iters=a.max_iters)
        lay.sync()
        t64 = time.perf_counter() - t
        npt = m.num_ports
        zm = Zm[fidx].reshape(len(fs), npt, npt)[:, 0, :]   # column of source port 0: index resp + np*0
        zd = Zd.reshape(len(fs), npt, npt)[:, 0, :]
        zdiff = float(np.abs(zm - zd).max() / np.abs(zd).max())
        fp64 = {"value": len(fs) / t64, "unit": "solves/s", "columns": len(fs),
                "freqs_hz": [float(f) for f in fs], "port": 0, "iters": std["iters"],
                "converged": f"{std['converged']}/{len(fs)}"}
    # The FP64 complex SpMV y = A(w) x (gp_apply_fullwave's kernel) on resident
    # vectors, priced on the CSR-equivalent bytes of a plain SpMV:
    # 12 nnz + 4 (n+1) + k (16 n read + 16 n written) + 16 n (complex diagonal)
    spmv = {}
    n_u, nnz_s = m.n, int(m.sizes.nnz_s)
    for kk in (1, 16):
        ms = C.c_double()
        capi.check(lay.lib.gp_bench_spmv(lay.h, kk, capi.GP_PREC_FP64, 50, C.byref(ms)), "gp_bench_spmv")
        byts = 12.0 * nnz_s + 4.0 * (n_u + 1) + kk * 32.0 * n_u + 16.0 * n_u
        gbs = byts / (ms.value * 1e-3) / 1e9
        spmv[f"k{kk}"] = {"avg_us": ms.value * 1e3, "bytes": byts, "achieved_GBs": gbs, "frac": gbs / 8000.0}
    spmv["kernel"] = ("k_amul_sell (FP64 complex A(w) X, S in sliced ELL: the reference's row order for one "
                      "column, 16 x 16-cell tile order for 16; average of 50 launches)")
    # HBM traffic of the dominant kernel, from the committed counter profile of this command
    traffic, traffic_src = None, None
    tp = os.path.join(ROOT, "profiles", "edge0_traffic.json")
    if os.path.exists(tp):
        with open(tp) as f:
            tj = json.load(f)
        traffic, traffic_src = tj.get("bytes_per_launch"), "profiles/edge0_traffic.json"

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex fp64 outer (residual, products, iterate) / complex fp32 preconditioner and Krylov basis storage"
                 if not a.basis64 else "complex fp64 outer / complex fp32 preconditioner",
        "data": "synthetic (seeded Manhattan layout generator)",
        "config": {"workload": a.workload, "cells": list(spec.cells), "edges": int(m.n), "ports": len(ports),
                   "freqs": len(freqs), "f_range_hz": list(spec.f_range), "columns_per_step": ncol,
                   "batch_slots": a.batch or "auto", "restart": a.restart, "rtol": 1e-8,
                   "krylov_basis": "fp64" if a.basis64 else "fp32"},
        "converged": f"{conv}/{a.steps * ncol}",
        "max_rel_residual": res_max,
        "iters": {"p50": float(np.percentile(iters, 50)), "max": int(max(iters))},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "setup_s": setup_s,
        "fp64_single": fp64,
        "speedup_vs_fp64_single": value / fp64["value"] if fp64 else None,
        "z_rel_diff_mixed_vs_fp64": zdiff,
        "copy_probe_GBs": copy_gbs,
        "spmv": spmv,
        "roofline": {"bound": "hbm", "kernel": "level-0 edge smoother (k_edge_sell<float, 4, 4>: 4 lanes per row, 4 slots per lane)",
                     "achieved": kbytes / (kms * 1e-3) / 1e9 if kms > 0 else None, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s",
                     "frac": (kbytes / (kms * 1e-3) / 1e9) / HBM_PEAK_GBS if kms > 0 else None,
                     "traffic": traffic, "traffic_source": traffic_src, "launches": launches,
                     "avg_launch_us": kms / launches * 1e3 if launches else None,
                     "bytes_per_launch": kbytes / launches if launches else None},
    }
    lay.close()
    if a.serve_requests > 0:
        print(f"[bench] serving leg: {a.serve_requests} requests", file=sys.stderr, flush=True)
        line["serving"] = serving_leg_guarded(a.serve_requests, a.serve_rate)
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

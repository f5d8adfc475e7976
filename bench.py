"""Benchmark: port x frequency full-wave solves per second on one MI355X.

Workload (headline): the seeded Manhattan layout at 400x400x32 cells (15.4M
unknowns), 4 metal layers, 16 ports, 20 log-spaced frequencies 1e7..1e10 Hz
(320 columns), the largest of the configurations that fits one GPU.  A step
is the 16 ports at one frequency, solved to a 1e-8 FP64 relative residual by
the mixed-precision batched path (FP32 multigrid preconditioner and Krylov
basis storage, FP64 FGMRES residuals, products, sums and iterate) on 16 device
slots; timed step t takes frequency t mod 20, so 20 timed steps cover every
(port, frequency) column once.  The stiffness, V0 basis and hierarchy are
resident in HBM (assembled before timing, as the reference assembles once per
layout before its frequency loop).

Besides the headline rate the line reports:
  fp64_single     the plain FP64 single-RHS path over the same frequency mix:
                  one column per frequency (port t mod 16 at frequency t), so
                  every frequency is weighted as in the timed steps, and the
                  speedup of the mixed batched path over it;
  latency         p50/p99 of the per-column service time in the timed steps;
  secondary       the same measurement on the 200x200x24 workload (8 ports x
                  10 frequencies, one step = all 80 columns);
  serving         a prefix of the seeded serving stream (24 layouts of 5e4..2e7
                  edges, 1-64 ports x 1-40 frequencies per request, Poisson
                  arrivals) served with continuous batching under a 260 GB
                  budget: per-request latency p50/p99;
  roofline        achieved algorithmic bandwidth of the dominant kernel (the
                  level-0 edge smoother), timed with HIP events on its stream
                  over the timed region, against 8 TB/s;
  cpu_baseline    MKL PARDISO (the reference's own direct solver, called as
                  reference() calls it) on the host cores, on a 48x48x12
                  instance of the same generator, and its Z error against the
                  oracle's refined solve.

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


def _median_time(fn, reps):
    fn()                                  # warm-up
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts)), ts


def cpu_baseline():
    """MKL PARDISO, mtype 13 with the reference's iparm (src/generateStiff.cpp:
    638-672), on the host cores: the 60x20x8 microstrip (1 port, 1 GHz; median
    of 5 after a warm-up) and a 48x48x12 instance of the 200x200x24 generator
    (4 ports at 1 GHz as one 4-RHS factorization, as reference() solves all
    ports of a frequency; median of 3 after a warm-up).  value = columns per
    second of the latter.  Falls back to SuperLU on one core, and says so, if
    the MKL runtime is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from gds2para_amd import mesh as M
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    threads = min(threads, os.cpu_count() or threads)
    out = {"unit": "solves/s"}
    try:
        import pardiso
        pardiso.load()
        pardiso.set_threads(threads)
        out.update(kind="pardiso", cores=threads, library=pardiso.version())
    except OSError as e:
        pardiso = None
        out.update(kind="port", cores=1, library=f"SciPy SuperLU (MKL runtime unavailable: {e})")
    res = {}
    for name, spec, ports, reps in (("microstrip_60x20x8", M.microstrip_spec(), [0], 5),
                                    ("manhattan_48x48x12", M.WORKLOADS["manhattan_48x48x12"](), [0, 1, 2, 3], 3)):
        m = M.Mesh(spec)
        S = oracle.stiffness_csr(m)
        f = 1e9
        A = oracle.system(m, f, S)
        B = np.stack([oracle.rhs(m, f, p) for p in ports], 1)
        box = {}

        def run():
            if pardiso is not None:
                box["X"], _ = pardiso.solve(A, B)
            else:
                import scipy.sparse.linalg as spla
                lu = spla.splu(A, permc_spec="COLAMD")
                box["X"] = np.stack([lu.solve(B[:, q]) for q in range(B.shape[1])], 1)
        med, ts = _median_time(run, reps)
        Z = np.zeros(m.num_ports ** 2, complex)
        for q, p in enumerate(ports):
            Z = oracle.extract_z(m, box["X"][:, q], p, Z)
        zerr = None
        gname = {"microstrip_60x20x8": "microstrip_60x20x8.npz",
                 "manhattan_48x48x12": "manhattan_48x48x12_lo.npz"}[name]
        gpath = os.path.join(ROOT, "tests", "golden", gname)
        if os.path.exists(gpath):
            g = np.load(gpath, allow_pickle=False)
            k = int(np.argmin(np.abs(g["freqs"] - f)))
            zerr = float(np.abs(Z - g["Z"][k]).max() / np.abs(g["Z"][k]).max())
        res[name] = {"edges": m.n, "columns": len(ports), "median_s": med, "runs_s": ts,
                     "solves_per_s": len(ports) / med, "z_rel_err_vs_refined_oracle": zerr}
    big = res["manhattan_48x48x12"]
    out["value"] = big["solves_per_s"]
    out["sample"] = (f"{out['library']} on {out['cores']} host threads: one 4-RHS factorization and solve "
                     f"(4 ports at 1 GHz) of a 48x48x12 instance of the 200x200x24 generator "
                     f"({big['edges']} unknowns), median of 3 after a warm-up: {big['median_s']:.2f} s; "
                     f"the 60x20x8 microstrip in {res['microstrip_60x20x8']['median_s']:.2f} s")
    out["cases"] = res
    return out


def serving_leg(n_requests, slots=16, python_server=False, progress=False):
    """The first n_requests of the seeded serving stream (gds2para_amd.stream:
    24 Manhattan layouts, 5e4..2e7 edges log-uniform, 64 ports and a 40-point
    10 MHz..10 GHz grid each; 1-64 ports x 1-40 frequencies per request,
    log-uniform), arriving as a Poisson stream at 0.5 requests/s and served with
    continuous batching under a 260 GB budget by the library's server
    (gp_server_*: admission, eviction and one stream thread per layout in C++),
    or by the Python scheduler serve.Server with python_server.  The layouts
    the prefix names are opened (assembly, V0 basis, hierarchy) before the
    first arrival, which is what a warm server holds; latency = completion -
    arrival per request.  progress prints each completed request to stderr."""
    from gds2para_amd import stream as ST
    from gds2para_amd.serve import CServer, GpuBackend, Request, Server, run_stream
    sp = ST.StreamSpec(n_requests=n_requests, rate_hz=0.5)
    specs = ST.layouts(sp)
    raw, offs = ST.requests(sp)
    keys = sorted({k for k, _, _ in raw})
    t_start = time.perf_counter()

    def done(r):
        if progress:
            print(f"[serve] {r.layout} {len(r.ports)} ports x {len(r.freqs)} freqs done at "
                  f"{time.perf_counter() - t_start:.1f} s{', ' + r.error if r.error else ''}", file=sys.stderr, flush=True)

    if python_server:
        srv = Server(GpuBackend({k: specs[k] for k in keys}, slots=slots), budget_bytes=260 << 30, on_done=done)
    else:
        srv = CServer(None, 260 << 30, slots=slots, layouts={k: specs[k] for k in keys}, on_done=done)
    t = time.perf_counter()
    for k in keys:                       # warm server: the named layouts resident
        srv.submit(Request(k, [0], [1e9]))
    srv.wait()
    warm_s = time.perf_counter() - t
    opened = srv.evictions
    reqs = [Request(k, p, f) for k, p, f in raw]
    t = time.perf_counter()
    run_stream(srv, reqs, offs)
    dt = time.perf_counter() - t
    evictions = srv.evictions - opened
    srv.close()
    lat = [r.latency for r in reqs]
    cols = sum(len(r.ports) * len(r.freqs) for r in reqs)
    if python_server:
        conv = sum(1 for r in reqs for rep in r.reports.values() if rep["converged"] == 1)
    else:
        conv = sum(int(r.reports.get("converged", 0)) for r in reqs)
    sizes = {k: ST.edges_of(specs[k]) for k in keys}
    return {"generator": "gds2para_amd.stream (24 layouts 5e4-2e7 edges, 1-64 ports x 1-40 freqs, Poisson)",
            "server": "serve.Server (Python)" if python_server else "gp_server (C++)",
            "requests": len(reqs), "of_stream": 1000, "columns": cols, "converged": f"{conv}/{cols}",
            "layouts": sizes, "arrival_rate_hz": sp.rate_hz, "slots": slots, "budget_GB": 260,
            "evictions": evictions, "warm_open_s": warm_s,
            "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
            "columns_per_s": cols / dt, "wall_s": dt}


def serving_leg_guarded(n_requests, limit_s=240, python_server=False):
    """serving_leg in a child process with a time limit: a stream that stalls
    is stopped and reported as such instead of holding up the benchmark line."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--serving-only", "--serve-requests", str(n_requests)]
    if python_server:
        cmd.append("--serve-python")
    try:
        p = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=limit_s)
    except subprocess.TimeoutExpired:
        return {"error": f"the serving stream did not finish within {limit_s} s and was stopped"}
    if p.returncode != 0:
        return {"error": f"the serving process ended with status {p.returncode}"}
    return json.loads(p.stdout.strip().splitlines()[-1])


def graphs_leg(cells=(64, 64, 12), reps=3):
    """Per-column wall time of one layout's 80 columns (8 ports x 10
    frequencies) with the preconditioner of each Krylov step replayed from
    captured HIP graphs and with its launches issued directly, on a small
    layout where launch overhead matters (no kernel timer: graphs are on)."""
    from gds2para_amd import _capi as capi
    from gds2para_amd import mesh as M
    from gds2para_amd.layout import Layout
    m = M.Mesh(M.manhattan_spec(cells, 8, 10, (1e7, 1e10), 2001))
    lay = Layout(m)
    lay.prepare(capi.GP_PREC_MIXED)
    fs = m.freqs()
    out = {"workload": f"manhattan_{cells[0]}x{cells[1]}x{cells[2]}", "edges": m.n, "columns": 80}
    for g in (1, 0, 1):
        lay.solve(fs, precision=capi.GP_PREC_MIXED, batch=16, graphs=g)   # capture / warm
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            Z, st, _ = lay.solve(fs, precision=capi.GP_PREC_MIXED, batch=16, graphs=g)
            ts.append(time.perf_counter() - t)
        out["graphs_ms" if g else "direct_ms"] = 1e3 * float(np.median(ts))
    info = lay.hierarchy_info()
    out["graph_captures"], out["graph_replays"] = info["graph_captures"], info["graph_replays"]
    lay.close()
    out["speedup"] = out["direct_ms"] / out["graphs_ms"]
    return out


def measure(workload, steps, warmup, a, per_frequency):
    """Timed steps of the mixed batched path on one workload, then the plain
    FP64 single-RHS path over the same frequency mix.

    per_frequency: a step is all ports at one frequency (timed step t takes
    frequency t mod nf); otherwise a step is every column of the workload."""
    import ctypes as C
    from gds2para_amd import _capi as capi
    from gds2para_amd import mesh as M
    from gds2para_amd.layout import Layout

    spec = M.WORKLOADS[workload]()
    t0 = time.perf_counter()
    m = M.Mesh(spec)
    lay = Layout(m)
    lay.assemble()
    lay.build_v0()
    lay.prepare(capi.GP_PREC_MIXED)
    lay.sync()
    setup_s = time.perf_counter() - t0
    freqs = m.freqs()
    nf, npt = len(freqs), m.num_ports
    ports = list(range(npt))
    kw = dict(ports=ports, precision=capi.GP_PREC_MIXED, rtol=a.rtol, batch=a.batch, restart=a.restart,
              max_iters=a.max_iters, sell_min_rows=a.sell_min_rows, smooth_edge=a.smooth_edge,
              smooth_aux=a.smooth_aux, basis=1 if a.basis64 else 0, graphs=1 if a.graphs else 0,
              edge_lds=1 if a.edge_lds else 0)

    def step(t):
        fs = freqs[[t % nf]] if per_frequency else freqs
        return fs, lay.solve(fs, **kw)

    tuned = None
    if a.batch == 0:
        tuned = lay.autotune(a.restart)
        print(f"[bench] {workload} autotuned: {tuned['slots']} slots, lanes cap {tuned['lanes_cap']} "
              f"({tuned['tune_ms']:.0f} ms)", file=sys.stderr, flush=True)
    for w in range(warmup):
        print(f"[bench] {workload} warmup step {w + 1}/{warmup}", file=sys.stderr, flush=True)
        step(w)
    lay.sync()
    lay.kernel_timer(True)
    lat, iters, conv, res_max = [], [], 0, 0.0
    zmix = {}                            # frequency index -> np x np block
    it_f = {}
    t = time.perf_counter()
    for k in range(steps):
        fs, (Zm, st, _) = step(k)
        lat += [d - s0 for d, s0 in zip(st["done_ms"], st["start_ms"])]
        iters += st["iters"]
        conv += st["converged"]
        res_max = max(res_max, st["max_rel_residual"])
        for q, f in enumerate(fs):
            fi = int(np.argmin(np.abs(freqs - f)))
            zmix[fi] = Zm[q]
            it_f.setdefault(f"{f:.4g}", []).extend(st["iters"][q * npt:(q + 1) * npt])
        print(f"[bench] {workload} step {k + 1}/{steps}: {st['solve_ms'] / 1e3:.2f} s, "
              f"converged {st['converged']}/{st['columns']}, max iters {st['max_iters']}", file=sys.stderr, flush=True)
    lay.sync()
    elapsed = time.perf_counter() - t
    launches, kms, kbytes = lay.kernel_timer_read()
    lay.kernel_timer(False)
    hinfo = lay.hierarchy_info()
    ncol = npt * (1 if per_frequency else nf)
    value = steps * ncol / elapsed

    # plain FP64 single-RHS path over the same frequency mix: one column per
    # frequency, ports taken in turn (port t mod np at frequency t)
    fp64, zdiff = None, None
    if a.fp64 and len(zmix) > 0:
        lay.prepare(capi.GP_PREC_FP64)
        cols = [(t % npt, t) for t in range(nf) if t in zmix]
        its, convd, tsum, dz = [], 0, 0.0, 0.0
        dz_f = {}
        for p, fi in cols:
            t1 = time.perf_counter()
            Zd, std, _ = lay.solve(freqs[[fi]], ports=[p], precision=capi.GP_PREC_FP64, rtol=a.rtol, batch=1,
                                   restart=a.restart, max_iters=a.max_iters)
            lay.sync()
            tsum += time.perf_counter() - t1
            its.append(std["iters"][0])
            convd += std["converged"]
            zd = Zd[0].reshape(npt, npt)[p]           # index resp + np*src: source p's column
            zm = zmix[fi].reshape(npt, npt)[p]
            dzf = float(np.abs(zm - zd).max() / np.abs(zd).max())
            dz = max(dz, dzf)
            dz_f[f"{freqs[fi]:.4g}"] = dzf
            print(f"[bench] {workload} FP64 column port {p} f={freqs[fi]:.4g}: {std['iters'][0]} iters",
                  file=sys.stderr, flush=True)
        zdiff = dz
        fp64 = {"value": len(cols) / tsum, "unit": "solves/s", "columns": len(cols),
                "sample": f"port t mod {npt} at frequency t, every one of the {nf} frequencies once",
                "iters": {f"{freqs[fi]:.4g}": i for (p, fi), i in zip(cols, its)},
                "converged": f"{convd}/{len(cols)}", "seconds": tsum, "z_rel_diff_per_freq": dz_f}

    spmv = {}
    n_u, nnz_s = m.n, int(m.sizes.nnz_s)
    for kk in (1, 16):
        ms = C.c_double()
        capi.check(lay.lib.gp_bench_spmv(lay.h, kk, capi.GP_PREC_FP64, 50, C.byref(ms)), "gp_bench_spmv")
        byts = 12.0 * nnz_s + 4.0 * (n_u + 1) + kk * 32.0 * n_u + 16.0 * n_u
        gbs = byts / (ms.value * 1e-3) / 1e9
        spmv[f"k{kk}"] = {"avg_us": ms.value * 1e3, "bytes": byts, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}
    spmv["kernel"] = ("k_amul_sell (FP64 complex A(w) X, S in sliced ELL: the reference's row order for one "
                      "column, 16 x 16-cell tile order for 16; average of 50 launches)")
    lay.close()
    out = {
        "value": value, "ms_per_step": elapsed / steps * 1e3, "steps": steps, "warmup": warmup,
        "config": {"workload": workload, "cells": list(spec.cells), "edges": int(m.n), "ports": npt, "freqs": nf,
                   "f_range_hz": list(spec.f_range),
                   "step": (f"{npt} ports at one frequency; timed step t takes frequency t mod {nf}"
                            if per_frequency else f"all {npt * nf} columns"),
                   "columns_per_step": ncol,
                   "batch_slots": a.batch if a.batch else f"autotuned: {tuned['slots']} (of 8/16/32), "
                                                          f"lanes cap {tuned['lanes_cap']}",
                   "restart": a.restart, "rtol": a.rtol,
                   "krylov_basis": "fp64" if a.basis64 else "fp32"},
        "converged": f"{conv}/{steps * ncol}", "max_rel_residual": res_max,
        "iters": {"p50": float(np.percentile(iters, 50)), "max": int(max(iters)),
                  "per_freq_max": {f: int(max(v)) for f, v in it_f.items()}},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "setup_s": setup_s,
        "autotune": tuned,
        "coarse": {"form_ms_total": hinfo["coarse_form_ms"], "frequencies": hinfo["coarse_entries"],
                   "aux_rows": hinfo["aux"][-1][0] if hinfo["aux"] else 0,
                   "aux_blocks": hinfo["aux_coarse_blocks"], "aux_max_block": hinfo["aux_coarse_max_block"],
                   "edge_rows": hinfo["edge"][-1][0] if hinfo["edge"] else 0},
        "fp64_single": fp64,
        "speedup_vs_fp64_single": value / fp64["value"] if fp64 else None,
        "z_rel_diff_mixed_vs_fp64": zdiff,
        "spmv": spmv,
        "edge0": {"launches": launches, "ms": kms, "bytes": kbytes,
                  "kernel": "k_edge_march (u staged plane by plane in LDS) for the K u launches"
                            if a.edge_lds and hinfo["march_block_rows"] else "k_edge_sell",
                  "march_block_rows": hinfo["march_block_rows"],
                  "march_outside_share": (hinfo["march_outside"] / hinfo["march_entries"]
                                          if hinfo["march_entries"] else None)},
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="manhattan_400x400x32")
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--restart", type=int, default=16)
    ap.add_argument("--max-iters", type=int, default=1500)
    ap.add_argument("--rtol", type=float, default=1e-9,
                    help="relative residual of every column, both paths (1e-9 is within the 1e-8 the metric names; "
                         "at 1e-8 the 10 GHz columns of the two paths differ by 1.7e-6 in Z, at 1e-9 by 1.6e-7)")
    ap.add_argument("--no-fp64", dest="fp64", action="store_false", help="skip the FP64 single-RHS comparison")
    ap.add_argument("--no-secondary", dest="secondary", action="store_false",
                    help="skip the 200x200x24 block")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=12, help="requests of the serving stream (0 skips it)")
    ap.add_argument("--smooth-edge", type=int, default=0, help="edge/nodal smoothing sweeps per level (0: library default)")
    ap.add_argument("--smooth-aux", type=int, default=0, help="auxiliary AMG sweeps per level (0: library default)")
    ap.add_argument("--sell-min-rows", type=int, default=0,
                    help="operators with at least this many rows use the sliced-ELL kernels (0: library default)")
    ap.add_argument("--basis64", action="store_true", help="FP64 Krylov basis on the mixed path (default FP32)")
    ap.add_argument("--edge-lds", action="store_true",
                    help="level-0 edge smoother's K u products from u staged plane by plane in LDS (k_edge_march) "
                         "instead of gathered through the caches")
    ap.add_argument("--no-graphs", dest="graphs", action="store_false",
                    help="launch the preconditioner directly instead of replaying captured HIP graphs")
    ap.add_argument("--serve-python", action="store_true",
                    help="serve the stream with the Python scheduler instead of the library's server")
    ap.add_argument("--serve-progress", action="store_true", help="print each completed serving request to stderr")
    ap.add_argument("--serving-only", action="store_true", help=argparse.SUPPRESS)
    a = ap.parse_args()
    if a.serving_only:                   # the child process of serving_leg_guarded
        print(json.dumps(serving_leg(a.serve_requests, python_server=a.serve_python, progress=a.serve_progress)),
              flush=True)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {a.gpus} is not supported "
              "(the solve path is single-device)", file=sys.stderr)
        sys.exit(2)

    import ctypes as C
    from gds2para_amd import _capi as capi
    cms = C.c_double()
    lib = capi.load()
    capi.check(lib.gp_bench_copy(1 << 30, 5, C.byref(cms)), "gp_bench_copy")
    copy_gbs = 2.0 * (1 << 30) / (cms.value * 1e-3) / 1e9

    per_freq = a.workload == "manhattan_400x400x32"
    h = measure(a.workload, a.steps, a.warmup, a, per_freq)
    traffic, traffic_src = None, None
    tp = os.path.join(ROOT, "profiles", f"edge0_traffic_{a.workload}.json")
    if os.path.exists(tp):
        with open(tp) as f:
            tj = json.load(f)
        traffic, traffic_src = tj.get("bytes_per_launch"), os.path.relpath(tp, ROOT)
    e0 = h.pop("edge0")
    launches, kms, kbytes = e0["launches"], e0["ms"], e0["bytes"]
    ach = kbytes / (kms * 1e-3) / 1e9 if kms > 0 else None
    line = {
        "metric": METRIC,
        "value": h.pop("value"),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": h.pop("ms_per_step"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex fp64 outer (residual, products, iterate) / complex fp32 preconditioner and Krylov basis storage"
                 if not a.basis64 else "complex fp64 outer / complex fp32 preconditioner",
        "data": "synthetic (seeded Manhattan layout generator)",
        "config": h.pop("config"),
    }
    h.pop("steps"), h.pop("warmup")
    line.update(h)
    line["copy_probe_GBs"] = copy_gbs
    line["roofline"] = {"bound": "hbm",
                        "kernel": "level-0 edge smoother: k_edge_sell<float, 4, 4> for the K u launches, k_edge_nok2<float, 4, 4> for the others (4 lanes per row, 4 slots per lane)",
                        "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS if ach else None,
                        "traffic": traffic, "traffic_source": traffic_src, "launches": launches,
                        "avg_launch_us": kms / launches * 1e3 if launches else None,
                        "bytes_per_launch": kbytes / launches if launches else None}
    if a.secondary and a.workload != "manhattan_200x200x24":
        sec = measure("manhattan_200x200x24", 2, 1, a, False)
        sec.pop("spmv")
        e = sec.pop("edge0")
        sec["edge0_frac"] = (e["bytes"] / (e["ms"] * 1e-3) / 1e9) / HBM_PEAK_GBS if e["ms"] > 0 else None
        line["secondary"] = sec
    if a.secondary:
        print("[bench] graphs leg", file=sys.stderr, flush=True)
        line["graphs"] = graphs_leg()
    if a.serve_requests > 0:
        print(f"[bench] serving leg: {a.serve_requests} requests", file=sys.stderr, flush=True)
        line["serving"] = serving_leg_guarded(a.serve_requests, python_server=a.serve_python)
    if not a.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

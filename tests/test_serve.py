"""Serving loop: continuous batching over gp_solve_stream.

CPU tests drive the scheduler with a mock backend (admission under a memory
budget, eviction of idle layouts, completion and latency bookkeeping).  The
GPU test streams overlapping requests over two layouts through the library and
checks every request's Z against the same request solved alone: within 1e-6
relative (the columns restart at different points in the mix, so they agree
to the solve tolerance rather than bit for bit)."""
import ctypes as C
import threading
import time

import numpy as np
import pytest

from gds2para_amd import _capi as capi
from gds2para_amd.serve import CServer, Request, Server, poisson_arrivals, run_stream


@pytest.fixture(params=["python", "library"])
def make_server(request):
    """The scheduler under test: serve.Server, or the library's gp_server
    (admission and streams in C++) driving the same Python backend."""
    if request.param == "python":
        return lambda be, budget: Server(be, budget)
    return lambda be, budget: CServer(be, budget)


class MockBackend:
    """Solves a column by sleeping `delay` s; its Z column is a fixed function
    of (layout, port, freq).  Runs `slots` columns at a time like the library."""

    def __init__(self, sizes, nports=3, slots=2, delay=0.01):
        self.sizes, self.nports, self.slots, self.delay = sizes, nports, slots, delay
        self.opened, self.closed, self.live = [], [], set()
        self.lock = threading.Lock()

    def estimate(self, key):
        return self.sizes[key]

    def open(self, key):
        with self.lock:
            self.opened.append(key)
            self.live.add(key)
        return key

    def bytes(self, key):
        return self.sizes[key]

    def num_ports(self, key):
        return self.nports

    def close(self, key):
        with self.lock:
            self.closed.append(key)
            self.live.discard(key)

    @staticmethod
    def zcol(key, port, freq, n):
        return np.array([hash((key, port, round(freq), r)) % 1000 + 1j for r in range(n)], complex)

    def stream(self, key, pull, waiting, push):
        active = []
        closed = False
        while True:
            while len(active) < self.slots and not closed:
                col, r = pull(len(active) == 0)
                if r == -1:
                    closed = True
                if r != 1:
                    break
                active.append(col)
            if not active:
                if closed:
                    return
                continue
            time.sleep(self.delay)
            p, f, tag = active.pop(0)
            push(tag, self.zcol(key, p, f, self.nports), dict(iters=1, converged=1, rel_residual=0.0))


def test_serve_completes_and_orders(make_server):
    be = MockBackend({"a": 10, "b": 10})
    srv = make_server(be, 100)
    reqs = [Request("a" if i % 2 else "b", [i % 3], [1e9, 2e9]) for i in range(8)]
    run_stream(srv, reqs, poisson_arrivals(200.0, len(reqs), seed=1))
    for r in reqs:
        assert r.remaining == 0 and r.latency >= 0 and not r.error
        for fi, f in enumerate(r.freqs):
            for p in r.ports:
                assert np.array_equal(r.Z[fi, :, p], MockBackend.zcol(r.layout, p, f, 3))
    srv.close()
    assert sorted(be.closed) == ["a", "b"] and not be.live


def test_serve_budget_evicts_idle_layouts(make_server):
    be = MockBackend({"a": 60, "b": 60, "c": 60}, delay=0.002)
    srv = make_server(be, 100)               # one layout at a time
    reqs = [Request(k, [0, 1], [1e8]) for k in ("a", "b", "c", "a")]
    run_stream(srv, reqs, [0.0, 0.05, 0.1, 0.15])
    assert all(r.remaining == 0 for r in reqs)
    assert be.opened.count("a") == 2          # evicted and reopened
    assert len(be.live) <= 1
    srv.close()


def test_serve_backlog_admitted_when_last_layout_goes_idle(make_server):
    """Budget for one layout: a slow request for a, then at once one for b.
    b can only be admitted once a is idle, and a goes idle inside its own
    stream's last push; admission must still run then (it once ran only on
    the stream thread, which could not evict its own layout, and nothing
    retried the backlog afterwards)."""
    be = MockBackend({"a": 60, "b": 60}, slots=1, delay=0.05)
    srv = make_server(be, 100)
    a = Request("a", [0], [1e9])
    b = Request("b", [1], [1e9])
    srv.submit(a)
    srv.submit(b)
    try:
        srv.wait(timeout=10)
        ev = srv.evictions
    finally:
        srv.close()
    assert a.remaining == 0 and b.remaining == 0
    assert be.opened == ["a", "b"] and be.closed[0] == "a"
    assert ev == 1


def test_serve_open_runs_off_the_lock(make_server):
    """Opening a layout (seconds on the GPU) holds neither the server lock nor
    another layout's stream: requests on a resident layout complete while a
    second layout is still being opened."""
    class SlowOpen(MockBackend):
        def open(self, key):
            if key == "slow":
                time.sleep(1.0)
            return super().open(key)

    be = SlowOpen({"fast": 1, "slow": 1}, slots=2, delay=0.005)
    srv = make_server(be, 100)
    warm = Request("fast", [0], [1e9])
    srv.submit(warm)
    srv.wait(timeout=5)
    srv.submit(Request("slow", [0], [1e9]))
    time.sleep(0.05)                                 # the slow open is under way
    t = time.monotonic()
    quick = Request("fast", [1, 2], [1e9, 2e9])
    srv.submit(quick)
    while quick.remaining and time.monotonic() - t < 5:
        time.sleep(0.005)
    dt = time.monotonic() - t
    srv.wait(timeout=10)
    srv.close()
    assert dt < 0.5, dt


def test_serve_same_alone_and_in_mix_mock(make_server):
    be = MockBackend({"a": 1, "b": 1}, slots=3)
    alone = make_server(be, 100)
    r0 = Request("a", [0, 2], [5e8, 1e9])
    run_stream(alone, [r0], [0.0])
    alone.close()
    mix = make_server(be, 100)
    reqs = [Request("b", [1], [1e9]), Request("a", [0, 2], [5e8, 1e9]), Request("a", [1], [2e9])]
    run_stream(mix, reqs, [0.0, 0.001, 0.002])
    mix.close()
    assert np.array_equal(reqs[1].Z, r0.Z)


class SlotLoopBackend(MockBackend):
    """The library's slot loop (solve_slots with CallbackIO): columns enter
    at restarts; with every slot idle the refill waits for the first column
    only and takes whatever else is already queued without waiting."""

    def __init__(self, sizes, slots=4, steps=(2, 6), block_every_slot=False):
        super().__init__(sizes, slots=slots)
        self.steps, self.block_every_slot = steps, block_every_slot

    def stream(self, key, pull, waiting, push):
        k = self.slots
        cid, need = [None] * k, [0] * k
        closed = False
        rng = np.random.default_rng(len(key))

        def refill(mask, block):
            nonlocal closed
            placed = False
            for j in range(k):
                if not mask[j]:
                    continue
                col, r = (None, -1) if closed else pull(block and (self.block_every_slot or not placed))
                closed = closed or r == -1
                cid[j] = col if r == 1 else None
                need[j] = int(rng.integers(*self.steps))
                placed = placed or r == 1

        refill([1] * k, False)
        while True:
            while all(c is None for c in cid) and not closed:
                refill([1] * k, True)
            if all(c is None for c in cid):
                return
            time.sleep(0.002)                       # one restart cycle
            done = [0] * k
            for j in range(k):
                if cid[j] is not None:
                    need[j] -= 1
                    if need[j] <= 0:
                        p, f, tag = cid[j]
                        push(tag, self.zcol(key, p, f, self.nports), dict(iters=1, converged=1, rel_residual=0.0))
                        done[j] = 1
            refill([1 if done[j] or (cid[j] is None and not closed and waiting() > 0) else 0 for j in range(k)],
                   False)


def test_cserver_request_collected_once():
    """gp_server_wait hands a request's Z to one collector: a second collect
    of the same request is an error, not a stale or empty result."""
    import ctypes as C
    be = MockBackend({"a": 10})
    srv = CServer(be, 100)
    try:
        ports = np.array([0, 1], np.int32)
        freqs = np.array([1e9], np.float64)
        rid = C.c_int64()
        capi.check(srv.lib.gp_server_submit(srv.h, srv._key("a"), ports.ctypes.data_as(C.POINTER(C.c_int32)), 2,
                                            freqs.ctypes.data_as(C.POINTER(C.c_double)), 1, C.byref(rid)),
                   "gp_server_submit")
        rep = capi.GpRequestReport()
        z = np.zeros(2 * 2 * 3)
        assert srv.lib.gp_server_wait(srv.h, rid.value, 10.0, z.ctypes.data_as(C.POINTER(C.c_double)),
                                      C.byref(rep)) == capi.GP_OK
        assert rep.converged == 2 and np.abs(z).sum() > 0
        assert srv.lib.gp_server_wait(srv.h, rid.value, 1.0, z.ctypes.data_as(C.POINTER(C.c_double)),
                                      C.byref(rep)) != capi.GP_OK
    finally:
        srv.close()


def _serve_idle_layout(be):
    """A layout goes idle, then gets a request with fewer columns than slots
    and nothing after it: the request must still finish."""
    srv = Server(be, 100)
    first = Request("a", [0, 1], [1e9])
    srv.submit(first)
    srv.wait(timeout=10)
    time.sleep(0.05)                                # every slot idle, the loop waits
    last = Request("a", [2], [5e8, 1e9])
    srv.submit(last)
    try:
        srv.wait(timeout=5)
    finally:
        with srv.lock:
            stuck = last.remaining
        for r in list(srv.res.values()):            # unblock the loop either way
            with r.cv:
                r.closed = True
                r.cv.notify_all()
    return stuck


def test_serve_idle_layout_last_request_finishes():
    assert _serve_idle_layout(SlotLoopBackend({"a": 1})) == 0


def test_serve_idle_layout_stalls_when_every_slot_blocks():
    """The defect the library had: its refill waited for every idle slot, so
    the columns already placed sat unsolved until more columns arrived."""
    with pytest.raises(TimeoutError):
        _serve_idle_layout(SlotLoopBackend({"a": 1}, block_every_slot=True))


@pytest.mark.gpu
def test_serve_gpu_same_alone_and_in_mix():
    from gds2para_amd import mesh as M
    from gds2para_amd.layout import Layout
    from gds2para_amd.serve import GpuBackend
    specs = {"m11": M.manhattan_spec((24, 20, 10), n_ports=3, nfreq=3, f_range=(1e8, 1e10), seed=11),
             "m12": M.manhattan_spec((20, 24, 10), n_ports=3, nfreq=3, f_range=(1e8, 1e10), seed=12)}
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(10):
        key = "m11" if i % 3 else "m12"
        ports = sorted(rng.choice(3, size=int(rng.integers(1, 4)), replace=False).tolist())
        freqs = sorted(rng.choice([1e8, 1e9, 3e9, 1e10], size=int(rng.integers(1, 3)), replace=False).tolist())
        reqs.append(Request(key, ports, freqs))
    print("[serve] streaming 10 requests over 2 layouts, 4 slots", flush=True)
    srv = Server(GpuBackend(specs, slots=4), budget_bytes=200 << 30)
    run_stream(srv, reqs, poisson_arrivals(20.0, len(reqs), seed=3))
    srv.close()
    lats = [r.latency for r in reqs]
    print(f"[serve] latency p50 {np.percentile(lats, 50):.3f} s p99 {np.percentile(lats, 99):.3f} s", flush=True)
    for key, spec in specs.items():
        m = M.Mesh(spec)
        lay = Layout(m)
        for r in (q for q in reqs if q.layout == key):
            assert not r.error
            for rep in r.reports.values():
                assert rep["converged"] == 1 and rep["rel_residual"] <= 1e-8
            Z, st, _ = lay.solve(r.freqs, ports=r.ports, precision=capi.GP_PREC_MIXED, rtol=1e-8, batch=4)
            Za = Z.reshape(len(r.freqs), 3, 3)     # [f, src, resp]: index resp + np*src
            for fi in range(len(r.freqs)):
                for p in r.ports:
                    a, b = r.Z[fi, :, p], Za[fi, p, :]
                    assert np.abs(a - b).max() / np.abs(b).max() < 1e-6, (key, r.freqs[fi], p)
        lay.close()


@pytest.mark.gpu
def test_serve_gpu_evicts_under_budget():
    """Three layouts under a budget that holds one at a time (each layout's
    estimate is ~94 MB, the budget 110 MB): requests alternate between the
    layouts, so idle layouts are evicted (their device memory freed while
    another layout's stream runs) and reopened, requests for a busy layout's
    rivals wait in the backlog, and every request still finishes with the Z
    of the same columns solved alone."""
    from gds2para_amd import mesh as M
    from gds2para_amd.layout import Layout
    from gds2para_amd.serve import GpuBackend
    specs = {"a": M.manhattan_spec((24, 20, 10), n_ports=3, nfreq=3, f_range=(1e8, 1e10), seed=11),
             "b": M.manhattan_spec((20, 24, 10), n_ports=3, nfreq=3, f_range=(1e8, 1e10), seed=12),
             "c": M.manhattan_spec((28, 16, 10), n_ports=3, nfreq=3, f_range=(1e8, 1e10), seed=13)}

    class Counting(GpuBackend):
        opened = []

        def open(self, key):
            self.opened.append(key)
            return super().open(key)

    be = Counting(specs, slots=4)
    keys = ["a", "b", "c", "a", "c", "b", "a", "b"]
    reqs = [Request(k, [i % 3], [1e9] if i % 2 else [1e8, 1e10]) for i, k in enumerate(keys)]
    print(f"[serve] {len(reqs)} requests over 3 layouts, budget 110 MB", flush=True)
    srv = Server(be, budget_bytes=110 << 20)
    run_stream(srv, reqs, poisson_arrivals(10.0, len(reqs), seed=7))
    srv.close()
    print(f"[serve] opened {be.opened}", flush=True)
    assert len(be.opened) > 3                       # some layout was evicted and reopened
    for key, spec in specs.items():
        lay = Layout(M.Mesh(spec))
        for r in (q for q in reqs if q.layout == key):
            assert not r.error and r.remaining == 0
            for rep in r.reports.values():
                assert rep["converged"] == 1 and rep["rel_residual"] <= 1e-8
            Z, st, _ = lay.solve(r.freqs, ports=r.ports, precision=capi.GP_PREC_MIXED, rtol=1e-8, batch=4)
            Za = Z.reshape(len(r.freqs), 3, 3)
            for fi in range(len(r.freqs)):
                for p in r.ports:
                    a, b = r.Z[fi, :, p], Za[fi, p, :]
                    assert np.abs(a - b).max() / np.abs(b).max() < 1e-6, (key, r.freqs[fi], p)
        lay.close()


@pytest.mark.gpu
def test_stream_rejects_malformed_columns():
    """gp_solve_stream hands a column whose port is out of range or whose
    frequency is not positive straight back as rejected (no Z, negative
    residual in its report) and solves the good columns around it."""
    from gds2para_amd import mesh as M
    from gds2para_amd.layout import Layout
    m = M.Mesh(M.manhattan_spec((24, 20, 10), n_ports=3, nfreq=3, f_range=(1e8, 1e10), seed=11))
    lay = Layout(m)
    lay.assemble()
    lay.prepare(capi.GP_PREC_MIXED)
    cols = [(0, 1e9, 0), (7, 1e9, 1), (1, -5.0, 2), (2, 1e9, 3)]
    queue = list(cols)
    got = {}

    def pull(user, colp, block):
        if not queue:
            return -1
        p, f, tag = queue.pop(0)
        colp.contents.port, colp.contents.freq, colp.contents.tag = p, f, tag
        return 1

    def waiting(user):
        return len(queue)

    def push(user, colp, zp, repp):
        r = repp.contents
        z = None if not zp else [zp[2 * q] + 1j * zp[2 * q + 1] for q in range(m.num_ports)]
        got[colp.contents.tag] = (z, r.converged, r.rel_residual)

    cb = (capi.PULL_FN(pull), capi.WAITING_FN(waiting), capi.PUSH_FN(push))
    o = capi.GpSolveOpts()
    lay.lib.gp_solve_opts_default(C.byref(o))
    st = capi.GpSolveStats()
    print("[serve] stream with two malformed columns", flush=True)
    capi.check(lay.lib.gp_solve_stream(lay.h, C.byref(o), 4, cb[0], cb[1], cb[2], None, C.byref(st)),
               "gp_solve_stream")
    assert sorted(got) == [0, 1, 2, 3]
    for tag in (1, 2):
        z, conv, rr = got[tag]
        assert z is None and conv == 0 and rr < 0
    Z, _, _ = lay.solve([1e9], ports=[0, 2], precision=capi.GP_PREC_MIXED, rtol=1e-8, batch=4)
    Zr = Z[0].reshape(3, 3)
    for tag, p in ((0, 0), (3, 2)):
        z, conv, rr = got[tag]
        assert conv == 1 and rr <= 1e-8
        assert np.abs(np.array(z) - Zr[p]).max() / np.abs(Zr[p]).max() < 1e-6
    lay.close()


@pytest.mark.gpu
def test_serve_gpu_stream_generator():
    """The serving workload's generator (gds2para_amd.stream) at a size a test
    can run: 6 layouts of 3e4-3e5 edges with 8 ports each, 16 requests of 1-8
    ports x 1-6 frequencies at 8 requests/s, under a budget that holds about
    two layouts.  Every column converges, layouts are evicted and reopened
    (admission from the backlog), and a sample of requests solved alone on a
    fresh layout gives the same Z to 1e-6."""
    from gds2para_amd import stream as ST
    from gds2para_amd.layout import Layout
    from gds2para_amd import mesh as M
    from gds2para_amd.serve import GpuBackend
    sp = ST.StreamSpec(n_layouts=6, edges=(3e4, 3e5), max_ports=8, max_freqs=6, n_requests=16, rate_hz=8.0,
                       seed=77)
    specs = ST.layouts(sp)
    raw, offs = ST.requests(sp)

    class Counting(GpuBackend):
        opened = []

        def open(self, key):
            self.opened.append(key)
            return super().open(key)

    be = Counting(specs, slots=8)
    est = sorted(be.estimate(k) for k in specs)
    budget = est[-1] + est[-2]
    reqs = [Request(k, p, f) for k, p, f in raw]
    print(f"[serve] stream generator: {len(reqs)} requests over {len({k for k, _, _ in raw})} layouts, "
          f"budget {budget >> 20} MB", flush=True)
    srv = Server(be, budget_bytes=budget)
    run_stream(srv, reqs, offs)
    ev = srv.evictions
    srv.close()
    lats = [r.latency for r in reqs]
    print(f"[serve] opened {len(be.opened)}, evictions {ev}, latency p50 {np.percentile(lats, 50):.3f} s "
          f"p99 {np.percentile(lats, 99):.3f} s", flush=True)
    assert ev >= 1 and len(be.opened) > len(set(be.opened))
    for r in reqs:
        assert not r.error and r.remaining == 0
        for rep in r.reports.values():
            assert rep["converged"] == 1 and rep["rel_residual"] <= 1e-8
    for r in reqs[:3]:
        lay = Layout(M.Mesh(specs[r.layout]))
        Z, st, _ = lay.solve(r.freqs, ports=r.ports, precision=capi.GP_PREC_MIXED, rtol=1e-8, batch=8)
        lay.close()
        npt = specs[r.layout].n_ports
        Za = Z.reshape(len(r.freqs), npt, npt)
        for fi in range(len(r.freqs)):
            for p in r.ports:
                a, b = r.Z[fi, :, p], Za[fi, p, :]
                assert np.abs(a - b).max() / np.abs(b).max() < 1e-6, (r.layout, r.freqs[fi], p)


def _check_against_alone(specs, reqs, nports=3):
    """Each request's Z equals the same columns solved alone by Layout.solve."""
    from gds2para_amd import mesh as M
    from gds2para_amd.layout import Layout
    for key, spec in specs.items():
        lay = Layout(M.Mesh(spec))
        for r in (q for q in reqs if q.layout == key):
            Z, st, _ = lay.solve(r.freqs, ports=r.ports, precision=capi.GP_PREC_MIXED, rtol=1e-8, batch=4)
            Za = Z.reshape(len(r.freqs), nports, nports)     # [f, src, resp]
            for fi in range(len(r.freqs)):
                for p in r.ports:
                    a, b = r.Z[fi, :, p], Za[fi, p, :]
                    assert np.abs(a - b).max() / np.abs(b).max() < 1e-6, (key, r.freqs[fi], p)
        lay.close()


@pytest.mark.gpu
def test_cserver_gpu_same_alone_and_in_mix():
    """The library's server with its own backend (gp_server_add_mesh layouts,
    gp_solve_stream on the server's stream threads): a Poisson stream of
    requests over two layouts returns, for every request, the Z of the same
    columns solved alone."""
    from gds2para_amd import mesh as M
    specs = {"m11": M.manhattan_spec((24, 20, 10), n_ports=3, nfreq=3, f_range=(1e8, 1e10), seed=11),
             "m12": M.manhattan_spec((20, 24, 10), n_ports=3, nfreq=3, f_range=(1e8, 1e10), seed=12)}
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(10):
        key = "m11" if i % 3 else "m12"
        ports = sorted(rng.choice(3, size=int(rng.integers(1, 4)), replace=False).tolist())
        freqs = sorted(rng.choice([1e8, 1e9, 3e9, 1e10], size=int(rng.integers(1, 3)), replace=False).tolist())
        reqs.append(Request(key, ports, freqs))
    print("[cserver] streaming 10 requests over 2 layouts, 4 slots", flush=True)
    srv = CServer(None, 200 << 30, slots=4, rtol=1e-8, layouts=specs)
    run_stream(srv, reqs, poisson_arrivals(20.0, len(reqs), seed=3))
    ev = srv.evictions
    srv.close()
    lats = [r.latency for r in reqs]
    print(f"[cserver] latency p50 {np.percentile(lats, 50):.3f} s p99 {np.percentile(lats, 99):.3f} s, "
          f"evictions {ev}", flush=True)
    assert ev == 0
    for r in reqs:
        assert not r.error and r.remaining == 0, r.error
        assert r.reports["converged"] == len(r.ports) * len(r.freqs)
        assert r.reports["max_rel_residual"] <= 1e-8
    _check_against_alone(specs, reqs)


@pytest.mark.gpu
def test_cserver_gpu_evicts_under_budget():
    """The library's server under a budget that holds one layout at a time
    (each estimate ~94 MB, budget 110 MB): requests alternate between three
    layouts, so idle layouts are evicted and reopened, and every request still
    returns the Z of its columns solved alone."""
    from gds2para_amd import mesh as M
    specs = {"a": M.manhattan_spec((24, 20, 10), n_ports=3, nfreq=3, f_range=(1e8, 1e10), seed=11),
             "b": M.manhattan_spec((20, 24, 10), n_ports=3, nfreq=3, f_range=(1e8, 1e10), seed=12),
             "c": M.manhattan_spec((28, 16, 10), n_ports=3, nfreq=3, f_range=(1e8, 1e10), seed=13)}
    keys = ["a", "b", "c", "a", "c", "b", "a", "b"]
    reqs = [Request(k, [i % 3], [1e9] if i % 2 else [1e8, 1e10]) for i, k in enumerate(keys)]
    print(f"[cserver] {len(reqs)} requests over 3 layouts, budget 110 MB", flush=True)
    srv = CServer(None, 110 << 20, slots=4, rtol=1e-8, layouts=specs)
    run_stream(srv, reqs, poisson_arrivals(10.0, len(reqs), seed=7))
    ev = srv.evictions
    srv.close()
    print(f"[cserver] evictions {ev}", flush=True)
    assert ev > 0
    for r in reqs:
        assert not r.error and r.remaining == 0, r.error
        assert r.reports["converged"] == len(r.ports) * len(r.freqs)
    _check_against_alone(specs, reqs)


@pytest.mark.gpu
def test_cserver_gpu_rejects_bad_port_and_keeps_serving():
    """A column with a port the layout does not have comes back rejected
    (counted in the request's report, its Z column left zero); the request's
    valid columns are solved, and the server serves the next request."""
    from gds2para_amd import mesh as M
    spec = M.manhattan_spec((24, 20, 10), n_ports=3, nfreq=3, f_range=(1e8, 1e10), seed=11)
    print("[cserver] request with port 5 on a 3-port layout", flush=True)
    srv = CServer(None, 200 << 30, slots=4, rtol=1e-8, layouts={"m": spec})
    try:
        bad = Request("m", [0, 5], [1e9])
        srv.submit(bad)
        srv.wait(120)
        assert bad.reports["rejected"] == 1 and bad.reports["converged"] == 1
        assert bad.error == "rejected column"
        good = Request("m", [1], [1e9])
        srv.submit(good)
        srv.wait(120)
        assert not good.error and good.reports["converged"] == 1
    finally:
        srv.close()
    _check_against_alone({"m": spec}, [good])
    # the rejected request's valid column equals the same column solved alone
    from gds2para_amd.layout import Layout
    lay = Layout(M.Mesh(spec))
    Z, st, _ = lay.solve([1e9], ports=[0], precision=capi.GP_PREC_MIXED, rtol=1e-8, batch=4)
    lay.close()
    Za = Z.reshape(1, 3, 3)
    assert np.abs(bad.Z[0, :, 0] - Za[0, 0, :]).max() / np.abs(Za[0, 0, :]).max() < 1e-6

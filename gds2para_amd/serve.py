"""Request-serving loop with continuous batching.

A request asks for the Z columns of some ports of one layout at some
frequencies.  The server keeps layouts resident on the GPU (stiffness, V0
basis, multigrid hierarchy, per-frequency coarse inverses) while their memory
fits a budget, and feeds every resident layout's columns to one long-running
gp_solve_stream call on that layout's own HIP stream: a column that arrives
while others are on the GPU takes the next slot that frees up at an FGMRES
restart, and layouts run side by side on the device.

Admission: a request for a layout that is not resident waits until the
layout's bytes (hierarchy plus slot workspace) fit the budget, evicting idle
resident layouts least recently used first.

The reference has no serving path (it solves one layout, one frequency at a
time); this module is the host side of the gp_solve_stream entry point.
"""
from __future__ import annotations

import collections
import ctypes as C
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _capi as capi


@dataclass
class Request:
    """Z columns of `ports` at `freqs` (Hz) for layout `layout`."""
    layout: str
    ports: list
    freqs: list
    rid: int = -1
    arrival: float = 0.0
    done: float = 0.0
    Z: np.ndarray = None           # [nf, np_layout, np_layout] complex, Z[f, resp, src]
    reports: dict = field(default_factory=dict)
    remaining: int = 0
    error: str = ""

    @property
    def latency(self):
        return self.done - self.arrival


class _Resident:
    """One layout on the device, its column queue and its solve thread."""

    def __init__(self, key, handle, nbytes):
        self.key, self.handle, self.nbytes = key, handle, nbytes
        self.q = collections.deque()       # (port, freq, tag)
        self.inflight = 0
        self.closed = False
        self.cv = threading.Condition()
        self.last_used = time.monotonic()
        self.thread = None
        self.error = None


class Server:
    """Continuous-batching server over a backend (GpuBackend for the library).

    backend.estimate(key) -> bytes before opening; backend.open(key) ->
    handle; backend.bytes(handle) -> device bytes incl. the slot workspace;
    backend.stream(handle, pull, waiting, push) runs until pull returns -1;
    backend.close(handle).

    Admission runs on a thread of its own.  It is woken when a request
    arrives, when a layout goes idle and when a layout has been opened, and it
    is the only thread that opens, evicts or queues columns.  Opening a layout
    (assembly and multigrid setup, seconds at millions of edges) therefore
    holds neither the server lock nor another layout's stream thread."""

    def __init__(self, backend, budget_bytes, clock=time.monotonic, on_done=None):
        self.backend = backend
        self.budget = budget_bytes
        self.clock = clock
        self.on_done = on_done             # called with each request once it completes
        self.res = {}                      # key -> _Resident
        self.backlog = collections.deque() # requests not yet queued on a layout
        self.requests = {}
        self.tags = {}                     # tag -> (rid, freq index, port)
        self.lock = threading.RLock()
        self.all_done = threading.Condition(self.lock)
        self._kick = threading.Condition(self.lock)
        self._dirty = False
        self._stop = False
        self._opening = 0                  # bytes reserved for a layout being opened
        self.evictions = 0
        self.error = None
        self._next_rid = 0
        self._next_tag = 0
        self._admitter = threading.Thread(target=self._admission_loop, daemon=True)
        self._admitter.start()

    # ---- public ----------------------------------------------------------
    def submit(self, req: Request) -> int:
        with self.lock:
            req.rid = self._next_rid
            self._next_rid += 1
            req.arrival = self.clock()
            req.remaining = len(req.ports) * len(req.freqs)
            self.requests[req.rid] = req
            r = self.res.get(req.layout)
            if r is not None and not any(q.layout == req.layout for q in self.backlog):
                self._enqueue(r, req)       # resident: no wait for the admission thread
            else:
                self.backlog.append(req)
                self._wake()
            return req.rid

    def wait(self, timeout=None):
        """Blocks until every submitted request has finished."""
        end = None if timeout is None else self.clock() + timeout
        with self.lock:
            while any(r.remaining > 0 for r in self.requests.values()):
                left = None if end is None else end - self.clock()
                if left is not None and left <= 0:
                    raise TimeoutError("requests still running")
                self.all_done.wait(0.05 if left is None else min(0.05, left))
                if self.error:
                    raise RuntimeError(f"admission: {self.error}")
                for r in self.res.values():
                    if r.error:
                        raise RuntimeError(f"layout {r.key}: {r.error}")

    def close(self):
        with self.lock:
            self._stop = True
            self._kick.notify_all()
        self._admitter.join()
        with self.lock:
            keys = list(self.res)
        for k in keys:
            self._evict(k)

    def resident_bytes(self):
        with self.lock:
            return sum(r.nbytes for r in self.res.values()) + self._opening

    # ---- admission ---------------------------------------------------------
    def _wake(self):
        """Asks the admission thread for a pass (caller holds the lock)."""
        self._dirty = True
        self._kick.notify_all()

    def _admission_loop(self):
        try:
            while True:
                with self.lock:
                    while not self._dirty and not self._stop:
                        self._kick.wait()
                    if self._stop:
                        return
                    self._dirty = False
                    plan = self._plan()
                if plan is None:
                    continue
                key, victims, need = plan
                for v in victims:           # already out of self.res: idle, nothing refills them
                    self._shutdown(v)
                try:
                    handle = self.backend.open(key)
                    nbytes = self.backend.bytes(handle)
                except Exception as e:     # surfaced by wait()
                    with self.lock:
                        self._opening -= need
                        self.error = f"opening {key}: {e}"
                        self.all_done.notify_all()
                    return
                r = _Resident(key, handle, nbytes)
                r.thread = threading.Thread(target=self._run, args=(r,), daemon=True)
                with self.lock:
                    self._opening -= need
                    self.res[key] = r
                    self._wake()            # queue the columns that waited for it
                r.thread.start()
        except Exception as e:
            with self.lock:
                self.error = str(e)
                self.all_done.notify_all()

    def _plan(self):
        """Caller holds the lock.  Queues every backlog request whose layout
        is resident, then returns (key, victims, bytes) for the oldest request
        whose layout is not: victims are idle layouts (least recently used
        first), already taken out of self.res, whose eviction makes it fit;
        None if nothing is to be opened now (the request waits for a layout to
        go idle).  A layout larger than the whole budget is opened only when
        nothing else is resident."""
        waiting = collections.deque()
        plan, blocked = None, False
        for req in self.backlog:
            r = self.res.get(req.layout)
            if r is not None:
                self._enqueue(r, req)
                continue
            waiting.append(req)
            # layouts open one at a time, oldest request first (a large layout
            # is not starved by smaller ones that keep fitting)
            if plan is not None or blocked or self._opening:
                continue
            need = self.backend.estimate(req.layout)
            used = sum(x.nbytes for x in self.res.values())
            idle = sorted((x for x in self.res.values() if not x.q and x.inflight == 0), key=lambda x: x.last_used)
            victims = []
            while used + need > self.budget and idle:
                v = idle.pop(0)
                victims.append(v)
                used -= v.nbytes
            if used + need > self.budget and len(victims) < len(self.res):
                blocked = True              # retried when a layout goes idle
                continue
            for v in victims:
                del self.res[v.key]
                self.evictions += 1
            self._opening = need
            plan = (req.layout, victims, need)
        self.backlog = waiting
        return plan

    def _enqueue(self, r, req):
        nps = self.backend.num_ports(r.handle)
        req.Z = np.zeros((len(req.freqs), nps, nps), complex)
        with r.cv:
            for fi, f in enumerate(req.freqs):
                for p in req.ports:
                    tag = self._next_tag
                    self._next_tag += 1
                    self.tags[tag] = (req.rid, fi, p)
                    r.q.append((p, float(f), tag))
            r.last_used = self.clock()
            r.cv.notify_all()

    def _shutdown(self, r):
        with r.cv:
            r.closed = True
            r.cv.notify_all()
        r.thread.join()
        self.backend.close(r.handle)

    def _evict(self, key):
        with self.lock:
            r = self.res.pop(key, None)
        if r is not None:
            self._shutdown(r)

    # ---- the stream of one layout -------------------------------------------
    def _run(self, r: _Resident):
        def pull(block):
            with r.cv:
                while not r.q:
                    if r.closed:
                        return None, -1
                    if not block:
                        return None, 0
                    r.cv.wait(0.05)
                col = r.q.popleft()
                r.inflight += 1
                return col, 1

        def waiting():
            return len(r.q)

        def push(tag, z, rep):
            finished = None
            with self.lock:
                rid, fi, p = self.tags.pop(tag)
                req = self.requests[rid]
                if z is None:
                    req.error = "rejected column"
                else:
                    req.Z[fi, :, p] = z
                req.reports[(fi, p)] = rep
                req.remaining -= 1
                if req.remaining == 0:
                    req.done = self.clock()
                    finished = req
                with r.cv:
                    r.inflight -= 1
                    r.last_used = self.clock()
                    idle = not r.q and r.inflight == 0
                if idle and self.backlog:
                    self._wake()            # a waiting request may now evict this layout
                self.all_done.notify_all()
            if finished is not None and self.on_done:
                self.on_done(finished)

        try:
            self.backend.stream(r.handle, pull, waiting, push)
        except Exception as e:  # surfaced by wait()
            r.error = str(e)
            with self.lock:
                self.all_done.notify_all()


class GpuBackend:
    """Backend over libgds2para: one Layout per key, solved with gp_solve_stream."""

    def __init__(self, specs, slots=16, precision=capi.GP_PREC_MIXED, rtol=1e-8, restart=16, max_iters=2000):
        from . import mesh as M
        self.M = M
        self.specs = specs                 # key -> GenSpec
        self.slots, self.precision, self.rtol = slots, precision, rtol
        self.restart, self.max_iters = restart, max_iters
        self.lib = capi.load()

    def estimate(self, key):
        """Bytes before the layout exists: ~2.3 KB per edge of hierarchy plus
        the slot workspace (basis counted at FP64, an upper bound; FP32
        preconditioned basis)."""
        cx, cy, cz = self.specs[key].cells
        n = 3 * (cx + 1) * (cy + 1) * (cz + 1)
        ws = self.slots * n * ((self.restart + 1) * 16 + self.restart * 8 * 2 + 256)
        return int(n * 2300 + ws)

    def open(self, key):
        from .layout import Layout
        m = self.M.Mesh(self.specs[key])
        lay = Layout(m)
        lay.assemble()
        lay.prepare(self.precision)
        return lay

    def bytes(self, lay):
        ws = C.c_int64()
        capi.check(self.lib.gp_solve_workspace_bytes(lay.h, self.precision, self.slots, self.restart, C.byref(ws)),
                   "gp_solve_workspace_bytes")
        return lay.device_bytes + ws.value

    def num_ports(self, lay):
        return lay.mesh.num_ports

    def close(self, lay):
        lay.close()

    def stream(self, lay, pull, waiting, push):
        npt = lay.mesh.num_ports

        def c_pull(user, colp, block):
            col, r = pull(bool(block))
            if r == 1:
                colp.contents.port, colp.contents.freq, colp.contents.tag = col
            return r

        def c_waiting(user):
            return waiting()

        def c_push(user, colp, zp, repp):
            z = None
            if zp:
                a = np.ctypeslib.as_array(zp, shape=(2 * npt,))
                z = a[0::2] + 1j * a[1::2]
            rp = repp.contents
            push(colp.contents.tag, z, dict(iters=rp.iters, converged=rp.converged, rel_residual=rp.rel_residual))

        cb = (capi.PULL_FN(c_pull), capi.WAITING_FN(c_waiting), capi.PUSH_FN(c_push))
        o = capi.GpSolveOpts()
        self.lib.gp_solve_opts_default(C.byref(o))
        o.rtol, o.precision, o.restart, o.max_iters = self.rtol, self.precision, self.restart, self.max_iters
        st = capi.GpSolveStats()
        capi.check(self.lib.gp_solve_stream(lay.h, C.byref(o), self.slots, cb[0], cb[1], cb[2], None, C.byref(st)),
                   "gp_solve_stream")


class CServer:
    """The library's server (gp_server_*: admission, eviction and the layouts'
    stream threads in C++), with Server's interface: submit(Request),
    wait(timeout), close(), evictions.

    backend=None serves the library's own layouts: `layouts` maps each key
    to a GenSpec, generated here and handed to gp_server_add_mesh.
    Otherwise `backend` follows Server's protocol (estimate, open, bytes,
    num_ports, close, stream) and is called from the server's threads.  Each
    submitted request is collected by a thread of its own blocked in
    gp_server_wait, which fills Z, the reports and `remaining` when it ends."""

    def __init__(self, backend, budget_bytes, slots=16, precision=capi.GP_PREC_MIXED, rtol=1e-8, restart=16,
                 max_iters=2000, layouts=None, clock=time.monotonic, on_done=None):
        self.lib = capi.load()
        self.clock = clock
        self.on_done = on_done              # called with each request once it is collected
        self.backend = backend
        self.keys, self.names = {}, []      # request key <-> server key
        self.handles = {}                   # handle id -> backend handle
        self._next_handle = 1
        self.meshes = []
        self._collectors = []
        self.error = None
        opts = capi.GpSolveOpts()
        self.lib.gp_solve_opts_default(C.byref(opts))
        opts.rtol, opts.precision, opts.restart, opts.max_iters = rtol, precision, restart, max_iters
        h = C.c_void_p()
        if backend is None:
            capi.check(self.lib.gp_server_create(None, C.byref(opts), slots, int(budget_bytes), C.byref(h)),
                       "gp_server_create")
            self.h = h
            from . import mesh as M
            for name, spec in (layouts or {}).items():
                m = M.Mesh(spec)
                k = C.c_int32()
                capi.check(self.lib.gp_server_add_mesh(self.h, m.ptr, C.byref(k)), "gp_server_add_mesh")
                self.meshes.append(m)
                self.keys[name] = k.value
                self.names.append(name)
        else:
            self._cb = self._callbacks()
            be = capi.GpServerBackend(None, *self._cb)
            self._be = be
            capi.check(self.lib.gp_server_create(C.byref(be), C.byref(opts), slots, int(budget_bytes), C.byref(h)),
                       "gp_server_create")
            self.h = h

    # ---- a Python backend behind the C callbacks ----------------------------
    def _key(self, name):
        if name not in self.keys:
            if self.backend is None:
                raise KeyError(f"unknown layout {name}")
            self.keys[name] = len(self.names)
            self.names.append(name)
        return self.keys[name]

    def _callbacks(self):
        be = self.backend

        def estimate(user, key):
            return int(be.estimate(self.names[key]))

        def open_(user, key, hp):
            try:
                obj = be.open(self.names[key])
            except Exception as e:           # reported by wait() as the open's failure
                self.error = f"opening {self.names[key]}: {e}"
                return capi.GP_ERR_STATE
            hid = self._next_handle
            self._next_handle += 1
            self.handles[hid] = obj
            hp[0] = hid
            return capi.GP_OK

        def nbytes(user, hid):
            return int(be.bytes(self.handles[hid]))

        def num_ports(user, key, hid):
            return int(be.num_ports(self.handles[hid]))

        def close(user, hid):
            be.close(self.handles.pop(hid))

        def stream(user, hid, pull_fp, waiting_fp, push_fp, ctx):
            col = capi.GpStreamColumn()

            def pull(block):
                r = pull_fp(ctx, C.byref(col), 1 if block else 0)
                return ((col.port, col.freq, col.tag) if r == 1 else None), r

            def waiting():
                return waiting_fp(ctx)

            def push(tag, z, rep):
                c = capi.GpStreamColumn(0, 0.0, tag)
                r = capi.GpColumnReport(int(rep.get("iters", 0)), int(rep.get("converged", 0)),
                                        float(rep.get("rel_residual", -1.0)), 0.0, 0.0)
                if z is None:
                    push_fp(ctx, C.byref(c), None, C.byref(r))
                else:
                    a = np.ascontiguousarray(np.stack([np.real(z), np.imag(z)], -1).ravel())
                    push_fp(ctx, C.byref(c), a.ctypes.data_as(C.POINTER(C.c_double)), C.byref(r))

            try:
                be.stream(self.handles[hid], pull, waiting, push)
            except Exception as e:
                self.error = f"stream: {e}"
                return capi.GP_ERR_STATE
            return capi.GP_OK

        return (capi.SB_ESTIMATE(estimate), capi.SB_OPEN(open_), capi.SB_BYTES(nbytes),
                capi.SB_NUM_PORTS(num_ports), capi.SB_CLOSE(close), capi.SB_STREAM(stream))

    # ---- public ------------------------------------------------------------
    def submit(self, req: Request) -> int:
        key = self._key(req.layout)
        ports = np.ascontiguousarray(np.asarray(req.ports, np.int32))
        freqs = np.ascontiguousarray(np.asarray(req.freqs, np.float64))
        rid = C.c_int64()
        req.arrival = self.clock()
        req.remaining = len(ports) * len(freqs)
        capi.check(self.lib.gp_server_submit(self.h, key, ports.ctypes.data_as(C.POINTER(C.c_int32)), len(ports),
                                             freqs.ctypes.data_as(C.POINTER(C.c_double)), len(freqs),
                                             C.byref(rid)), "gp_server_submit")
        req.rid = rid.value
        t = threading.Thread(target=self._collect, args=(req,), daemon=True)
        self._collectors.append(t)
        t.start()
        return req.rid

    def _collect(self, req):
        rep = capi.GpRequestReport()
        st = self.lib.gp_server_wait(self.h, req.rid, -1.0, None, C.byref(rep))     # done; npt sizes Z
        if st == capi.GP_OK:
            npt = max(rep.npt, 1)
            buf = np.zeros(2 * len(req.freqs) * len(req.ports) * npt)
            st = self.lib.gp_server_wait(self.h, req.rid, -1.0, buf.ctypes.data_as(C.POINTER(C.c_double)),
                                         C.byref(rep))
        if st != capi.GP_OK:
            req.error = self.error or self.lib.gp_last_error().decode(errors="replace")
            self.error = self.error or req.error
            return
        npt = rep.npt
        z = (buf[0::2] + 1j * buf[1::2]).reshape(len(req.freqs), len(req.ports), max(npt, 1))
        req.Z = np.zeros((len(req.freqs), npt, npt), complex)
        for pi, p in enumerate(req.ports):
            if 0 <= p < npt:
                req.Z[:, :, p] = z[:, pi, :]
        req.reports = dict(converged=rep.converged, rejected=rep.rejected, max_iters=rep.max_iters,
                           max_rel_residual=rep.max_rel_residual)
        if rep.rejected:
            req.error = "rejected column"
        req.done = req.arrival + (rep.done_s - rep.arrival_s)
        req.remaining = 0
        if self.on_done:
            self.on_done(req)

    def wait(self, timeout=None):
        """Blocks until every submitted request has been collected."""
        end = None if timeout is None else self.clock() + timeout
        for t in list(self._collectors):
            left = None if end is None else max(0.0, end - self.clock())
            t.join(left)
            if t.is_alive():
                raise TimeoutError("requests still running")
        if self.error:
            raise RuntimeError(self.error)

    @property
    def evictions(self):
        info = capi.GpServerInfo()
        capi.check(self.lib.gp_server_info_get(self.h, C.byref(info)), "gp_server_info_get")
        return info.evictions

    def close(self):
        if getattr(self, "h", None):
            self.lib.gp_server_destroy(self.h)
            self.h = None


def poisson_arrivals(rate_hz, n, seed=0):
    """Arrival offsets (s) of a seeded Poisson stream."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_hz, n))


def run_stream(server, requests, offsets, clock=time.monotonic, sleep=time.sleep):
    """Submits requests at their offsets from now, waits, returns them with latencies."""
    t0 = clock()
    for req, off in zip(requests, offsets):
        dt = t0 + off - clock()
        if dt > 0:
            sleep(dt)
        server.submit(req)
    server.wait()
    return requests

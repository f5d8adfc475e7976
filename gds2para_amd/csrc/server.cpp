// server.cpp -- request serving with continuous batching over resident
// layouts, in the library: gp_server_*.
//
// A request asks for the Z columns of some ports of one layout at some
// frequencies.  Layouts stay resident while their bytes fit a budget; every
// resident layout has one thread that runs a long gp_solve_stream call on the
// layout's own HIP stream and takes columns from the layout's queue, so a
// column that arrives while others are on the GPU takes the next slot that
// frees up at an FGMRES restart, and layouts run side by side on the device.
//
// Admission runs on a thread of its own.  It is woken when a request arrives
// for a layout that is not resident (or one with older requests waiting),
// when a layout goes idle and when a layout has been opened, and it is the
// only thread that opens or evicts layouts.  It queues backlog requests whose
// layout is resident, then takes the oldest request whose layout is not: its
// layout is opened once idle layouts (least recently used first) are evicted
// to make it fit; a layout larger than the whole budget is opened only when
// nothing else is resident.  Layouts open one at a time, oldest request
// first, so a large layout is not starved by smaller ones that keep fitting.
// Opening (assembly and multigrid setup, seconds at millions of edges) holds
// neither the server lock nor any layout's stream thread.
//
// Lock order: the server's mutex, then a layout's mutex.  The reference has no
// serving path; there is nothing here to cite.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gds2para.h"
#include "gp_status.hpp"

namespace {

typedef std::chrono::steady_clock Clock;

struct Column {
  int32_t port;
  double freq;
  int64_t tag;
};

struct Resident {
  int32_t key = 0;
  void *handle = nullptr;
  int64_t nbytes = 0;
  int32_t npt = 0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Column> q;
  int64_t inflight = 0;
  bool closed = false;
  double last_used = 0;
  std::thread th;
  std::string error;
};

struct Request {
  int32_t key = 0;
  std::vector<int32_t> ports;
  std::vector<double> freqs;
  int32_t npt = 0;
  std::vector<double> Z;          // (fi * np + pi) * npt + resp, complex interleaved
  gp_request_report rep{};
  int64_t remaining = 0;
};

struct TagRef {
  int64_t rid;
  int32_t fi, pi;
};

struct Plan {
  bool open = false;
  int32_t key = 0;
  int64_t need = 0;
  std::vector<std::unique_ptr<Resident>> victims;
};

// The library's own backend: key k is the k-th mesh given to
// gp_server_add_mesh; a layout is created, assembled and prepared on open and
// solved with gp_solve_stream.
struct LibBackend {
  std::vector<const gp_mesh *> meshes;
  gp_solve_opts opts{};
  int32_t slots = 16;
};

int64_t lib_estimate(void *user, int32_t key) {
  const LibBackend *b = static_cast<const LibBackend *>(user);
  if (key < 0 || key >= (int32_t)b->meshes.size()) return 0;
  // ~2.3 KB per edge of hierarchy and operators, plus the slot workspace with
  // the basis counted at FP64 (an upper bound) and the FP32 preconditioned blocks
  const gp_mesh *m = b->meshes[key];
  const int64_t n = 3 * m->nx * m->ny * m->nz;
  const int64_t r = std::max(b->opts.restart, 2);
  const int64_t ws = (int64_t)b->slots * n * ((r + 1) * 16 + r * 8 * 2 + 256);
  return n * 2300 + ws;
}

int32_t lib_open(void *user, int32_t key, void **handle) {
  const LibBackend *b = static_cast<const LibBackend *>(user);
  if (key < 0 || key >= (int32_t)b->meshes.size()) return gp::fail(GP_ERR_ARG, "unknown layout key");
  gp_layout *lay = nullptr;
  gp_status st = gp_layout_create(b->meshes[key], &lay);
  if (st == GP_OK) st = gp_assemble_stiffness(lay);
  if (st == GP_OK) st = gp_prepare_solver(lay, b->opts.precision);
  if (st != GP_OK) {
    if (lay) gp_layout_destroy(lay);
    return st;
  }
  *handle = lay;
  return GP_OK;
}

int64_t lib_bytes(void *user, void *handle) {
  const LibBackend *b = static_cast<const LibBackend *>(user);
  gp_layout *lay = static_cast<gp_layout *>(handle);
  int64_t ws = 0;
  if (gp_solve_workspace_bytes(lay, b->opts.precision, b->slots, b->opts.restart, &ws) != GP_OK) ws = 0;
  return gp_layout_device_bytes(lay) + ws;
}

int32_t lib_num_ports(void *user, int32_t key, void *) {
  const LibBackend *b = static_cast<const LibBackend *>(user);
  return key >= 0 && key < (int32_t)b->meshes.size() ? b->meshes[key]->num_ports : 0;
}

void lib_close(void *, void *handle) { gp_layout_destroy(static_cast<gp_layout *>(handle)); }

int32_t lib_stream(void *user, void *handle, gp_pull_fn pull, gp_waiting_fn waiting, gp_push_fn push, void *ctx) {
  const LibBackend *b = static_cast<const LibBackend *>(user);
  gp_solve_stats st{};
  return gp_solve_stream(static_cast<gp_layout *>(handle), &b->opts, b->slots, pull, waiting, push, ctx, &st);
}

}  // namespace

struct gp_server {
  gp_server_backend be{};
  std::unique_ptr<LibBackend> lib;  // set when the library's own backend is used
  int64_t budget = 0;
  Clock::time_point t0 = Clock::now();

  std::mutex mu;
  std::condition_variable kick, done;
  std::map<int32_t, std::unique_ptr<Resident>> res;
  std::deque<int64_t> backlog;                 // request ids not yet queued on a layout
  std::unordered_map<int64_t, Request> reqs;
  std::unordered_map<int64_t, TagRef> tags;
  bool dirty = false, stop = false;
  int64_t opening = 0;                         // bytes reserved for the layout being opened
  int64_t evictions = 0;
  std::string error;
  int64_t next_rid = 0, next_tag = 0;
  std::thread admitter;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
  void wake() { dirty = true; kick.notify_all(); }   // caller holds mu

  void enqueue(Resident &r, int64_t rid) {       // caller holds mu
    Request &q = reqs[rid];
    q.npt = r.npt;
    q.rep.npt = r.npt;
    q.Z.assign(2 * q.freqs.size() * q.ports.size() * (size_t)std::max(r.npt, 0), 0.0);
    std::lock_guard<std::mutex> g(r.mu);
    for (size_t fi = 0; fi < q.freqs.size(); ++fi)
      for (size_t pi = 0; pi < q.ports.size(); ++pi) {
        const int64_t tag = next_tag++;
        tags[tag] = {rid, (int32_t)fi, (int32_t)pi};
        r.q.push_back({q.ports[pi], q.freqs[fi], tag});
      }
    r.last_used = now();
    r.cv.notify_all();
  }

  Plan plan() {                                   // caller holds mu
    Plan p;
    std::deque<int64_t> waiting;
    bool blocked = false;
    for (int64_t rid : backlog) {
      const int32_t key = reqs[rid].key;
      auto it = res.find(key);
      if (it != res.end()) {
        enqueue(*it->second, rid);
        continue;
      }
      waiting.push_back(rid);
      if (p.open || blocked || opening) continue;
      const int64_t need = be.estimate(be.user, key);
      int64_t used = 0;
      std::vector<Resident *> idle;
      for (auto &kv : res) {
        used += kv.second->nbytes;
        std::lock_guard<std::mutex> g(kv.second->mu);
        if (kv.second->q.empty() && kv.second->inflight == 0) idle.push_back(kv.second.get());
      }
      std::sort(idle.begin(), idle.end(), [](const Resident *a, const Resident *b) { return a->last_used < b->last_used; });
      std::vector<int32_t> vk;
      for (Resident *v : idle) {
        if (used + need <= budget) break;
        vk.push_back(v->key);
        used -= v->nbytes;
      }
      if (used + need > budget && vk.size() < res.size()) {
        blocked = true;                             // retried when a layout goes idle
        continue;
      }
      for (int32_t k : vk) {
        p.victims.push_back(std::move(res[k]));
        res.erase(k);
        ++evictions;
      }
      opening = need;
      p.open = true;
      p.key = key;
      p.need = need;
    }
    backlog.swap(waiting);
    return p;
  }

  void shutdown(Resident &r) {                    // r is out of res; no lock held
    {
      std::lock_guard<std::mutex> g(r.mu);
      r.closed = true;
      r.cv.notify_all();
    }
    if (r.th.joinable()) r.th.join();
    be.close(be.user, r.handle);
  }

  void admission_loop() {
    for (;;) {
      Plan p;
      {
        std::unique_lock<std::mutex> g(mu);
        kick.wait(g, [&] { return dirty || stop; });
        if (stop) return;
        dirty = false;
        p = plan();
      }
      if (!p.open) continue;
      for (auto &v : p.victims) shutdown(*v);
      p.victims.clear();
      void *h = nullptr;
      const int32_t st = be.open(be.user, p.key, &h);
      if (st != GP_OK) {
        std::lock_guard<std::mutex> g(mu);
        opening -= p.need;
        error = "opening layout " + std::to_string(p.key) + ": " + gp_last_error();
        done.notify_all();
        return;
      }
      auto r = std::make_unique<Resident>();
      r->key = p.key;
      r->handle = h;
      r->nbytes = be.bytes(be.user, h);
      r->npt = be.num_ports(be.user, p.key, h);
      r->last_used = now();
      Resident *rp = r.get();
      {
        std::lock_guard<std::mutex> g(mu);
        opening -= p.need;
        res[p.key] = std::move(r);
        wake();                                    // queue the columns that waited for it
        rp->th = std::thread([this, rp] { run(*rp); });
      }
    }
  }

  // ---- the stream of one layout ----------------------------------------
  struct StreamCtx {
    gp_server *srv;
    Resident *r;
  };

  static int32_t c_pull(void *user, gp_stream_column *col, int32_t block) {
    Resident &r = *static_cast<StreamCtx *>(user)->r;
    std::unique_lock<std::mutex> g(r.mu);
    while (r.q.empty()) {
      if (r.closed) return -1;
      if (!block) return 0;
      // short waits on the system clock (pthread_cond_timedwait), which
      // ThreadSanitizer follows; the steady-clock form is invisible to it
      r.cv.wait_until(g, std::chrono::system_clock::now() + std::chrono::milliseconds(50));
    }
    const Column c = r.q.front();
    r.q.pop_front();
    ++r.inflight;
    col->port = c.port;
    col->freq = c.freq;
    col->tag = c.tag;
    return 1;
  }

  static int64_t c_waiting(void *user) {
    Resident &r = *static_cast<StreamCtx *>(user)->r;
    std::lock_guard<std::mutex> g(r.mu);
    return (int64_t)r.q.size();
  }

  static void c_push(void *user, const gp_stream_column *col, const double *z, const gp_column_report *rep) {
    StreamCtx &cx = *static_cast<StreamCtx *>(user);
    gp_server &s = *cx.srv;
    Resident &r = *cx.r;
    std::lock_guard<std::mutex> g(s.mu);
    auto it = s.tags.find(col->tag);
    if (it != s.tags.end()) {
      const TagRef t = it->second;
      s.tags.erase(it);
      auto qi = s.reqs.find(t.rid);
      if (qi != s.reqs.end()) {
        Request &q = qi->second;
        if (!z) {
          ++q.rep.rejected;
        } else {
          const size_t o = 2 * ((size_t)t.fi * q.ports.size() + (size_t)t.pi) * (size_t)q.npt;
          std::memcpy(&q.Z[o], z, sizeof(double) * 2 * (size_t)q.npt);
          if (rep->converged) ++q.rep.converged;
          q.rep.max_iters = std::max(q.rep.max_iters, rep->iters);
          q.rep.max_rel_residual = std::max(q.rep.max_rel_residual, rep->rel_residual);
        }
        if (--q.remaining == 0) q.rep.done_s = s.now();
      }
    }
    bool idle;
    {
      std::lock_guard<std::mutex> gr(r.mu);
      --r.inflight;
      r.last_used = s.now();
      idle = r.q.empty() && r.inflight == 0;
    }
    if (idle && !s.backlog.empty()) s.wake();      // a waiting request may now evict this layout
    s.done.notify_all();
  }

  void run(Resident &r) {
    StreamCtx cx{this, &r};
    const int32_t st = be.stream(be.user, r.handle, &c_pull, &c_waiting, &c_push, &cx);
    if (st != GP_OK) {
      const std::string msg = gp_last_error();
      std::lock_guard<std::mutex> g(mu);
      r.error = "layout " + std::to_string(r.key) + ": " + msg;
      done.notify_all();
    }
  }
};

using gp::fail;

extern "C" gp_status gp_server_create(const gp_server_backend *backend, const gp_solve_opts *opts, int32_t slots,
                                      int64_t budget_bytes, gp_server **out) {
  if (!out || budget_bytes <= 0 || slots < 1 || slots > 64) return fail(GP_ERR_ARG, "gp_server_create: bad arguments");
  if (backend && !(backend->estimate && backend->open && backend->bytes && backend->num_ports && backend->close &&
                   backend->stream))
    return fail(GP_ERR_ARG, "gp_server_create: every backend callback is required");
  auto s = std::make_unique<gp_server>();
  s->budget = budget_bytes;
  if (backend) {
    s->be = *backend;
  } else {
    s->lib = std::make_unique<LibBackend>();
    if (opts) s->lib->opts = *opts;
    else gp_solve_opts_default(&s->lib->opts);
    s->lib->slots = slots;
    s->be.user = s->lib.get();
    s->be.estimate = &lib_estimate;
    s->be.open = &lib_open;
    s->be.bytes = &lib_bytes;
    s->be.num_ports = &lib_num_ports;
    s->be.close = &lib_close;
    s->be.stream = &lib_stream;
  }
  gp_server *sp = s.get();
  s->admitter = std::thread([sp] { sp->admission_loop(); });
  *out = s.release();
  return GP_OK;
}

extern "C" gp_status gp_server_add_mesh(gp_server *srv, const gp_mesh *mesh, int32_t *key) {
  if (!srv || !mesh || !key) return fail(GP_ERR_ARG, "gp_server_add_mesh: null argument");
  if (!srv->lib) return fail(GP_ERR_STATE, "gp_server_add_mesh: the server has a caller backend");
  const gp_status vs = gp_mesh_validate(mesh);
  if (vs != GP_OK) return vs;
  std::lock_guard<std::mutex> g(srv->mu);
  *key = (int32_t)srv->lib->meshes.size();
  srv->lib->meshes.push_back(mesh);
  return GP_OK;
}

extern "C" gp_status gp_server_submit(gp_server *srv, int32_t key, const int32_t *ports, int32_t np,
                                      const double *freqs, int32_t nf, int64_t *request) {
  if (!srv || !ports || np < 1 || !freqs || nf < 1 || !request) return fail(GP_ERR_ARG, "gp_server_submit: bad arguments");
  std::lock_guard<std::mutex> g(srv->mu);
  if (!srv->error.empty()) return fail(GP_ERR_STATE, "gp_server_submit: " + srv->error);
  if (srv->lib && (key < 0 || key >= (int32_t)srv->lib->meshes.size()))
    return fail(GP_ERR_ARG, "gp_server_submit: unknown layout key");
  const int64_t rid = srv->next_rid++;
  Request &q = srv->reqs[rid];
  q.key = key;
  q.ports.assign(ports, ports + np);
  q.freqs.assign(freqs, freqs + nf);
  q.remaining = (int64_t)np * nf;
  q.rep.columns = np * nf;
  q.rep.arrival_s = srv->now();
  auto it = srv->res.find(key);
  bool older = false;
  for (int64_t b : srv->backlog)
    if (srv->reqs[b].key == key) { older = true; break; }
  if (it != srv->res.end() && !older) {
    srv->enqueue(*it->second, rid);               // resident: no wait for the admission thread
  } else {
    srv->backlog.push_back(rid);
    srv->wake();
  }
  *request = rid;
  return GP_OK;
}

extern "C" gp_status gp_server_wait(gp_server *srv, int64_t request, double timeout_s, double *Z,
                                    gp_request_report *rep) {
  if (!srv) return fail(GP_ERR_ARG, "gp_server_wait: null server");
  std::unique_lock<std::mutex> g(srv->mu);
  auto qi = srv->reqs.find(request);
  if (qi == srv->reqs.end()) return fail(GP_ERR_ARG, "gp_server_wait: unknown or already collected request");
  const auto end = Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(
                                      timeout_s < 0 ? 1e9 : timeout_s));
  for (;;) {
    if (!srv->error.empty()) return fail(GP_ERR_STATE, "gp_server_wait: admission: " + srv->error);
    for (auto &kv : srv->res)
      if (!kv.second->error.empty()) return fail(GP_ERR_STATE, "gp_server_wait: " + kv.second->error);
    auto cur = srv->reqs.find(request);  // another waiter may have collected it meanwhile
    if (cur == srv->reqs.end()) return fail(GP_ERR_STATE, "gp_server_wait: request collected by another waiter");
    Request &q = cur->second;
    if (q.remaining == 0) {
      if (rep) *rep = q.rep;
      if (Z) {
        std::copy(q.Z.begin(), q.Z.end(), Z);
        srv->reqs.erase(request);
      }
      return GP_OK;
    }
    if (Clock::now() >= end) return fail(GP_ERR_STATE, "gp_server_wait: timed out");
    srv->done.wait_until(g, std::chrono::system_clock::now() + std::chrono::milliseconds(50));  // end: steady clock
  }
}

extern "C" gp_status gp_server_info_get(gp_server *srv, gp_server_info *out) {
  if (!srv || !out) return fail(GP_ERR_ARG, "gp_server_info_get: null argument");
  std::lock_guard<std::mutex> g(srv->mu);
  *out = gp_server_info{};
  for (auto &kv : srv->res) out->resident_bytes += kv.second->nbytes;
  out->resident_bytes += srv->opening;
  out->resident_layouts = (int32_t)srv->res.size();
  out->evictions = srv->evictions;
  out->backlog = (int64_t)srv->backlog.size();
  return GP_OK;
}

extern "C" void gp_server_destroy(gp_server *srv) {
  if (!srv) return;
  {
    std::lock_guard<std::mutex> g(srv->mu);
    srv->stop = true;
    srv->kick.notify_all();
  }
  if (srv->admitter.joinable()) srv->admitter.join();
  std::vector<std::unique_ptr<Resident>> all;
  {
    std::lock_guard<std::mutex> g(srv->mu);
    for (auto &kv : srv->res) all.push_back(std::move(kv.second));
    srv->res.clear();
  }
  for (auto &r : all) srv->shutdown(*r);
  delete srv;
}

// Request loop for many BVPs on one GPU (continuous batching).
//
// The reference solves one BVP at a time per process (BEM_BVP::solve, BEM_solveBVP.hpp:624, called
// once per RK stage by main_time_domain.cpp:2095; solveForPhiPhin_t, :999, for the second BVP of a
// stage). Here a pool of worker threads drives one bemfmm context each at a time, so the solves of
// different requests overlap on the device (one HIP stream per context). Requests queue FIFO; a worker
// takes the oldest request that (a) is not behind another request of its session and (b) fits the
// device-memory budget, so requests arriving while others run are admitted as slots and memory free up.
// A session keeps its context (tree, operators, near field, preconditioner) resident between its
// requests: a request with the same geometry only refreshes the boundary values (the phi_t solve of an
// RK stage), one with moved nodes re-runs the setup in place.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bemfmm.h"
#include "../../include/bemfmm_server.h"

namespace {

double wall() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

struct Request {
  int64_t ticket = 0;
  int64_t session = -1;
  int n_nodes = 0, n_faces = 0, n_dof = 0;
  std::vector<double> X, phi, phin, x0;
  std::vector<int32_t> tri, face_dof, kind, point;
  bemfmm_opts opts{};
  double tol = 1e-10;
  int max_iter = 1000, restart = 1000;
  double est_bytes = 0;
  double reserved = 0;  // bytes reserved at admission for a context this request does not hold yet
  // results
  std::vector<double> x;
  bemfmm_result res{};
  int rc = 0;
  std::string err;
  double t_submit = 0, t_start = 0, t_end = 0;
  bool done = false;
};

struct Session {
  bemfmm_ctx* ctx = nullptr;
  std::vector<double> X;
  std::vector<int32_t> tri, face_dof, kind, point;
  bemfmm_opts opts{};
  double bytes = 0;       // device bytes held while resident
  bool busy = false;
  bool closing = false;
  double last_use = 0;    // for eviction of idle resident sessions
  int64_t pending = 0;    // requests queued or running
};

// field by field: the structure has padding bytes, which a byte comparison would include
bool same_opts(const bemfmm_opts& a, const bemfmm_opts& b) {
  return a.p == b.p && a.theta == b.theta && a.leaf_max == b.leaf_max && a.max_level == b.max_level && a.q == b.q &&
         a.precision == b.precision && a.preconditioner == b.preconditioner && a.refine_eta == b.refine_eta &&
         a.schwarz_max_core == b.schwarz_max_core && a.schwarz_max_block == b.schwarz_max_block &&
         a.schwarz_overlap == b.schwarz_overlap && a.m2l_tile_pairs == b.m2l_tile_pairs &&
         a.dof_reorder == b.dof_reorder && a.use_graphs == b.use_graphs && a.m2l_kernel == b.m2l_kernel &&
         a.overlap_near == b.overlap_near && a.refine_recycle == b.refine_recycle &&
         a.ilu_neighborhood == b.ilu_neighborhood && a.ilu_kring == b.ilu_kring &&
         a.schwarz_variant == b.schwarz_variant && a.schwarz_core_k == b.schwarz_core_k &&
         a.milu_omega == b.milu_omega && a.ilut_drop_tol == b.ilut_drop_tol && a.ilut_max_entries == b.ilut_max_entries &&
         a.ilut_pivot_min == b.ilut_pivot_min && a.plane_wave_digits == b.plane_wave_digits &&
         a.tree_on_gpu == b.tree_on_gpu && a.element_order == b.element_order;
}

}  // namespace

struct bemfmm_server {
  bemfmm_server_opts o{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::map<int64_t, std::shared_ptr<Request>> finished;
  std::set<int64_t> collected;  // tickets already returned by bemfmm_server_wait
  std::map<int64_t, Session> sessions;
  // Idle contexts kept for reuse: a finished request without a session, a closed session or a session whose
  // options changed leaves its context here, and a request that needs a context takes one with the same
  // options instead of creating one (its device buffers are grow-only, so a new geometry of a similar size
  // reuses them). Destroying a context frees its buffers, and every free synchronises the whole device, so
  // with other slots' work queued the pool saves more than the allocations. Pooled bytes stay in used_bytes;
  // under memory pressure pooled contexts are released first.
  struct Pooled {
    bemfmm_ctx* ctx;
    bemfmm_opts opts;
    double bytes;
    double last_use;
  };
  std::vector<Pooled> pool;
  std::vector<std::thread> workers;
  double used_bytes = 0, peak_bytes = 0;
  int running = 0;
  int64_t next_ticket = 1;
  bool stop = false;
  std::vector<double> latencies;  // seconds, completed requests
  int64_t n_failed = 0, n_evicted = 0;
  std::string err;

  // device bytes a new context of this problem will hold (plan + near field + Krylov basis), from the
  // measured footprint of the 200k-node tank (~72 KB per node in FP64, ~96 KB with the FP32 copies)
  static double estimate(const Request& r) {
    const double per_node = (r.opts.precision == BEMFMM_PRECISION_FP32_REFINE) ? 96e3 : 72e3;
    return per_node * r.n_nodes + 8.0 * r.n_dof * std::min(r.max_iter + 2, 512);
  }

  // next admissible request (FIFO); must hold mu
  std::shared_ptr<Request> pick() {
    for (auto it = queue.begin(); it != queue.end(); ++it) {
      auto& r = *it;
      double need = 0;
      if (r->session >= 0) {
        Session& s = sessions[r->session];
        if (s.busy) continue;
        // only the oldest queued request of a session may run
        bool older = false;
        for (auto jt = queue.begin(); jt != it; ++jt)
          if ((*jt)->session == r->session) older = true;
        if (older) continue;
        if (!s.ctx) need = r->est_bytes;
      } else {
        need = r->est_bytes;
      }
      // over budget: release idle resident sessions (least recently used first), else wait for memory
      while (used_bytes + need > o.hbm_budget_bytes && evict_one(r->session)) {
      }
      if (used_bytes + need > o.hbm_budget_bytes && running > 0) continue;
      auto out = r;
      queue.erase(it);
      used_bytes += need;
      out->reserved = need;
      peak_bytes = std::max(peak_bytes, used_bytes);
      if (out->session >= 0) sessions[out->session].busy = true;
      return out;
    }
    return nullptr;
  }

  // the most recently pooled context with these options (must hold mu); nullptr if none
  bemfmm_ctx* take_pooled_locked(const bemfmm_opts& opts, double& bytes) {
    for (size_t k = pool.size(); k-- > 0;)
      if (same_opts(pool[k].opts, opts)) {
        bemfmm_ctx* c = pool[k].ctx;
        bytes = pool[k].bytes;
        pool.erase(pool.begin() + (long)k);
        return c;
      }
    return nullptr;
  }

  void run(Request& r, Session* s) {
    auto fail = [&](bemfmm_ctx* c, int rc) {
      r.rc = rc;
      r.err = c ? bemfmm_last_error(c) : "bemfmm_ctx_create failed";
    };
    bemfmm_ctx* c = s ? s->ctx : nullptr;
    double held = s ? s->bytes : 0.;  // bytes accounted to c before this request
    bool fresh = false;
    if (c && !same_opts(s->opts, r.opts)) {  // options changed: the session's context goes to the pool
      std::lock_guard<std::mutex> lk(mu);
      pool.push_back(Pooled{c, s->opts, s->bytes, wall()});
      s->ctx = nullptr;
      s->bytes = 0;
      s->X.clear();
      c = nullptr;
      held = 0;
    }
    if (!c) {
      {
        std::lock_guard<std::mutex> lk(mu);
        c = take_pooled_locked(r.opts, held);
      }
      if (!c) {
        c = bemfmm_ctx_create(o.device, &r.opts);
        held = 0;
      }
      fresh = true;
      if (s) {
        s->ctx = c;
        s->opts = r.opts;
        s->bytes = held;
      }
    }
    int rc = BEMFMM_OK;
    if (!c) {
      fail(nullptr, BEMFMM_E_DEVICE);
      rc = BEMFMM_E_DEVICE;
    } else {
      const bool same_topology = s && !fresh && s->tri == r.tri && s->face_dof == r.face_dof &&
                                 s->kind == r.kind && s->point == r.point;
      const bool same_geometry = same_topology && s->X == r.X;
      if (!same_geometry) {
        rc = bemfmm_set_geometry(c, r.n_nodes, r.X.data(), r.n_faces, r.tri.data(), r.face_dof.data(), r.n_dof,
                                 r.kind.data(), r.point.data());
        if (rc == BEMFMM_OK) rc = bemfmm_set_bc(c, r.phi.data(), r.phin.data());
        if (rc == BEMFMM_OK) rc = bemfmm_prepare(c);
        if (s) {  // a failed setup leaves no geometry the next request could reuse
          s->X = rc == BEMFMM_OK ? r.X : std::vector<double>();
          s->tri = r.tri;
          s->face_dof = r.face_dof;
          s->kind = r.kind;
          s->point = r.point;
        }
      } else {
        rc = bemfmm_set_bc(c, r.phi.data(), r.phin.data());  // same geometry: only the knowns change
      }
      if (rc == BEMFMM_OK) {
        r.x.assign(r.n_dof, 0.);
        if (r.x0.empty()) r.x0.assign(r.n_dof, 0.);
        rc = bemfmm_solve(c, r.x0.data(), r.x.data(), r.tol, r.max_iter, r.restart, &r.res);
        if (rc == BEMFMM_E_NOT_CONVERGED) rc = BEMFMM_OK;  // reported through res.converged
      }
      if (rc != BEMFMM_OK) fail(c, rc);
    }
    // accounting: the admission reservation is replaced by the context's measured footprint
    bemfmm_stats st{};
    const double actual = (c && bemfmm_get_stats(c, &st) == BEMFMM_OK)
                              ? st.device_bytes + 8.0 * r.n_dof * std::min(r.max_iter + 2, 512)
                              : held;
    std::lock_guard<std::mutex> lk(mu);
    used_bytes += actual - held - r.reserved;
    r.reserved = 0;
    peak_bytes = std::max(peak_bytes, used_bytes);
    if (s) {
      s->bytes = c ? actual : 0.;
    } else if (c) {
      pool.push_back(Pooled{c, r.opts, actual, wall()});
    }
  }

  void worker() {
    if (o.omp_threads_per_slot > 0) omp_set_num_threads(o.omp_threads_per_slot);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      std::shared_ptr<Request> r;
      cv_work.wait(lk, [&] { return stop || (r = pick()) != nullptr; });
      if (!r) return;  // stop with nothing admissible
      ++running;
      Session* s = r->session >= 0 ? &sessions[r->session] : nullptr;
      r->t_start = wall();
      lk.unlock();
      run(*r, s);
      lk.lock();
      r->t_end = wall();
      --running;
      if (s) {
        s->busy = false;
        s->last_use = r->t_end;
        if (--s->pending == 0 && s->closing) close_locked(r->session);
      }
      r->done = true;
      if (r->rc == BEMFMM_OK) latencies.push_back(r->t_end - r->t_submit);
      else ++n_failed;
      finished[r->ticket] = r;
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }

  // frees the least recently used pooled context, else the context of the least recently used idle session
  // (its next request sets it up again)
  bool evict_one(int64_t keep) {
    if (!pool.empty()) {
      size_t k = 0;
      for (size_t i = 1; i < pool.size(); ++i)
        if (pool[i].last_use < pool[k].last_use) k = i;
      bemfmm_ctx_destroy(pool[k].ctx);
      used_bytes -= pool[k].bytes;
      pool.erase(pool.begin() + (long)k);
      ++n_evicted;
      return true;
    }
    Session* best = nullptr;
    for (auto& kv : sessions) {
      Session& ss = kv.second;
      if (kv.first == keep || ss.busy || !ss.ctx) continue;
      if (!best || ss.last_use < best->last_use) best = &ss;
    }
    if (!best) return false;
    bemfmm_ctx_destroy(best->ctx);
    best->ctx = nullptr;
    best->X.clear();
    used_bytes -= best->bytes;
    best->bytes = 0;
    ++n_evicted;
    return true;
  }

  // a closed session's context goes to the pool (its bytes stay accounted there)
  void close_locked(int64_t id) {
    auto it = sessions.find(id);
    if (it == sessions.end()) return;
    if (it->second.ctx) pool.push_back(Pooled{it->second.ctx, it->second.opts, it->second.bytes, wall()});
    else used_bytes -= it->second.bytes;
    sessions.erase(it);
  }
};

extern "C" {

void bemfmm_server_default_opts(bemfmm_server_opts* o) {
  if (!o) return;
  o->device = 0;
  o->slots = 4;
  o->hbm_budget_bytes = 200e9;
  o->omp_threads_per_slot = 0;
}

bemfmm_server* bemfmm_server_create(const bemfmm_server_opts* opts) {
  auto* s = new bemfmm_server();
  if (opts) s->o = *opts;
  else bemfmm_server_default_opts(&s->o);
  if (s->o.slots < 1 || s->o.slots > 64 || !(s->o.hbm_budget_bytes > 0)) {
    delete s;
    return nullptr;
  }
  // host threads for each request's setup: the process's OpenMP thread budget (OMP_NUM_THREADS, else the
  // processors available) shared by the slots
  if (s->o.omp_threads_per_slot <= 0) s->o.omp_threads_per_slot = std::max(1, omp_get_max_threads() / s->o.slots);
  for (int i = 0; i < s->o.slots; ++i) s->workers.emplace_back([s] { s->worker(); });
  return s;
}

void bemfmm_server_destroy(bemfmm_server* s) {
  if (!s) return;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    // finish what is queued, then stop
    s->cv_done.wait(lk, [&] { return s->queue.empty() && s->running == 0; });
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->workers) t.join();
  for (auto& kv : s->sessions)
    if (kv.second.ctx) bemfmm_ctx_destroy(kv.second.ctx);
  for (auto& pc : s->pool) bemfmm_ctx_destroy(pc.ctx);
  delete s;
}

int64_t bemfmm_server_submit(bemfmm_server* s, const bemfmm_problem* p, int64_t session) {
  if (!s || !p) return BEMFMM_E_INVALID;
  if (p->n_nodes <= 0 || p->n_faces <= 0 || p->n_dof <= 0 || !p->X || !p->tri || !p->face_dof || !p->dof_kind ||
      !p->dof_point || !p->phi || !p->phin)
    return BEMFMM_E_INVALID;
  auto r = std::make_shared<Request>();
  r->session = session;
  r->n_nodes = p->n_nodes;
  r->n_faces = p->n_faces;
  r->n_dof = p->n_dof;
  r->X.assign(p->X, p->X + 3 * (size_t)p->n_nodes);
  r->tri.assign(p->tri, p->tri + 3 * (size_t)p->n_faces);
  r->face_dof.assign(p->face_dof, p->face_dof + 3 * (size_t)p->n_faces);
  r->kind.assign(p->dof_kind, p->dof_kind + p->n_dof);
  r->point.assign(p->dof_point, p->dof_point + p->n_dof);
  r->phi.assign(p->phi, p->phi + p->n_dof);
  r->phin.assign(p->phin, p->phin + p->n_dof);
  if (p->x0) r->x0.assign(p->x0, p->x0 + p->n_dof);
  r->opts = p->opts;
  r->tol = p->tol;
  r->max_iter = p->max_iter;
  r->restart = p->restart;
  r->est_bytes = bemfmm_server::estimate(*r);
  r->t_submit = wall();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (session >= 0) {
      Session& ss = s->sessions[session];
      if (ss.closing) return BEMFMM_E_STATE;
      ss.pending++;
    }
    r->ticket = s->next_ticket++;
    s->queue.push_back(r);
  }
  s->cv_work.notify_all();
  return r->ticket;
}

int bemfmm_server_wait(bemfmm_server* s, int64_t ticket, double* x, bemfmm_result* res, double* times3) {
  if (!s || ticket <= 0) return BEMFMM_E_INVALID;
  std::shared_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    // unknown tickets and tickets already collected would never finish
    if (ticket >= s->next_ticket || s->collected.count(ticket)) return BEMFMM_E_INVALID;
    s->collected.insert(ticket);
    s->cv_done.wait(lk, [&] { return s->finished.count(ticket) > 0; });
    r = s->finished[ticket];
    s->finished.erase(ticket);
  }
  if (r->rc != BEMFMM_OK) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->err = r->err;
    return r->rc;
  }
  if (x) std::copy(r->x.begin(), r->x.end(), x);
  if (res) *res = r->res;
  if (times3) {
    times3[0] = r->t_start - r->t_submit;  // queueing
    times3[1] = r->t_end - r->t_start;     // service (setup if any + solve)
    times3[2] = r->t_end - r->t_submit;    // latency
  }
  return BEMFMM_OK;
}

int bemfmm_server_close_session(bemfmm_server* s, int64_t session) {
  if (!s || session < 0) return BEMFMM_E_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->sessions.find(session);
  if (it == s->sessions.end()) return BEMFMM_E_INVALID;
  if (it->second.pending == 0) s->close_locked(session);
  else it->second.closing = true;  // released after its last request
  return BEMFMM_OK;
}

int bemfmm_server_get_stats(bemfmm_server* s, bemfmm_server_stats* st) {
  if (!s || !st) return BEMFMM_E_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  std::vector<double> l = s->latencies;
  std::sort(l.begin(), l.end());
  auto pct = [&](double q) {
    if (l.empty()) return 0.0;
    const double pos = q * (double)(l.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < l.size() ? l[i] * (1 - f) + l[i + 1] * f : l[i];
  };
  st->completed = (int64_t)l.size();
  st->failed = s->n_failed;
  st->queued = (int64_t)s->queue.size();
  st->running = s->running;
  st->resident_sessions = (int)s->sessions.size();
  st->pooled_contexts = (int)s->pool.size();
  st->p50_latency_s = pct(0.50);
  st->p99_latency_s = pct(0.99);
  st->used_bytes = s->used_bytes;
  st->peak_bytes = s->peak_bytes;
  st->evicted = s->n_evicted;
  return BEMFMM_OK;
}

const char* bemfmm_server_last_error(bemfmm_server* s) { return s ? s->err.c_str() : "null server"; }

}  // extern "C"

// cv_serve.hip -- continuous-batching request loop for independent cases on
// one GPU.  The reference has no counterpart: it runs one case per process
// (main.c:1094-1170).
//
// A request is one case (grid, nvert, initial fluxes, solver options) and a
// step count.  Requests queue in arrival order.  A request is admitted when
// a slot is free and its context fits the HBM budget next to the cases in
// flight; its fields then stay resident on the device until its last step.
// Every admitted case owns one context and one HIP stream.  Two modes:
//  * streams: one worker thread per slot steps its case with cv_flow_step,
//    so the kernels of all cases in flight overlap on the GPU;
//  * batched: one scheduler thread advances all admitted cases together,
//    one cv_batch per group of cases with equal solver options, rebuilt
//    whenever cases join or leave; a case the batch cannot take is stepped
//    alone in the same tick.
// A case's results do not depend on what else runs (each context's
// arithmetic and reduction order is fixed, and a batch reproduces the
// single-case bits).  Each step's wall-clock latency is recorded.
#include "cv_internal.h"
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <thread>

namespace {

using clk = std::chrono::steady_clock;

struct Case {
  long id = 0;
  cv_grid_desc g{};
  cv_solver_opts o{};
  std::vector<double> xyz, nv, uc;
  std::vector<cv_ib_node> ib;   // IB node records (none: IB cells held)
  int ib_mode = CV_IB_NOSLIP;
  int nsteps = 0;
  long bytes = 0;     // resident device bytes of the case's context
  long scratch = 0;   // cv_set_geometry's transient scratch, held from admission until the metrics are built
  double t_submit = 0, t_admit = 0, t_done = 0;
  int status = CV_CASE_QUEUED;
  int rc = CV_OK;
  std::vector<double> step_ms;
  std::vector<int> step_its;
  std::vector<double> p, ucont;
  cv_step_stats last{};
};

}  // namespace

struct cv_server {
  int device = 0;
  int mode = CV_SERVE_STREAMS;
  long budget = 0, in_use = 0, peak = 0;
  int slots = 0, running = 0;
  bool stop = false;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<long> pending;
  std::map<long, std::unique_ptr<Case>> cases;
  long next_id = 1;
  long done = 0;
  std::vector<std::thread> workers;
  clk::time_point t0;
  double now() const { return std::chrono::duration<double>(clk::now() - t0).count(); }
};

// the admission reserved the geometry scratch beside the context; once the
// metrics are built (or failed) it is free again
static void release_scratch(cv_server *S, Case *cs) {
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->in_use -= cs->scratch;
    cs->scratch = 0;
  }
  S->cv_work.notify_all();
}

// the case's device context from its inputs (which are then dropped)
static int open_case(cv_server *S, Case *cs, cv_ctx **out) {
  (void)hipSetDevice(S->device);
  cv_ctx *c = nullptr;
  int rc = cv_ctx_create(&cs->g, S->device, &c);
  if (rc == CV_OK) rc = cv_set_solver(c, &cs->o);
  if (rc == CV_OK) rc = cv_set_geometry(c, cs->xyz.data());
  release_scratch(S, cs);
  if (rc == CV_OK) rc = cv_set_nvert(c, cs->nv.data());
  if (rc == CV_OK && !cs->ib.empty()) rc = cv_set_ib_nodes(c, cs->ib.data(), (int)cs->ib.size(), cs->ib_mode);
  if (rc == CV_OK) rc = cv_put_field(c, CV_F_UCONT, -1, cs->uc.data());
  std::vector<double>().swap(cs->xyz);
  std::vector<double>().swap(cs->nv);
  std::vector<double>().swap(cs->uc);
  std::vector<cv_ib_node>().swap(cs->ib);
  *out = c;
  return rc;
}

// the final p and Ucont to the case record, then the context goes
static int close_case(Case *cs, cv_ctx *c, int rc) {
  if (rc == CV_OK && c) {
    const long n = (long)(cs->g.IM + 1) * (cs->g.JM + 1) * (cs->g.KM + 1);
    cs->p.resize(n);
    cs->ucont.resize(3 * n);
    rc = cv_get_field(c, CV_F_P, -1, cs->p.data());
    if (rc == CV_OK) rc = cv_get_field(c, CV_F_UCONT, -1, cs->ucont.data());
  }
  if (c) cv_ctx_destroy(c);
  return rc;
}

static int run_case(cv_server *S, Case *cs) {
  cv_ctx *c = nullptr;
  int rc = open_case(S, cs, &c);
  for (int s = 0; rc == CV_OK && s < cs->nsteps; s++) {
    cv_step_stats st{};
    const auto a = clk::now();
    int r = cv_flow_step(c, &st);
    const double ms = 1e3 * std::chrono::duration<double>(clk::now() - a).count();
    if (r != CV_OK && r != CV_ERR_NOT_CONVERGED) { rc = r; break; }
    std::lock_guard<std::mutex> lk(S->mu);
    cs->step_ms.push_back(ms);
    cs->step_its.push_back(st.ksp_its);
    cs->last = st;
    if (r == CV_ERR_NOT_CONVERGED) cs->rc = r;
  }
  return close_case(cs, c, rc);
}

// bookkeeping of a finished case (lock held by the caller)
static void retire(cv_server *S, Case *cs, int rc) {
  S->in_use -= cs->bytes;
  S->running--;
  S->done++;
  if (rc != CV_OK) cs->rc = rc;
  cs->t_done = S->now();
  cs->status = rc == CV_OK ? CV_CASE_DONE : CV_CASE_FAILED;
}

// admission of the head of the queue (lock held): a free slot and room in
// the budget
static Case *admit_one(cv_server *S) {
  if (S->pending.empty() || S->running >= S->slots) return nullptr;
  Case *h = S->cases[S->pending.front()].get();
  if (S->in_use + h->bytes + h->scratch > S->budget) return nullptr;
  S->pending.pop_front();
  S->in_use += h->bytes + h->scratch;
  if (S->in_use > S->peak) S->peak = S->in_use;
  S->running++;
  h->status = CV_CASE_RUNNING;
  h->t_admit = S->now();
  return h;
}

static void worker(cv_server *S) {
  for (;;) {
    Case *cs;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait(lk, [&] {
        if (S->stop) return true;
        if (S->pending.empty()) return false;
        const Case *h = S->cases[S->pending.front()].get();
        return S->in_use + h->bytes + h->scratch <= S->budget;
      });
      if (S->stop) return;
      const long id = S->pending.front();
      S->pending.pop_front();
      cs = S->cases[id].get();
      S->in_use += cs->bytes + cs->scratch;
      if (S->in_use > S->peak) S->peak = S->in_use;
      S->running++;
      cs->status = CV_CASE_RUNNING;
      cs->t_admit = S->now();
    }
    const int rc = run_case(S, cs);
    {
      std::lock_guard<std::mutex> lk(S->mu);
      retire(S, cs, rc);
    }
    S->cv_done.notify_all();
    S->cv_work.notify_all();
  }
}

// ---------------- batched mode -----------------------------------------
namespace {
struct Live {
  Case *cs;
  cv_ctx *c;
  int steps = 0;
};
struct Group {
  cv_solver_opts o{};
  std::vector<Live *> members;
  cv_batch *b = nullptr;
  bool batchable = true;
};
// field by field: the struct has padding a caller need not clear
bool same_solver(const cv_solver_opts &a, const cv_solver_opts &b) {
  return a.precision == b.precision && a.restart == b.restart && a.max_it == b.max_it && a.rtol == b.rtol &&
         a.atol == b.atol && a.pre_sweeps == b.pre_sweeps && a.post_sweeps == b.post_sweeps &&
         a.coarse_sweeps == b.coarse_sweeps && a.smoother == b.smoother && a.omega == b.omega;
}
}  // namespace

// record one step of a case (lock taken here)
static void record(cv_server *S, Live *l, double ms, const cv_step_stats &st, int r) {
  std::lock_guard<std::mutex> lk(S->mu);
  l->cs->step_ms.push_back(ms);
  l->cs->step_its.push_back(st.ksp_its);
  l->cs->last = st;
  if (r == CV_ERR_NOT_CONVERGED) l->cs->rc = r;
  l->steps++;
}

static void scheduler(cv_server *S) {
  std::vector<std::unique_ptr<Live>> live;
  std::vector<Group> groups;
  bool changed = false;
  for (;;) {
    std::vector<Case *> admitted;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait(lk, [&] {
        if (S->stop || !live.empty()) return true;
        if (S->pending.empty()) return false;
        const Case *h = S->cases[S->pending.front()].get();
        return S->in_use + h->bytes + h->scratch <= S->budget;
      });
      if (S->stop && live.empty()) break;
      if (!S->stop)   // at destroy: finish the cases in flight, admit no more
        while (Case *h = admit_one(S)) admitted.push_back(h);
    }
    for (Case *cs : admitted) {
      cv_ctx *c = nullptr;
      const int rc = open_case(S, cs, &c);
      if (rc != CV_OK) {
        close_case(cs, c, rc);
        std::lock_guard<std::mutex> lk(S->mu);
        retire(S, cs, rc);
        S->cv_done.notify_all();
        continue;
      }
      live.push_back(std::make_unique<Live>(Live{cs, c, 0}));
      changed = true;
    }
    if (changed) {   // regroup by solver options and rebuild the batches
      for (auto &g : groups)
        if (g.b) cv_batch_destroy(g.b);
      groups.clear();
      for (auto &l : live) {
        Group *g = nullptr;
        for (auto &x : groups)
          if (same_solver(x.o, l->cs->o) && x.members.size() < 64) { g = &x; break; }
        if (!g) { groups.emplace_back(); g = &groups.back(); g->o = l->cs->o; }
        g->members.push_back(l.get());
      }
      for (auto &g : groups) {
        std::vector<cv_ctx *> cs;
        for (Live *l : g.members) cs.push_back(l->c);
        g.batchable = cv_batch_create(cs.data(), (int)cs.size(), &g.b) == CV_OK;
        if (!g.batchable) g.b = nullptr;
      }
      changed = false;
    }
    // one step of every live case
    std::vector<Live *> failed;
    for (auto &g : groups) {
      if (g.batchable) {
        std::vector<cv_step_stats> st(g.members.size());
        const auto a = clk::now();
        const int r = cv_batch_step(g.b, st.data());
        const double ms = 1e3 * std::chrono::duration<double>(clk::now() - a).count();
        if (r != CV_OK && r != CV_ERR_NOT_CONVERGED && r != CV_ERR_NONFINITE) {
          for (Live *l : g.members) { l->cs->rc = r; failed.push_back(l); }
          continue;
        }
        // a case that went non-finite fails alone; the batch carried the
        // others through the step with the bits they get without it
        for (size_t q = 0; q < g.members.size(); q++) {
          const cv_step_stats &x = st[q];
          if (!std::isfinite(x.maxdiv) || !std::isfinite(x.ksp_bnorm) || !std::isfinite(x.ksp_rnorm)) {
            g.members[q]->cs->rc = CV_ERR_NONFINITE;
            failed.push_back(g.members[q]);
            continue;
          }
          record(S, g.members[q], ms, x, x.converged ? CV_OK : CV_ERR_NOT_CONVERGED);
        }
      } else {
        for (Live *l : g.members) {
          cv_step_stats st{};
          const auto a = clk::now();
          const int r = cv_flow_step(l->c, &st);
          const double ms = 1e3 * std::chrono::duration<double>(clk::now() - a).count();
          if (r != CV_OK && r != CV_ERR_NOT_CONVERGED) { l->cs->rc = r; failed.push_back(l); continue; }
          record(S, l, ms, st, r);
        }
      }
    }
    // retire finished and failed cases
    for (size_t q = 0; q < live.size();) {
      Live *l = live[q].get();
      const bool bad = std::find(failed.begin(), failed.end(), l) != failed.end();
      if (!bad && l->steps < l->cs->nsteps) { q++; continue; }
      const int rc = close_case(l->cs, l->c, bad ? l->cs->rc : CV_OK);
      {
        std::lock_guard<std::mutex> lk(S->mu);
        retire(S, l->cs, bad ? (rc != CV_OK ? rc : l->cs->rc) : rc);
      }
      S->cv_done.notify_all();
      live.erase(live.begin() + q);
      changed = true;
    }
    if (changed)
      for (auto &g : groups) {   // contexts may be gone: drop the batches now
        if (g.b) cv_batch_destroy(g.b);
        g.b = nullptr;
      }
  }
  for (auto &g : groups)
    if (g.b) cv_batch_destroy(g.b);
  S->cv_done.notify_all();
}

extern "C" int cv_server_create_mode(int device, long hbm_budget_bytes, int slots, int mode, cv_server **out) {
  if (!out || hbm_budget_bytes <= 0 || slots < 1 || slots > 64) return CV_ERR_ARG;
  if (mode != CV_SERVE_STREAMS && mode != CV_SERVE_BATCHED) return CV_ERR_ARG;
  *out = nullptr;
  CV_CHECK(hipSetDevice(device));
  cv_server *S = new cv_server();
  S->device = device;
  S->mode = mode;
  S->budget = hbm_budget_bytes;
  S->slots = slots;
  S->t0 = clk::now();
  if (mode == CV_SERVE_BATCHED) S->workers.emplace_back(scheduler, S);
  else
    for (int q = 0; q < slots; q++) S->workers.emplace_back(worker, S);
  *out = S;
  return CV_OK;
}

extern "C" int cv_server_create(int device, long hbm_budget_bytes, int slots, cv_server **out) {
  return cv_server_create_mode(device, hbm_budget_bytes, slots, CV_SERVE_STREAMS, out);
}

extern "C" int cv_server_submit(cv_server *S, const cv_grid_desc *g, const cv_solver_opts *o, const double *xyz,
                                const double *nvert, const double *ucont, int nsteps, long *case_id) {
  return cv_server_submit_ib(S, g, o, xyz, nvert, ucont, nsteps, nullptr, 0, CV_IB_NOSLIP, case_id);
}

// device bytes of the IB node buffers and the face and cell lists: at most
// 6 faces and 7 cells per node
static long ib_bytes(int n) { return (long)n * (4 * 4 + 12 * 8 + 4 * 8 + 6 * 4 + 7 * 4); }

extern "C" int cv_server_submit_ib(cv_server *S, const cv_grid_desc *g, const cv_solver_opts *o, const double *xyz,
                                   const double *nvert, const double *ucont, int nsteps, const cv_ib_node *ib,
                                   int n_ib, int ib_mode, long *case_id) {
  if (!S || !g || !o || !xyz || !nvert || !ucont || nsteps < 1 || !case_id || n_ib < 0 || (n_ib > 0 && !ib) ||
      (ib_mode != CV_IB_NOSLIP && ib_mode != CV_IB_FREESLIP))
    return CV_ERR_ARG;
  long bytes = cv_ctx_bytes(g, o->restart);
  if (bytes < 0) return CV_ERR_ARG;
  bytes += ib_bytes(n_ib);
  if (o->precision == CV_MIXED16) bytes += cv_half_bytes(g);
  const long scratch = cv_geometry_scratch_bytes(g);
  if (bytes + scratch > S->budget) {
    cv_set_error("cv_server_submit: case needs %ld bytes (%ld while its metrics are built), budget is %ld", bytes,
                 bytes + scratch, S->budget);
    return CV_ERR_OOM;
  }
  auto cs = std::make_unique<Case>();
  cs->g = *g;
  cs->o = *o;
  const size_t nn = (size_t)g->IM * g->JM * g->KM, ns = (size_t)(g->IM + 1) * (g->JM + 1) * (g->KM + 1);
  cs->xyz.assign(xyz, xyz + 3 * nn);
  cs->nv.assign(nvert, nvert + ns);
  cs->uc.assign(ucont, ucont + 3 * ns);
  if (n_ib > 0) cs->ib.assign(ib, ib + n_ib);
  cs->ib_mode = ib_mode;
  cs->nsteps = nsteps;
  cs->bytes = bytes;
  cs->scratch = scratch;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    cs->id = S->next_id++;
    cs->t_submit = S->now();
    *case_id = cs->id;
    S->pending.push_back(cs->id);
    S->cases[cs->id] = std::move(cs);
  }
  S->cv_work.notify_all();
  return CV_OK;
}

static Case *find_case(cv_server *S, long id) {
  auto it = S->cases.find(id);
  return it == S->cases.end() ? nullptr : it->second.get();
}

extern "C" int cv_server_wait(cv_server *S, long id, cv_case_info *info) {
  if (!S) return CV_ERR_ARG;
  std::unique_lock<std::mutex> lk(S->mu);
  Case *cs = find_case(S, id);
  if (!cs) return CV_ERR_ARG;
  S->cv_done.wait(lk, [&] { return cs->status == CV_CASE_DONE || cs->status == CV_CASE_FAILED; });
  if (info) {
    info->status = cs->status;
    info->rc = cs->rc;
    info->steps_done = (int)cs->step_ms.size();
    info->bytes = cs->bytes;
    info->t_submit = cs->t_submit;
    info->t_admit = cs->t_admit;
    info->t_done = cs->t_done;
    info->last = cs->last;
  }
  return cs->status == CV_CASE_DONE ? CV_OK : cs->rc;
}

extern "C" int cv_server_step_ms(cv_server *S, long id, double *ms, int *its, int max) {
  if (!S || max < 0) return CV_ERR_ARG;
  std::lock_guard<std::mutex> lk(S->mu);
  Case *cs = find_case(S, id);
  if (!cs) return CV_ERR_ARG;
  const int n = (int)cs->step_ms.size() < max ? (int)cs->step_ms.size() : max;
  for (int q = 0; q < n; q++) {
    if (ms) ms[q] = cs->step_ms[q];
    if (its) its[q] = cs->step_its[q];
  }
  return n;
}

extern "C" int cv_server_fetch(cv_server *S, long id, int field, double *out) {
  if (!S || !out) return CV_ERR_ARG;
  std::lock_guard<std::mutex> lk(S->mu);
  Case *cs = find_case(S, id);
  if (!cs || cs->status != CV_CASE_DONE) return CV_ERR_STATE;
  if (field == CV_F_P) std::copy(cs->p.begin(), cs->p.end(), out);
  else if (field == CV_F_UCONT) std::copy(cs->ucont.begin(), cs->ucont.end(), out);
  else return CV_ERR_ARG;
  return CV_OK;
}

extern "C" int cv_server_release(cv_server *S, long id) {
  if (!S) return CV_ERR_ARG;
  std::lock_guard<std::mutex> lk(S->mu);
  Case *cs = find_case(S, id);
  if (!cs || (cs->status != CV_CASE_DONE && cs->status != CV_CASE_FAILED)) return CV_ERR_STATE;
  S->cases.erase(id);
  return CV_OK;
}

extern "C" int cv_server_status(cv_server *S, cv_server_info *info) {
  if (!S || !info) return CV_ERR_ARG;
  std::lock_guard<std::mutex> lk(S->mu);
  info->queued = (long)S->pending.size();
  info->running = S->running;
  info->done = S->done;
  info->bytes_in_use = S->in_use;
  info->bytes_peak = S->peak;
  info->budget = S->budget;
  info->time = S->now();
  return CV_OK;
}

extern "C" int cv_server_destroy(cv_server *S) {
  if (!S) return CV_ERR_ARG;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  for (auto &t : S->workers) t.join();
  delete S;
  return CV_OK;
}

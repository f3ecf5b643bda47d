// cv_serve.hip -- continuous-batching request loop for independent cases on
// one GPU.  The reference has no counterpart: it runs one case per process
// (main.c:1094-1170).
//
// A request is one case (grid, nvert, initial fluxes, solver options) and a
// step count.  Requests queue in arrival order.  A request is admitted when
// a slot is free and its context fits the HBM budget next to the cases in
// flight; its fields then stay resident on the device until its last step.
// Every admitted case owns one context and one HIP stream, so the kernels
// of all cases in flight overlap on the GPU, and a case's results do not
// depend on what else runs (each context's arithmetic and reduction order
// is fixed).  Each step's wall-clock latency is recorded.
#include "cv_internal.h"
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

namespace {

using clk = std::chrono::steady_clock;

struct Case {
  long id = 0;
  cv_grid_desc g{};
  cv_solver_opts o{};
  std::vector<double> xyz, nv, uc;
  int nsteps = 0;
  long bytes = 0;
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

static int run_case(cv_server *S, Case *cs) {
  (void)hipSetDevice(S->device);
  cv_ctx *c = nullptr;
  int rc = cv_ctx_create(&cs->g, S->device, &c);
  if (rc == CV_OK) rc = cv_set_solver(c, &cs->o);
  if (rc == CV_OK) rc = cv_set_geometry(c, cs->xyz.data());
  if (rc == CV_OK) rc = cv_set_nvert(c, cs->nv.data());
  if (rc == CV_OK) rc = cv_put_field(c, CV_F_UCONT, -1, cs->uc.data());
  std::vector<double>().swap(cs->xyz);
  std::vector<double>().swap(cs->nv);
  std::vector<double>().swap(cs->uc);
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
  if (rc == CV_OK) {
    const long n = (long)(cs->g.IM + 1) * (cs->g.JM + 1) * (cs->g.KM + 1);
    cs->p.resize(n);
    cs->ucont.resize(3 * n);
    rc = cv_get_field(c, CV_F_P, -1, cs->p.data());
    if (rc == CV_OK) rc = cv_get_field(c, CV_F_UCONT, -1, cs->ucont.data());
  }
  if (c) cv_ctx_destroy(c);
  return rc;
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
        return S->in_use + h->bytes <= S->budget;
      });
      if (S->stop) return;
      const long id = S->pending.front();
      S->pending.pop_front();
      cs = S->cases[id].get();
      S->in_use += cs->bytes;
      if (S->in_use > S->peak) S->peak = S->in_use;
      S->running++;
      cs->status = CV_CASE_RUNNING;
      cs->t_admit = S->now();
    }
    const int rc = run_case(S, cs);
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->in_use -= cs->bytes;
      S->running--;
      S->done++;
      if (rc != CV_OK) cs->rc = rc;
      cs->t_done = S->now();
      cs->status = rc == CV_OK ? CV_CASE_DONE : CV_CASE_FAILED;
    }
    S->cv_done.notify_all();
    S->cv_work.notify_all();
  }
}

extern "C" int cv_server_create(int device, long hbm_budget_bytes, int slots, cv_server **out) {
  if (!out || hbm_budget_bytes <= 0 || slots < 1 || slots > 64) return CV_ERR_ARG;
  *out = nullptr;
  CV_CHECK(hipSetDevice(device));
  cv_server *S = new cv_server();
  S->device = device;
  S->budget = hbm_budget_bytes;
  S->slots = slots;
  S->t0 = clk::now();
  for (int q = 0; q < slots; q++) S->workers.emplace_back(worker, S);
  *out = S;
  return CV_OK;
}

extern "C" int cv_server_submit(cv_server *S, const cv_grid_desc *g, const cv_solver_opts *o, const double *xyz,
                                const double *nvert, const double *ucont, int nsteps, long *case_id) {
  if (!S || !g || !o || !xyz || !nvert || !ucont || nsteps < 1 || !case_id) return CV_ERR_ARG;
  const long bytes = cv_ctx_bytes(g, o->restart);
  if (bytes < 0) return CV_ERR_ARG;
  if (bytes > S->budget) {
    cv_set_error("cv_server_submit: case needs %ld bytes, budget is %ld", bytes, S->budget);
    return CV_ERR_OOM;
  }
  auto cs = std::make_unique<Case>();
  cs->g = *g;
  cs->o = *o;
  const size_t nn = (size_t)g->IM * g->JM * g->KM, ns = (size_t)(g->IM + 1) * (g->JM + 1) * (g->KM + 1);
  cs->xyz.assign(xyz, xyz + 3 * nn);
  cs->nv.assign(nvert, nvert + ns);
  cs->uc.assign(ucont, ucont + 3 * ns);
  cs->nsteps = nsteps;
  cs->bytes = bytes;
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

// Serving loop: fit requests and assignment queries against resident
// models, admitted under an HBM budget and run on worker contexts.
//
// Each worker owns one context (HIP stream + workspace) and grow-only
// device staging buffers for its requests' series and results.  Fit
// workers take the oldest fit (or, when no fit waits, the oldest query);
// query workers take only queries, so a burst of long fits cannot delay
// short queries.  A request is admitted when its staging bytes fit in the
// budget left after resident models and the requests already running; it
// waits in its queue otherwise (first come, first served per queue).
// No reference counterpart.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "tcx_host.hpp"
#include "tcx_internal.hpp"
#include "timesc.h"

using Clock = std::chrono::steady_clock;

namespace {

struct Model {
  int k = 0, m = 0;
  float* d_cent = nullptr;
  void* d_spec = nullptr;
  float* d_mag = nullptr;
  float* d_norm = nullptr;
  int* d_zero = nullptr;
  size_t bytes = 0;
};

struct Pending {
  tcx_request* req;
  Clock::time_point t_submit;
  size_t bytes;
};

struct Worker {
  tcx_ctx* ctx = nullptr;
  bool query_only = false;
  float* d_x = nullptr;
  size_t x_cap = 0;
  int* d_i = nullptr;  // labels, then shifts
  float* d_f = nullptr;  // distances
  size_t n_cap = 0;
  std::thread th;
};

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

size_t request_bytes(const tcx_request& r) {
  return (size_t)r.n * r.m * sizeof(float) + (size_t)r.n * 3 * sizeof(int);
}

}  // namespace

struct tcx_server {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Pending> fitq, asgq;
  std::vector<Model> models;
  std::vector<Worker*> workers;
  tcx_ctx* admin = nullptr;  // uploads models (under admin_mu)
  std::mutex admin_mu;
  int fit_workers = 0;
  size_t budget = 0, used = 0;
  int running = 0;
  bool stop = false;
};

namespace {

// Grows the worker's staging buffers to hold n series of length m.
int stage(Worker& w, long n, int m) {
  const size_t xb = (size_t)n * m * sizeof(float);
  if (xb > w.x_cap) {
    if (w.d_x) tcx_free(w.ctx, w.d_x);
    w.d_x = nullptr;
    w.x_cap = 0;
    int rc = tcx_malloc(w.ctx, (void**)&w.d_x, xb);
    if (rc != TCX_OK) return rc;
    w.x_cap = xb;
  }
  const size_t nb = (size_t)n;
  if (nb > w.n_cap) {
    if (w.d_i) tcx_free(w.ctx, w.d_i);
    if (w.d_f) tcx_free(w.ctx, w.d_f);
    w.d_i = nullptr;
    w.d_f = nullptr;
    w.n_cap = 0;
    int rc = tcx_malloc(w.ctx, (void**)&w.d_i, 2 * nb * sizeof(int));
    if (rc == TCX_OK) rc = tcx_malloc(w.ctx, (void**)&w.d_f, nb * sizeof(float));
    if (rc != TCX_OK) return rc;
    w.n_cap = nb;
  }
  return TCX_OK;
}

int run_fit(tcx_server* s, Worker& w, tcx_request& r) {
  int rc = stage(w, r.n, r.m);
  if (rc == TCX_OK)
    rc = tcx_memcpy_h2d(w.ctx, w.d_x, r.h_x, (size_t)r.n * r.m * sizeof(float));
  if (rc == TCX_OK && r.znorm_ddof >= 0) rc = tcx_znorm(w.ctx, w.d_x, r.n, r.m, r.znorm_ddof);
  std::vector<float> cent;
  float* hc = r.h_centroids;
  if (rc == TCX_OK && r.publish_model && !hc) {
    cent.resize((size_t)r.k * r.m);
    hc = cent.data();
  }
  if (rc == TCX_OK) rc = tcx_kshape_fit(w.ctx, w.d_x, r.n, r.m, r.k, &r.opts, r.h_labels, hc,
                                        &r.stats);
  if (rc == TCX_OK && r.publish_model) rc = tcx_server_add_model(s, hc, r.k, r.m, &r.model_out);
  return rc;
}

int run_assign(tcx_server* s, Worker& w, tcx_request& r) {
  Model md;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (r.model < 0 || r.model >= (int)s->models.size())
      return tcx::fail(TCX_ERR_ARG, "assign query: unknown model " + std::to_string(r.model));
    md = s->models[r.model];
  }
  if (md.m != r.m)
    return tcx::fail(TCX_ERR_ARG, "assign query: series length " + std::to_string(r.m) +
                                      " differs from the model's " + std::to_string(md.m));
  int rc = stage(w, r.n, r.m);
  if (rc == TCX_OK)
    rc = tcx_memcpy_h2d(w.ctx, w.d_x, r.h_x, (size_t)r.n * r.m * sizeof(float));
  if (rc == TCX_OK && r.znorm_ddof >= 0) rc = tcx_znorm(w.ctx, w.d_x, r.n, r.m, r.znorm_ddof);
  int* d_lab = w.d_i;
  int* d_sh = w.d_i + w.n_cap;
  if (rc == TCX_OK)
    rc = tcx::assign_resident(w.ctx, w.d_x, r.n, r.m, md.k, md.d_spec, md.d_mag, md.d_norm,
                              TCX_CONV_CPU, d_lab, w.d_f, d_sh);
  if (rc == TCX_OK) rc = tcx_memcpy_d2h(w.ctx, r.h_labels, d_lab, (size_t)r.n * sizeof(int));
  if (rc == TCX_OK && r.h_dist)
    rc = tcx_memcpy_d2h(w.ctx, r.h_dist, w.d_f, (size_t)r.n * sizeof(float));
  if (rc == TCX_OK && r.h_shift)
    rc = tcx_memcpy_d2h(w.ctx, r.h_shift, d_sh, (size_t)r.n * sizeof(int));
  return rc;
}

// Next admissible request for this worker (called with the lock held).
bool take(tcx_server* s, Worker& w, Pending* out) {
  auto fits = [&](const Pending& p) { return !s->budget || s->used + p.bytes <= s->budget; };
  if (!w.query_only && !s->fitq.empty() && fits(s->fitq.front())) {
    *out = s->fitq.front();
    s->fitq.pop_front();
    return true;
  }
  if (!s->asgq.empty() && fits(s->asgq.front())) {
    *out = s->asgq.front();
    s->asgq.pop_front();
    return true;
  }
  return false;
}

void worker_main(tcx_server* s, Worker* w) {
  for (;;) {
    Pending p{nullptr, Clock::time_point(), 0};
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] {
        return (s->stop && s->fitq.empty() && s->asgq.empty()) || take(s, *w, &p);
      });
      if (!p.req) return;  // stopping with empty queues
      s->used += p.bytes;
      ++s->running;
    }
    tcx_request& r = *p.req;
    const Clock::time_point ts = Clock::now();
    const int rc = r.kind == TCX_REQ_FIT ? run_fit(s, *w, r) : run_assign(s, *w, r);
    const Clock::time_point te = Clock::now();
    r.status = rc;
    r.message[0] = '\0';
    if (rc != TCX_OK) {
      std::snprintf(r.message, sizeof(r.message), "%s", tcx_last_error());
      tcx::clear_error();
    }
    r.ms_queued = ms_between(p.t_submit, ts);
    r.ms_run = ms_between(ts, te);
    r.ms_latency = ms_between(p.t_submit, te);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->used -= p.bytes;
      --s->running;
      r.done = 1;
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();  // budget freed
  }
}

void free_model(tcx_ctx* c, Model& md) {
  if (md.d_cent) tcx_free(c, md.d_cent);
  if (md.d_spec) tcx_free(c, md.d_spec);
  if (md.d_mag) tcx_free(c, md.d_mag);
  if (md.d_norm) tcx_free(c, md.d_norm);
  if (md.d_zero) tcx_free(c, md.d_zero);
  md = Model();
}

}  // namespace

extern "C" {

int tcx_server_create(int device, int fit_workers, int query_workers, size_t hbm_budget,
                      tcx_server** out) {
  tcx::clear_error();
  if (!out) return tcx::fail(TCX_ERR_ARG, "server: null output pointer");
  if (fit_workers < 0 || query_workers < 0 || fit_workers + query_workers < 1 ||
      fit_workers + query_workers > 64)
    return tcx::fail(TCX_ERR_ARG, "server: 1..64 workers");
  tcx_server* s = new (std::nothrow) tcx_server();
  if (!s) return tcx::fail(TCX_ERR_NOMEM, "server: out of host memory");
  s->budget = hbm_budget;
  s->fit_workers = fit_workers;
  int rc = tcx_create(device, 0, &s->admin);
  for (int i = 0; rc == TCX_OK && i < fit_workers + query_workers; ++i) {
    Worker* w = new Worker();
    w->query_only = i >= fit_workers;
    rc = tcx_create(device, 0, &w->ctx);
    if (rc != TCX_OK) {
      delete w;
      break;
    }
    s->workers.push_back(w);
  }
  if (rc != TCX_OK) {
    for (Worker* w : s->workers) {
      tcx_destroy(w->ctx);
      delete w;
    }
    if (s->admin) tcx_destroy(s->admin);
    delete s;
    return rc;
  }
  for (Worker* w : s->workers) w->th = std::thread(worker_main, s, w);
  *out = s;
  return TCX_OK;
}

int tcx_server_add_model(tcx_server* s, const float* h_centroids, int k, int m, int* model_id) {
  tcx::clear_error();
  if (!s || !h_centroids || !model_id || k < 1 || m < 2)
    return tcx::fail(TCX_ERR_ARG, "add_model: bad arguments");
  const int F = tcx::fft_length(m);
  Model md;
  md.k = k;
  md.m = m;
  md.bytes = (size_t)k * m * 4 + (size_t)k * F * 12 + (size_t)k * 8;
  {
    std::lock_guard<std::mutex> lk(s->mu);  // reserve the model's bytes
    if (s->budget && s->used + md.bytes > s->budget)
      return tcx::fail(TCX_ERR_BUDGET, "add_model: the model exceeds the HBM budget");
    s->used += md.bytes;
  }
  std::lock_guard<std::mutex> alk(s->admin_mu);  // one upload at a time
  tcx_ctx* c = s->admin;
  int rc = tcx_malloc(c, (void**)&md.d_cent, (size_t)k * m * 4);
  if (rc == TCX_OK) rc = tcx_malloc(c, &md.d_spec, (size_t)k * F * 8);
  if (rc == TCX_OK) rc = tcx_malloc(c, (void**)&md.d_mag, (size_t)k * F * 4);
  if (rc == TCX_OK) rc = tcx_malloc(c, (void**)&md.d_norm, (size_t)k * 4);
  if (rc == TCX_OK) rc = tcx_malloc(c, (void**)&md.d_zero, (size_t)k * 4);
  if (rc == TCX_OK) rc = tcx_memcpy_h2d(c, md.d_cent, h_centroids, (size_t)k * m * 4);
  if (rc == TCX_OK)
    rc = tcx::model_spectra(c, md.d_cent, k, m, md.d_spec, md.d_mag, md.d_norm, md.d_zero);
  if (rc == TCX_OK) rc = tcx::stream_sync(c);
  std::lock_guard<std::mutex> lk(s->mu);
  if (rc != TCX_OK) {
    s->used -= md.bytes;
    free_model(c, md);
    return rc;
  }
  s->models.push_back(md);
  *model_id = (int)s->models.size() - 1;
  return TCX_OK;
}

int tcx_server_submit(tcx_server* s, tcx_request* r) {
  tcx::clear_error();
  if (!s || !r) return tcx::fail(TCX_ERR_ARG, "submit: null pointer");
  r->done = 0;
  r->status = -1;
  r->message[0] = '\0';
  r->model_out = -1;
  std::memset(&r->stats, 0, sizeof(r->stats));
  if (r->kind != TCX_REQ_FIT && r->kind != TCX_REQ_ASSIGN)
    return tcx::fail(TCX_ERR_ARG, "submit: unknown request kind");
  if (!r->h_x || !r->h_labels || r->n < 1 || r->m < 2)
    return tcx::fail(TCX_ERR_ARG, "submit: bad series or label buffer");
  Pending p{r, Clock::now(), request_bytes(*r)};
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return tcx::fail(TCX_ERR_ARG, "submit: server is stopping");
    if (r->kind == TCX_REQ_FIT && s->fit_workers == 0)
      return tcx::fail(TCX_ERR_ARG, "submit: the server has no fit workers");
    if (s->budget && p.bytes > s->budget)
      return tcx::fail(TCX_ERR_BUDGET, "submit: request exceeds the HBM budget");
    (r->kind == TCX_REQ_FIT ? s->fitq : s->asgq).push_back(p);
  }
  s->cv_work.notify_all();
  return TCX_OK;
}

int tcx_server_wait(tcx_server* s, tcx_request* r) {
  tcx::clear_error();
  if (!s || !r) return tcx::fail(TCX_ERR_ARG, "wait: null pointer");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return r->done != 0; });
  if (r->status != TCX_OK) return tcx::fail(r->status, r->message);
  return TCX_OK;
}

int tcx_server_destroy(tcx_server* s) {
  tcx::clear_error();
  if (!s) return TCX_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  int rc = TCX_OK;
  for (Worker* w : s->workers) {
    w->th.join();
    if (w->d_x) tcx_free(w->ctx, w->d_x);
    if (w->d_i) tcx_free(w->ctx, w->d_i);
    if (w->d_f) tcx_free(w->ctx, w->d_f);
    int r = tcx_destroy(w->ctx);
    if (rc == TCX_OK) rc = r;
    delete w;
  }
  for (Model& md : s->models) free_model(s->admin, md);
  int r = tcx_destroy(s->admin);
  if (rc == TCX_OK) rc = r;
  delete s;
  return rc;
}

}  // extern "C"

// Serving loop: fit requests and assignment queries against resident
// models, admitted under an HBM budget and run on worker contexts.
//
// Each worker owns one context (HIP stream + workspace) and grow-only
// device staging buffers for its requests' series and results.  Fit
// workers take the oldest fit (or, when no fit waits, the oldest query);
// query workers take only queries, so a burst of long fits cannot delay
// short queries.
//
// HBM budget.  A request needs its staging bytes plus, for a fit, the
// workspace bound of a fit of its size (tcx_fit_workspace_bytes), for a
// query the twiddle table.  Charged against the budget are the resident
// models, every running worker's reservation and every idle worker's held
// bytes (workspace slots and staging buffers stay allocated between
// requests).  A worker takes the oldest request of its queues when the
// charge with the request added fits the budget, releasing its own held
// bytes first when only that makes it fit; when idle workers' held bytes
// block it, those workers release them.  A request that cannot fit even
// with every worker idle and empty (need > budget - resident models) is
// refused at submission, or failed when a later model makes it so; since
// running requests end and idle workers release, every queued request is
// eventually admitted.  No reference counterpart.
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
  size_t bytes;  // what the request needs (staging + workspace bound)
};

struct Worker {
  tcx_ctx* ctx = nullptr;
  bool query_only = false;
  // budget state, under the server's lock
  bool busy = false;      // running a request (or releasing its workspace)
  bool trim = false;      // asked to release its workspace while idle
  size_t held = 0;        // bytes its context holds while idle
  size_t reserved = 0;    // bytes charged while busy
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

size_t staging_bytes(const tcx_request& r) {
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
  size_t budget = 0;
  size_t resident = 0;  // bytes of the resident models (and those being added)
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

void free_staging(Worker& w) {
  if (w.d_x) tcx_free(w.ctx, w.d_x);
  if (w.d_i) tcx_free(w.ctx, w.d_i);
  if (w.d_f) tcx_free(w.ctx, w.d_f);
  w.d_x = nullptr;
  w.d_i = nullptr;
  w.d_f = nullptr;
  w.x_cap = w.n_cap = 0;
}

// Bytes charged against the budget, leaving out worker `except` (lock held).
size_t charged(const tcx_server* s, const Worker* except) {
  size_t b = s->resident;
  for (const Worker* v : s->workers)
    if (v != except) b += v->busy ? v->reserved : v->held;
  return b;
}

// Asks the idle workers other than `except` that hold workspace to release
// it (lock held).
void request_trims(tcx_server* s, const Worker* except) {
  bool any = false;
  for (Worker* v : s->workers)
    if (v != except && !v->busy && !v->trim && v->held > 0) {
      v->trim = true;
      any = true;
    }
  if (any) s->cv_work.notify_all();
}

// Fails the queued requests that can no longer fit the budget beside the
// resident models (lock held).  Returns true if any was failed.
bool fail_unfit(tcx_server* s) {
  if (!s->budget) return false;
  bool any = false;
  for (auto* q : {&s->fitq, &s->asgq})
    for (auto it = q->begin(); it != q->end();) {
      if (it->bytes + s->resident > s->budget) {
        tcx_request& r = *it->req;
        r.status = TCX_ERR_BUDGET;
        std::snprintf(r.message, sizeof(r.message),
                      "request needs %zu bytes; the HBM budget leaves %zu beside the resident "
                      "models", it->bytes, s->budget - std::min(s->budget, s->resident));
        r.ms_queued = r.ms_latency = ms_between(it->t_submit, Clock::now());
        r.done = 1;
        it = q->erase(it);
        any = true;
      } else {
        ++it;
      }
    }
  return any;
}

int add_model(tcx_server* s, const float* h_centroids, int k, int m, int* model_id);

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
  if (rc == TCX_OK && r.publish_model) {
    {
      // the fit is over: charge what this worker's context holds, not its
      // reservation, while the model is added
      std::lock_guard<std::mutex> lk(s->mu);
      w.reserved = tcx::held_bytes(w.ctx);
    }
    rc = add_model(s, hc, r.k, r.m, &r.model_out);
  }
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

// Takes the next admissible request for this worker (lock held).  Sets
// *self_trim when the worker must release its own workspace first.
bool take(tcx_server* s, Worker& w, Pending* out, bool* self_trim) {
  auto admit = [&](const Pending& p) {
    if (!s->budget) {
      *self_trim = false;
      w.reserved = 0;
      return true;
    }
    const size_t others = charged(s, &w);
    if (others + w.held + p.bytes <= s->budget) {
      *self_trim = false;
      w.reserved = w.held + p.bytes;
      return true;
    }
    if (others + p.bytes <= s->budget) {
      *self_trim = true;
      w.reserved = p.bytes;
      return true;
    }
    request_trims(s, &w);  // idle workers' held bytes may be what blocks it
    return false;
  };
  if (!w.query_only && !s->fitq.empty() && admit(s->fitq.front())) {
    *out = s->fitq.front();
    s->fitq.pop_front();
    return true;
  }
  if (!s->asgq.empty() && admit(s->asgq.front())) {
    *out = s->asgq.front();
    s->asgq.pop_front();
    return true;
  }
  return false;
}

void worker_main(tcx_server* s, Worker* w) {
  for (;;) {
    Pending p{nullptr, Clock::time_point(), 0};
    bool self_trim = false, trim_only = false;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] {
        if (s->stop) return true;
        if (w->trim) return trim_only = true;
        return take(s, *w, &p, &self_trim);
      });
      if (!trim_only && !p.req) return;  // stopping (queued requests were failed)
      w->trim = false;
      w->busy = true;
      if (trim_only) w->reserved = w->held;
      else ++s->running;
    }
    if (trim_only || self_trim) {
      free_staging(*w);
      (void)tcx::trim_workspace(w->ctx);
      tcx::clear_error();
    }
    if (trim_only) {
      {
        std::lock_guard<std::mutex> lk(s->mu);
        w->held = tcx::held_bytes(w->ctx);
        w->busy = false;
      }
      s->cv_work.notify_all();  // bytes released
      continue;
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
      w->held = tcx::held_bytes(w->ctx);
      w->busy = false;
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

int add_model(tcx_server* s, const float* h_centroids, int k, int m, int* model_id) {
  if (!s || !h_centroids || !model_id || k < 1 || m < 2)
    return tcx::fail(TCX_ERR_ARG, "add_model: bad arguments");
  if (m > 4096) return tcx::fail(TCX_ERR_UNSUPPORTED, "add_model: series length > 4096");
  const int F = tcx::fft_length(m);
  Model md;
  md.k = k;
  md.m = m;
  // centroids, spectra, magnitudes, norms, zero flags, and the upload
  // context's twiddle table
  md.bytes = (size_t)k * m * 4 + (size_t)k * F * 12 + (size_t)k * 8 + (size_t)F * 8;
  {
    std::lock_guard<std::mutex> lk(s->mu);  // reserve the model's bytes
    if (s->budget && charged(s, nullptr) + md.bytes > s->budget) {
      request_trims(s, nullptr);
      return tcx::fail(TCX_ERR_BUDGET, "add_model: the model (" + std::to_string(md.bytes) +
                                           " bytes) does not fit the HBM budget left");
    }
    s->resident += md.bytes;
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
  bool failed_some = false;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (rc != TCX_OK) {
      s->resident -= md.bytes;
      free_model(c, md);
      return rc;
    }
    s->models.push_back(md);
    *model_id = (int)s->models.size() - 1;
    failed_some = fail_unfit(s);
  }
  if (failed_some) s->cv_done.notify_all();
  return TCX_OK;
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
  return add_model(s, h_centroids, k, m, model_id);
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
  if (r->m > 4096) return tcx::fail(TCX_ERR_UNSUPPORTED, "submit: series length > 4096");
  if (r->kind == TCX_REQ_FIT && (r->k < 1 || r->k > r->n))
    return tcx::fail(TCX_ERR_ARG, "submit: fit needs 1 <= k <= n");
  size_t need = staging_bytes(*r);
  if (r->kind == TCX_REQ_FIT)
    need += tcx_fit_workspace_bytes(s->admin, r->n, r->m, r->k);
  else
    need += (size_t)tcx::fft_length(r->m) * 8;  // the worker's twiddle table
  Pending p{r, Clock::now(), need};
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return tcx::fail(TCX_ERR_ARG, "submit: server is stopping");
    if (r->kind == TCX_REQ_FIT && s->fit_workers == 0)
      return tcx::fail(TCX_ERR_ARG, "submit: the server has no fit workers");
    if (s->budget && p.bytes + s->resident > s->budget)
      return tcx::fail(TCX_ERR_BUDGET,
                       "submit: request needs " + std::to_string(p.bytes) +
                           " bytes; the HBM budget leaves " +
                           std::to_string(s->budget - std::min(s->budget, s->resident)) +
                           " beside the resident models");
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
    // requests still queued fail; running ones finish
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    for (auto* q : {&s->fitq, &s->asgq}) {
      for (Pending& p : *q) {
        p.req->status = TCX_ERR_ARG;
        std::snprintf(p.req->message, sizeof(p.req->message),
                      "the server stopped before the request ran");
        p.req->done = 1;
      }
      q->clear();
    }
  }
  s->cv_work.notify_all();
  s->cv_done.notify_all();
  int rc = TCX_OK;
  for (Worker* w : s->workers) {
    w->th.join();
    free_staging(*w);
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

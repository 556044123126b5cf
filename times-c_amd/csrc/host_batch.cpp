// Batched fits: a fixed pool of worker threads, each with its own engine
// context (HIP stream + workspace) on one device.  Jobs are independent
// fits; a worker takes the next job from a shared queue, z-normalizes its
// data if asked and runs tcx_kshape_fit on its own stream, so the small
// kernels and host synchronisations of different jobs overlap on the GPU.
// No reference counterpart.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "tcx_host.hpp"
#include "tcx_internal.hpp"
#include "timesc.h"

using Clock = std::chrono::steady_clock;

struct tcx_batch {
  std::vector<tcx_ctx*> ctx;
  size_t budget = 0;
  std::vector<std::thread> threads;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<int> queue;
  tcx_job* jobs = nullptr;
  int pending = 0;
  bool stop = false;
  bool running = false;
  Clock::time_point t0;
};

namespace {

double ms_since(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

void run_job(tcx_ctx* ctx, tcx_job& j) {
  j.status = TCX_OK;
  if (!j.d_x || !j.h_labels) {
    j.status = tcx::fail(TCX_ERR_ARG, "batch job: null data or label pointer");
    return;
  }
  if (j.znorm_ddof >= 0) {
    j.status = tcx_znorm(ctx, j.d_x, j.n, j.m, j.znorm_ddof);
    if (j.status != TCX_OK) return;
  }
  j.status = tcx_kshape_fit(ctx, j.d_x, j.n, j.m, j.k, &j.opts, j.h_labels, j.h_centroids,
                            &j.stats);
}

void worker_main(tcx_batch* b, int w) {
  for (;;) {
    int idx;
    {
      std::unique_lock<std::mutex> lk(b->mu);
      b->cv_work.wait(lk, [&] { return b->stop || !b->queue.empty(); });
      if (b->stop && b->queue.empty()) return;
      idx = b->queue.front();
      b->queue.pop_front();
    }
    tcx_job& j = b->jobs[idx];
    const Clock::time_point ts = Clock::now();
    run_job(b->ctx[w], j);
    const Clock::time_point te = Clock::now();
    j.message[0] = '\0';
    if (j.status != TCX_OK) {
      std::snprintf(j.message, sizeof(j.message), "%s", tcx_last_error());
      tcx::clear_error();
    }
    j.worker = w;
    j.ms_queued = ms_since(b->t0, ts);
    j.ms_run = ms_since(ts, te);
    j.ms_latency = ms_since(b->t0, te);
    {
      std::lock_guard<std::mutex> lk(b->mu);
      if (--b->pending == 0) b->cv_done.notify_all();
    }
  }
}

// Runs jobs[idx] on the worker pool (the caller set b->running); returns
// when every one has finished and clears b->running.
void run_pool(tcx_batch* b, tcx_job* jobs, const std::vector<int>& idx, Clock::time_point t0) {
  {
    std::unique_lock<std::mutex> lk(b->mu);
    b->jobs = jobs;
    b->pending = (int)idx.size();
    b->t0 = t0;
    for (int i : idx) {
      jobs[i].status = -1;
      b->queue.push_back(i);
    }
  }
  b->cv_work.notify_all();
  std::unique_lock<std::mutex> lk(b->mu);
  b->cv_done.wait(lk, [&] { return b->pending == 0; });
  b->running = false;
  b->jobs = nullptr;
}

int first_failure(const tcx_job* jobs, int njobs) {
  for (int i = 0; i < njobs; ++i)
    if (jobs[i].status != TCX_OK)
      return tcx::fail(jobs[i].status,
                       "batch: job " + std::to_string(i) + " failed: " + jobs[i].message);
  return TCX_OK;
}

double job_cost(const tcx_job& j) {
  int F = 1;
  while (F < 2 * j.m) F <<= 1;  // >= the FFT length
  return (double)j.n * (double)(j.k + 1) * F * std::log2((double)F) + (double)j.n * j.m * j.m;
}

}  // namespace

extern "C" {

int tcx_batch_create(int device, int workers, size_t hbm_budget, tcx_batch** out) {
  tcx::clear_error();
  if (!out) return tcx::fail(TCX_ERR_ARG, "batch: null output pointer");
  if (workers < 1 || workers > 64) return tcx::fail(TCX_ERR_ARG, "batch: workers must be 1..64");
  tcx_batch* b = new (std::nothrow) tcx_batch();
  if (!b) return tcx::fail(TCX_ERR_NOMEM, "batch: out of host memory");
  for (int w = 0; w < workers; ++w) {
    tcx_ctx* c = nullptr;
    int rc = tcx_create(device, hbm_budget, &c);
    if (rc != TCX_OK) {
      for (tcx_ctx* x : b->ctx) tcx_destroy(x);
      delete b;
      return rc;
    }
    b->ctx.push_back(c);
  }
  b->budget = hbm_budget;
  for (int w = 0; w < workers; ++w) b->threads.emplace_back(worker_main, b, w);
  *out = b;
  return TCX_OK;
}

int tcx_batch_run(tcx_batch* b, tcx_job* jobs, int njobs, int order) {
  tcx::clear_error();
  if (!b) return tcx::fail(TCX_ERR_ARG, "batch: null batch");
  if (njobs < 0 || (njobs > 0 && !jobs)) return tcx::fail(TCX_ERR_ARG, "batch: bad job list");
  if (order < TCX_ORDER_GIVEN || order > TCX_ORDER_LONGEST_FIRST)
    return tcx::fail(TCX_ERR_ARG, "batch: unknown order");
  std::vector<int> idx(njobs);
  for (int i = 0; i < njobs; ++i) idx[i] = i;
  if (order != TCX_ORDER_GIVEN) {
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int c) {
      const double ca = job_cost(jobs[a]), cc = job_cost(jobs[c]);
      return order == TCX_ORDER_SHORTEST_FIRST ? ca < cc : ca > cc;
    });
  }
  {
    std::unique_lock<std::mutex> lk(b->mu);
    if (b->running) return tcx::fail(TCX_ERR_ARG, "batch: a run is already in progress");
    b->running = true;
  }
  run_pool(b, jobs, idx, Clock::now());
  return first_failure(jobs, njobs);
}

constexpr size_t kGroupedCohortMin = 32;  // fewer jobs: one cohort

int tcx_batch_run_grouped(tcx_batch* b, tcx_job* jobs, int njobs) {
  tcx::clear_error();
  if (!b) return tcx::fail(TCX_ERR_ARG, "batch: null batch");
  if (njobs < 0 || (njobs > 0 && !jobs)) return tcx::fail(TCX_ERR_ARG, "batch: bad job list");
  {
    std::unique_lock<std::mutex> lk(b->mu);
    if (b->running) return tcx::fail(TCX_ERR_ARG, "batch: a run is already in progress");
    b->running = true;
  }
  const Clock::time_point t0 = Clock::now();
  std::vector<int> grouped, rest;
  for (int i = 0; i < njobs; ++i) {
    const char* why = nullptr;
    (tcx::grouped_supported(jobs[i], &why) ? grouped : rest).push_back(i);
  }
  int rc = TCX_OK;
  if (!grouped.empty()) {
    // Cohorts (default two, TCX_GROUPED_COHORTS) on as many contexts (their
    // own streams), run from one host thread each: while one cohort is in a
    // latency-bound phase (the Lanczos steps, the host's planning) the
    // others' launches fill the GPU.  Jobs go to the cohort with the
    // smaller summed cost, largest first; a job's results do not depend on
    // its cohort.
    int W = 1;
    if (grouped.size() >= kGroupedCohortMin) {
      const char* e = std::getenv("TCX_GROUPED_COHORTS");  // default 2
      W = e ? std::max(1, std::atoi(e)) : 2;
      W = std::min<int>(W, (int)b->ctx.size());
    }
    std::vector<std::vector<int>> cohort(W);
    std::vector<double> load(W, 0.0);
    std::vector<int> order = grouped;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return job_cost(jobs[a]) > job_cost(jobs[c]); });
    for (int i : order) {
      const int w = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      cohort[w].push_back(i);
      load[w] += job_cost(jobs[i]);
    }
    size_t budget = b->budget;
    if (!budget) {  // most of the free memory and of what the cohorts keep for reuse
      size_t kept = 0;
      for (int w = 0; w < W; ++w) kept += tcx::ctx_buffer_device_bytes(b->ctx[w]);
      budget = (tcx::device_free_bytes(b->ctx[0]) + kept) / 10 * 8;
    }
    std::vector<int> rcs(W, TCX_OK);
    std::vector<std::string> msgs(W);
    auto work = [&](int w) {
      std::sort(cohort[w].begin(), cohort[w].end());  // the caller's order inside a cohort
      // the second stream: an idle worker context's, when there are enough,
      // so that the cohorts' streams are the first 2W streams created (the
      // runtime maps streams onto its hardware queues in creation order)
      void* second = b->ctx.size() >= (size_t)(2 * W) ? tcx::stream_of(b->ctx[W + w]) : nullptr;
      rcs[w] = tcx::grouped_fit(b->ctx[w], jobs, cohort[w].data(), (int)cohort[w].size(),
                                budget / W, t0, second);
      if (rcs[w] != TCX_OK) msgs[w] = tcx_last_error();  // thread-local
    };
    // a cohort whose thread cannot be started runs on this thread afterwards
    std::vector<std::thread> th;
    std::vector<int> inline_w;
    for (int w = 1; w < W; ++w) {
      try {
        th.emplace_back(work, w);
      } catch (const std::system_error&) {
        inline_w.push_back(w);
      }
    }
    work(0);
    for (int w : inline_w) work(w);
    for (std::thread& t : th) t.join();
    for (int w = 0; w < W; ++w) {
      if (rcs[w] == TCX_OK) continue;
      if (rc == TCX_OK) rc = rcs[w];
      for (int i : cohort[w])
        if (jobs[i].status == -1) {
          jobs[i].status = rcs[w];
          std::snprintf(jobs[i].message, sizeof(jobs[i].message), "%s", msgs[w].c_str());
        }
    }
  }
  // options the grouped executor does not take (centre initialisation,
  // eigenvector centroids, k beyond the pruned assignment): worker pool
  if (!rest.empty()) {
    std::stable_sort(rest.begin(), rest.end(),
                     [&](int a, int c) { return job_cost(jobs[a]) > job_cost(jobs[c]); });
    run_pool(b, jobs, rest, t0);
  } else {
    std::lock_guard<std::mutex> lk(b->mu);
    b->running = false;
  }
  if (rc != TCX_OK) return tcx::fail(rc, "batch: grouped run failed");
  return first_failure(jobs, njobs);
}

int tcx_batch_destroy(tcx_batch* b) {
  tcx::clear_error();
  if (!b) return TCX_OK;
  {
    std::lock_guard<std::mutex> lk(b->mu);
    b->stop = true;
  }
  b->cv_work.notify_all();
  for (std::thread& t : b->threads) t.join();
  int rc = TCX_OK;
  for (tcx_ctx* c : b->ctx) {
    int r = tcx_destroy(c);
    if (rc == TCX_OK) rc = r;
  }
  delete b;
  return rc;
}

}  // extern "C"

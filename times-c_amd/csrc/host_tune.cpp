// Autotuner of the launch shapes that do not change results.
//
// Two choices, each measured on the caller's data with whole fits (host
// clock around tcx_kshape_fit, which ends synchronized) and kept in a
// process-wide table that every context and the grouped executor consult:
//   - the pruned assignment's grid cap, per (F, ceil(k/16), floor(log2 n)):
//     the kernel strides over series pairs, each pair evaluated by one
//     group, so the grid changes only which block takes a pair;
//   - the align + pack rows per block (R = 16, 32, 64), per m rounded up to
//     32 (m <= 1024): each packed column is computed the same way whatever
//     block holds it.
// The tuner checks that every candidate gives the labels and centroids of
// the default, and fails otherwise.  Table file: one entry per line,
// "prune <F> <kb> <nb> <cap>" or "pack <m32> <rows>"; '#' starts a comment.
// No reference counterpart (the reference has fixed launch shapes).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "tcx_host.hpp"
#include "tcx_internal.hpp"
#include "timesc.h"

namespace tcx {
namespace {

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int>, int> g_prune_cap;  // (F, kb, nb) -> grid cap
std::map<int, int> g_pack_rows;                  // m rounded up to 32 -> rows

int kbucket(int k) { return (k + 15) / 16; }
int nbucket(long n) {
  int b = 0;
  while (n > 1) {
    n >>= 1;
    ++b;
  }
  return b;
}
int mbucket(int m) { return (m + 31) / 32 * 32; }

bool valid_cap(int cap) { return cap >= 64 && cap <= (1 << 20); }
bool valid_rows(int m, int rows) {
  return (rows == 16 || rows == 32 || rows == 64) && m <= kTunePackMaxM &&
         (size_t)rows * (m + 4) * 4 <= (size_t)kTunePackMaxBytes;
}

int fit_ms(tcx_ctx* ctx, const float* d_x, int n, int m, int k, int iters, int reps,
           std::vector<int>& labels, std::vector<float>& cent, double* best) {
  tcx_fit_opts o;
  tcx_fit_opts_default(&o);
  o.max_iter = iters;
  o.seed = 1;
  labels.resize(n);
  cent.resize((size_t)k * m);
  *best = 1e300;
  for (int r = 0; r < reps; ++r) {
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = tcx_kshape_fit(ctx, d_x, n, m, k, &o, labels.data(), cent.data(), nullptr);
    if (rc != TCX_OK) return rc;
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *best = std::min(*best, ms);
  }
  return TCX_OK;
}

}  // namespace

int tuned_prune_cap(int F, int k, long n) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_prune_cap.find({F, kbucket(k), nbucket(n)});
  return it == g_prune_cap.end() ? 0 : it->second;
}

int tuned_pack_rows(int m) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_pack_rows.find(mbucket(m));
  return it == g_pack_rows.end() ? 0 : it->second;
}

}  // namespace tcx

using namespace tcx;

extern "C" {

void tcx_tune_reset(void) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_prune_cap.clear();
  g_pack_rows.clear();
}

int tcx_tune_set(int kind, int key0, int key1, int key2, int value) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (kind == TCX_TUNE_PRUNE_CAP) {
    if (key0 < 4 || (key0 & (key0 - 1)) != 0 || key1 < 1 || key2 < 0 || key2 > 40 ||
        !valid_cap(value))
      return fail(TCX_ERR_ARG, "tune_set: bad pruned-assignment entry");
    g_prune_cap[{key0, key1, key2}] = value;
    return TCX_OK;
  }
  if (kind == TCX_TUNE_PACK_ROWS) {
    if (key0 < 32 || key0 % 32 != 0 || !valid_rows(key0, value))
      return fail(TCX_ERR_ARG, "tune_set: bad align + pack entry");
    g_pack_rows[key0] = value;
    return TCX_OK;
  }
  return fail(TCX_ERR_ARG, "tune_set: unknown kind");
}

int tcx_tune_get(int kind, int key0, int key1, int key2, int* value) {
  if (!value) return fail(TCX_ERR_ARG, "tune_get: null value");
  *value = kind == TCX_TUNE_PRUNE_CAP
               ? tuned_prune_cap(key0, key1 * 16, key2 >= 0 && key2 < 62 ? 1L << key2 : 0)
               : (kind == TCX_TUNE_PACK_ROWS ? tuned_pack_rows(key0) : -1);
  return *value < 0 ? fail(TCX_ERR_ARG, "tune_get: unknown kind") : TCX_OK;
}

int tcx_tune_save(const char* path) {
  if (!path) return fail(TCX_ERR_ARG, "tune_save: null path");
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(TCX_ERR_IO, std::string("tune_save: cannot open ") + path);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  std::fprintf(f, "# times-c_amd launch-shape table\n");
  for (auto& e : g_prune_cap)
    std::fprintf(f, "prune %d %d %d %d\n", std::get<0>(e.first), std::get<1>(e.first),
                 std::get<2>(e.first), e.second);
  for (auto& e : g_pack_rows) std::fprintf(f, "pack %d %d\n", e.first, e.second);
  const bool ok = std::fclose(f) == 0;
  return ok ? TCX_OK : fail(TCX_ERR_IO, std::string("tune_save: write failed: ") + path);
}

int tcx_tune_load(const char* path) {
  if (!path) return fail(TCX_ERR_ARG, "tune_load: null path");
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(TCX_ERR_IO, std::string("tune_load: cannot open ") + path);
  std::vector<std::vector<int>> rows;
  char line[256];
  int lineno = 0, rc = TCX_OK;
  while (std::fgets(line, sizeof line, f)) {
    ++lineno;
    char* h = std::strchr(line, '#');
    if (h) *h = '\0';
    char kind[16];
    int a = 0, b = 0, c = 0, e = 0;
    const int got = std::sscanf(line, "%15s %d %d %d %d", kind, &a, &b, &c, &e);
    if (got <= 0) continue;  // blank
    if (got == 5 && std::strcmp(kind, "prune") == 0) {
      rows.push_back({TCX_TUNE_PRUNE_CAP, a, b, c, e});
    } else if (got == 3 && std::strcmp(kind, "pack") == 0) {
      rows.push_back({TCX_TUNE_PACK_ROWS, a, 0, 0, b});
    } else {
      rc = fail(TCX_ERR_IO, "tune_load: bad line " + std::to_string(lineno) + " of " + path);
      break;
    }
  }
  std::fclose(f);
  if (rc != TCX_OK) return rc;
  // every entry checked before any is applied
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto& r : rows) {
      const bool ok = r[0] == TCX_TUNE_PRUNE_CAP
                          ? (r[1] >= 4 && (r[1] & (r[1] - 1)) == 0 && r[2] >= 1 && r[3] >= 0 &&
                             r[3] <= 40 && valid_cap(r[4]))
                          : (r[1] >= 32 && r[1] % 32 == 0 && valid_rows(r[1], r[4]));
      if (!ok) return fail(TCX_ERR_IO, std::string("tune_load: invalid entry in ") + path);
    }
  }
  for (auto& r : rows) tcx_tune_set(r[0], r[1], r[2], r[3], r[4]);
  return TCX_OK;
}

int tcx_autotune(tcx_ctx* ctx, const float* d_x, int n, int m, int k, int reps,
                 tcx_tune_report* report) {
  if (!ctx || !d_x || n < 2 || m < 2 || k < 1 || reps < 1)
    return fail(TCX_ERR_ARG, "autotune: bad arguments");
  const int F = fft_length(m);
  const int kb = (k + 15) / 16, nb = nbucket(n);
  const int iters = 3;
  std::vector<int> lab0, lab;
  std::vector<float> c0, c;
  // the shapes in force before tuning, as the reference for equality
  int prev_cap = tuned_prune_cap(F, k, n), prev_rows = tuned_pack_rows(m);
  auto restore = [&]() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (prev_cap) g_prune_cap[{F, kb, nb}] = prev_cap; else g_prune_cap.erase({F, kb, nb});
    if (prev_rows) g_pack_rows[mbucket(m)] = prev_rows; else g_pack_rows.erase(mbucket(m));
  };
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_prune_cap.erase({F, kb, nb});
    g_pack_rows.erase(mbucket(m));
  }
  double ms_default = 0;
  int rc = fit_ms(ctx, d_x, n, m, k, iters, reps, lab0, c0, &ms_default);
  if (rc != TCX_OK) {
    restore();
    return rc;
  }
  auto same = [&]() {
    return lab == lab0 && std::memcmp(c.data(), c0.data(), sizeof(float) * c0.size()) == 0;
  };
  // pruned-assignment grid cap
  const long npairs = ((long)n + (long)k * 63 + 1) / 2;
  int best_cap = 0;
  double best = ms_default;
  for (int cap : {1024, 2048, 4096, 8192, 16384, 32768}) {
    if (cap > 2 * npairs) break;
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_prune_cap[{F, kb, nb}] = cap;
    }
    double ms = 0;
    rc = fit_ms(ctx, d_x, n, m, k, iters, reps, lab, c, &ms);
    if (rc == TCX_OK && !same())
      rc = fail(TCX_ERR_NUMERIC, "autotune: a grid cap changed the fit's results");
    if (rc != TCX_OK) {
      restore();
      return rc;
    }
    if (ms < best) {
      best = ms;
      best_cap = cap;
    }
  }
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (best_cap) g_prune_cap[{F, kb, nb}] = best_cap; else g_prune_cap.erase({F, kb, nb});
  }
  // align + pack rows per block
  int best_rows = 0;
  for (int rows : {16, 32, 64}) {
    if (!valid_rows(mbucket(m), rows)) continue;
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_pack_rows[mbucket(m)] = rows;
    }
    double ms = 0;
    rc = fit_ms(ctx, d_x, n, m, k, iters, reps, lab, c, &ms);
    if (rc == TCX_OK && !same())
      rc = fail(TCX_ERR_NUMERIC, "autotune: a block shape changed the fit's results");
    if (rc != TCX_OK) {
      restore();
      return rc;
    }
    if (ms < best) {
      best = ms;
      best_rows = rows;
    }
  }
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (best_rows) g_pack_rows[mbucket(m)] = best_rows; else g_pack_rows.erase(mbucket(m));
  }
  if (report) {
    report->fft = F;
    report->prune_cap = best_cap;
    report->pack_rows = best_rows;
    report->ms_default = ms_default;
    report->ms_tuned = best;
  }
  return TCX_OK;
}

}  // extern "C"

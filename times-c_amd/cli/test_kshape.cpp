// Drop-in replacement of the reference CLI (Code/test/test_kshape.cu):
//
//   test -d <dev> <file> <k> <flag> <init_chose> [--key=value ...]
//
// flag 0: labeled CSV (label in column 0, k taken from the labels and the
// CLI k ignored, test_kshape.cu:24-30), 10 iterations (:21);
// flag 1: unlabeled CSV, k from the command line, 20 iterations (:90).
// init_chose 0: random initial labels (k_shape); 1: random initial centres
// (k_shape_tsl).  Like the reference's k_shape and k_shape_tsl, a fit stops
// right after a refinement that leaves one non-empty cluster or at most k/2
// (kshape.cuh:248-249, 454-455), keeping the labels of the last assignment.
// When the fit ends with one non-empty cluster or fewer than k/2, it is
// rerun as k_shape with eigenvector centroids (cal_select 1), as
// test_kshape.cu:54-59 does (exactly k/2 stops without a rerun).  Output:
// one-based labels "l,l,...,l," (readfile.cuh:120).
//
// Optional trailing flags (no reference counterpart):
//   --out=PATH          label file (default ../out/cukshape/result.txt)
//   --max-iter=N        iteration cap
//   --seed=S            initial labels / centres and power start vectors
//   --ddof=D            ddof of the input z-normalization (default 0, the
//                       CPU path's zscore; the reference GPU CLI used 1 --
//                       NCC is scale invariant, so labels do not depend on it)
//   --convention=cpu|kshape   alignment sign (default cpu)
//   --precision=split|bf16|auto  S = Y^T Y precision (default split; auto: BF16
//                             once its centroids are within 1e-4 of split)
//   --dump-iters=PATH   labels after every iteration, one line each
//   --centroids=PATH    final centroids, one CSV line each
//   --budget=BYTES      HBM budget of the context
//   --cpu-format        write zero-based labels one per line, the CPU
//                       path's output (test_timesC.py:324-325)
//   --no-collapse-stop  run every iteration like the CPU path (_kshape has
//                       no collapse stop, test_timesC.py:223-264)
// <file> may also be a binary matrix (tcx_write_binary, "TCXB0001" magic);
// flag 0 then needs the file's labels.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "timesc.h"

namespace {

void die(const char* what, int rc) {
  std::fprintf(stderr, "%s failed (status %d): %s\n", what, rc, tcx_last_error());
  std::exit(1);
}

#define CHECKRC(call)              \
  do {                             \
    int rc_ = (call);              \
    if (rc_ != TCX_OK) die(#call, rc_); \
  } while (0)

bool opt(const char* a, const char* key, std::string* val) {
  size_t n = std::strlen(key);
  if (std::strncmp(a, key, n) == 0 && a[n] == '=') {
    *val = a + n + 1;
    return true;
  }
  return false;
}

}  // namespace

int main(int argc, char** argv) {
  // test_kshape.cu:162-210: "-d" first, then device, file, k, flag, init_chose
  if (argc < 2 || std::strcmp(argv[1], "-d") != 0) return 0;
  if (argc < 7) {
    std::fprintf(stderr,
                 "usage: %s -d <dev> <file> <k> <flag> <init_chose> [--out=PATH] "
                 "[--max-iter=N] [--seed=S] [--ddof=D] [--convention=cpu|kshape] "
                 "[--precision=split|bf16|auto] [--dump-iters=PATH] [--centroids=PATH] "
                 "[--budget=BYTES] [--cpu-format]\n",
                 argv[0]);
    return 2;
  }
  int dev = std::atoi(argv[2]);
  std::string file = argv[3];
  int k = std::atoi(argv[4]);
  int flag = std::atoi(argv[5]);
  int init_chose = std::atoi(argv[6]);
  std::string out = "../out/cukshape/result.txt", dump, cent_path;
  int max_iter = -1, ddof = 0;
  uint64_t seed = 0;
  size_t budget = 0;
  bool cpu_format = false;
  tcx_fit_opts o;
  tcx_fit_opts_default(&o);
  o.stop_on_collapse = 1;
  for (int i = 7; i < argc; ++i) {
    std::string v;
    if (opt(argv[i], "--out", &v)) out = v;
    else if (opt(argv[i], "--max-iter", &v)) max_iter = std::atoi(v.c_str());
    else if (opt(argv[i], "--seed", &v)) seed = std::strtoull(v.c_str(), nullptr, 10);
    else if (opt(argv[i], "--ddof", &v)) ddof = std::atoi(v.c_str());
    else if (opt(argv[i], "--convention", &v)) o.convention = v == "kshape" ? TCX_CONV_KSHAPE : TCX_CONV_CPU;
    else if (opt(argv[i], "--precision", &v))
      o.s_precision = v == "bf16" ? TCX_S_BF16 : (v == "auto" ? TCX_S_AUTO : TCX_S_SPLIT_BF16);
    else if (opt(argv[i], "--dump-iters", &v)) dump = v;
    else if (opt(argv[i], "--centroids", &v)) cent_path = v;
    else if (opt(argv[i], "--budget", &v)) budget = std::strtoull(v.c_str(), nullptr, 10);
    else if (std::strcmp(argv[i], "--cpu-format") == 0) cpu_format = true;
    else if (std::strcmp(argv[i], "--no-collapse-stop") == 0) o.stop_on_collapse = 0;
    else {
      std::fprintf(stderr, "unknown option %s\n", argv[i]);
      return 2;
    }
  }
  std::printf("device_id = %i\n", dev);
  std::printf("-------------- %s --------------\n", file.c_str());

  float* h = nullptr;
  int* lab = nullptr;
  int rows = 0, cols = 0, iters_default;
  bool binary = false;
  if (FILE* f = std::fopen(file.c_str(), "rb")) {
    char magic[8] = {0};
    binary = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "TCXB0001", 8) == 0;
    std::fclose(f);
  }
  if (binary) {
    CHECKRC(tcx_read_binary(file.c_str(), &h, &lab, &rows, &cols));
    if (flag == 0 && !lab) {
      std::fprintf(stderr, "flag 0 needs labels; %s has none\n", file.c_str());
      return 2;
    }
  }
  if (flag == 0) {
    if (!binary) CHECKRC(tcx_read_labeled_csv(file.c_str(), &h, &lab, &rows, &cols));
    int mx = lab[0], mn = lab[0];
    for (int i = 1; i < rows; ++i) {
      mx = lab[i] > mx ? lab[i] : mx;
      mn = lab[i] < mn ? lab[i] : mn;
    }
    k = mx;  // test_kshape.cu:24-30
    if (mn == 0) k += 1;
    else if (mn == -1) k = 2;
    iters_default = 10;
  } else {
    if (!binary) CHECKRC(tcx_read_unlabeled_csv(file.c_str(), &h, &rows, &cols));
    iters_default = 20;
  }
  std::printf("mat_rows%d mat_cols%d\n", rows, cols);
  if (k < 1) {
    std::fprintf(stderr, "k must be >= 1 (got %d)\n", k);
    return 2;
  }
  o.max_iter = max_iter > 0 ? max_iter : iters_default;
  o.seed = seed;
  o.init_mode = init_chose == 1 ? 1 : 0;
  std::vector<int> hist;
  if (!dump.empty()) {
    hist.assign((size_t)o.max_iter * rows, 0);
    o.labels_per_iter = hist.data();
  }

  tcx_ctx* ctx = nullptr;
  CHECKRC(tcx_create(dev, budget, &ctx));
  float* d = nullptr;
  auto t0 = std::chrono::steady_clock::now();
  CHECKRC(tcx_malloc(ctx, (void**)&d, sizeof(float) * (size_t)rows * cols));
  CHECKRC(tcx_memcpy_h2d(ctx, d, h, sizeof(float) * (size_t)rows * cols));
  CHECKRC(tcx_znorm(ctx, d, rows, cols, ddof));
  auto t1 = std::chrono::steady_clock::now();
  std::printf("first norm run on GPU: %f msecs.\n",
              std::chrono::duration<double, std::milli>(t1 - t0).count());
  std::printf(o.init_mode ? "exec k_shape_tsl\n" : "exec k_shape\n");
  std::vector<int> labels(rows);
  std::vector<float> cent((size_t)k * cols);
  tcx_fit_stats st;
  CHECKRC(tcx_kshape_fit(ctx, d, rows, cols, k, &o, labels.data(), cent.data(), &st));
  auto t2 = std::chrono::steady_clock::now();
  std::printf("iterations %d converged %d collapsed %d sum_class %d power_steps_max %d\n",
              st.iterations, st.converged, st.collapsed, st.sum_class, st.power_steps_max);
  std::printf("refine %.3f ms assign %.3f ms\n", st.ms_refine, st.ms_assign);
  std::printf("total time run on GPU: %f msecs.\n",
              std::chrono::duration<double, std::milli>(t2 - t0).count());
  if (st.sum_class == 1 || st.sum_class < k / 2) {
    // test_kshape.cu:54-59, 112-117: too few non-empty clusters -> rerun
    // k_shape (random initial labels) with eigenvector centroids
    std::printf("sum_class %d: rerun k_shape with cal_select=1\n", st.sum_class);
    tcx_fit_opts o2 = o;
    o2.init_mode = 0;
    o2.cal_select = 1;
    CHECKRC(tcx_kshape_fit(ctx, d, rows, cols, k, &o2, labels.data(), cent.data(), &st));
    t2 = std::chrono::steady_clock::now();
    std::printf("iterations %d converged %d collapsed %d sum_class %d\n", st.iterations,
                st.converged, st.collapsed, st.sum_class);
    std::printf("total time run on GPU: %f msecs.\n",
                std::chrono::duration<double, std::milli>(t2 - t0).count());
  }
  if (cpu_format) {
    CHECKRC(tcx_write_labels_lines(out.c_str(), labels.data(), rows));  // test_timesC.py:324-325
  } else {
    for (int& l : labels) l += 1;  // one-based output, kshape.cuh:91
    CHECKRC(tcx_write_labels(out.c_str(), labels.data(), rows));
  }
  if (!dump.empty()) {
    FILE* f = std::fopen(dump.c_str(), "wb");
    if (!f) die("open --dump-iters file", TCX_ERR_IO);
    for (int it = 0; it < st.iterations; ++it) {
      for (int i = 0; i < rows; ++i)
        std::fprintf(f, i ? ",%d" : "%d", hist[(size_t)it * rows + i]);
      std::fprintf(f, "\n");
    }
    std::fclose(f);
  }
  if (!cent_path.empty()) {
    FILE* f = std::fopen(cent_path.c_str(), "wb");
    if (!f) die("open --centroids file", TCX_ERR_IO);
    for (int j = 0; j < k; ++j) {
      for (int t = 0; t < cols; ++t)
        std::fprintf(f, t ? ",%.9g" : "%.9g", cent[(size_t)j * cols + t]);
      std::fprintf(f, "\n");
    }
    std::fclose(f);
  }
  CHECKRC(tcx_free(ctx, d));
  tcx_destroy(ctx);
  tcx_free_host(h);
  tcx_free_host(lab);
  std::printf("\n");
  return 0;
}

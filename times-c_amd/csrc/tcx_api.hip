// C ABI of the engine: context, workspace, primitives and the k-Shape fit
// loop.  Host code; every kernel runs on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "tcx_host.hpp"
#include "tcx_internal.hpp"
#include "tcx_kernels.hpp"
#include "timesc.h"

namespace tcx {

namespace {
thread_local std::string g_err;
}

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
void clear_error() { g_err.clear(); }

}  // namespace tcx

using tcx::fail;

#define TCX_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(TCX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
  } while (0)

#define TCX_TRY(call)            \
  do {                           \
    int rc_ = (call);            \
    if (rc_ != TCX_OK) return rc_; \
  } while (0)

// Workspace slots (grow-only device buffers held by the context).
enum Slot {
  kSlotCspec, kSlotCnorm, kSlotCzero, kSlotLabel, kSlotShift, kSlotDist, kSlotCent,
  kSlotCentNew, kSlotBlkCount, kSlotBlkBase, kSlotGroupInfo,
  kSlotPerm, kSlotYhi, kSlotYlo, kSlotA0, kSlotPlan, kSlotPart,
  kSlotS, kSlotInfo, kSlotDummy, kSlotWbuf, kSlotRbuf, kSlotBar, kSlotVout, kSlotSteps,
  kSlotConv, kSlotErr, kSlotCmag, kSlotEvals, kSlotSvout, kSlotLzW, kSlotLzState, kSlotMask,
  kSlotRowStats, kSlotFitRowStats, kSlotKeyCounts, kSlotKeyOff, kSlotMinRow, kSlotXs, kSlotXn,
  kSlotLzXch, kSlotXm, kSlotUb, kSlotCm, kSlotCheck, kSlotCheckInfo, kNumSlots
};

struct FitState {
  bool active = false;
  const float* d_x = nullptr;
  int n = 0, m = 0, k = 0;
  tcx_fit_opts o{};
  double tol = 1e-12;
  int split = 1;
  int it = 0, converged = 0, sum_class = 0, steps_max = 0, fallbacks = 0;
  int collapsed = 0;  // stopped on collapse (opts.stop_on_collapse)
  // TCX_S_AUTO: iteration from which S runs in BF16 (-1: not yet), checks
  // run so far, and the accepted BF16 - split centroid difference
  int s_bf16_from = -1, s_checks = 0;
  double s_budget = 1e-4, s_check_err = 0.0;
  long n_pad = 0;
  bool have_order = false;   // the refine's cluster-sorted rows are valid
  const double2* row_stats = nullptr;  // per-row sums of d_x (fixed during the fit)
  long long evals = 0, pairs = 0;
  float ms_ref = 0.f, ms_asg = 0.f, ms_asg_last = 0.f, ms_bnd = 0.f;
  bool ref_timed = false;  // events 0..1 hold a refine not yet added to ms_ref
  bool have_dist = false;  // dist holds an assignment's distances (orders the next grouping)
  float2* xs = nullptr;    // half spectra of the series (per fit; NULL: transformed per call)
  float* xn = nullptr;     // their norms
  uint16_t* xm = nullptr;  // their bound rows, binary16 (with xs)
  float* ub = nullptr;     // bounds of the current assignment [n][ub_ld] (with xs)
  int ub_ld = 0;
  uint16_t* cm = nullptr;  // centroid bound rows, binary16, fragment order (with xs)
  int *label = nullptr, *shift = nullptr, *zero = nullptr;
  float *dist = nullptr, *cent = nullptr, *norm = nullptr;
  float2* spec = nullptr;
  std::chrono::steady_clock::time_point t_start;
};

struct tcx_ctx {
  FitState fit;
  int device = 0;
  size_t budget = 0;
  hipStream_t stream = nullptr;
  std::map<int, float2*> twiddles;
  void* slot_ptr[kNumSlots] = {};
  size_t slot_bytes[kNumSlots] = {};
  size_t ws_bytes = 0;
  size_t user_bytes = 0;
  std::map<void*, size_t> user_allocs;
  hipEvent_t ev[5] = {};  // refine (0, 1), assignment (2, 3), bounds end (4)
  hipEvent_t coop_ev = nullptr;  // after this context's last multi-block Lanczos launch
  void* lz_xch = nullptr;        // exchange buffer (and its size) the generation below refers to
  size_t lz_xch_bytes = 0;
  uint32_t lz_gen = 0;           // generation of the last multi-block Lanczos launch
  // the last S plan (host copies; the upload is staged in `pinned`)
  std::vector<tcx::SyrkGroup> sgroups;
  std::vector<tcx::SyrkCluster> sclusters;
  int cus = 0;                 // compute units of the device
  long long wall_khz = 0;      // wall_clock64() rate
  void* pinned = nullptr;      // page-locked host area for small readbacks
  size_t pinned_bytes = 0;
  // the grouped executor's buffers kept between calls, and its second stream
  void* cb_ptr[tcx::kCtxBuffers] = {};
  size_t cb_bytes[tcx::kCtxBuffers] = {};
  bool cb_pinned[tcx::kCtxBuffers] = {};
  hipStream_t aux_stream = nullptr;
};

namespace {

int fft_size(int m) {
  unsigned v = (unsigned)(2 * m - 1);
  int bl = 0;
  while (v) {
    ++bl;
    v >>= 1;
  }
  return 1 << bl;
}

int set_dev(tcx_ctx* c) {
  if (!c) return fail(TCX_ERR_ARG, "null context");
  TCX_HIP(hipSetDevice(c->device));
  return TCX_OK;
}

// Page-locked host memory of at least `bytes` for device-to-host readbacks
// of a few hundred bytes (pageable targets make such copies synchronous
// and staged).  The area is reused by every call on this context; the
// caller has synchronized the stream before it reads or reuses it.
int pinned_area(tcx_ctx* c, size_t bytes, void** out) {
  if (bytes > c->pinned_bytes) {
    if (c->pinned) TCX_HIP(hipHostFree(c->pinned));
    c->pinned = nullptr;
    c->pinned_bytes = 0;
    const size_t want = std::max<size_t>(bytes, 64 * 1024);
    TCX_HIP(hipHostMalloc(&c->pinned, want, hipHostMallocDefault));
    c->pinned_bytes = want;
  }
  *out = c->pinned;
  return TCX_OK;
}

int check_budget(tcx_ctx* c, size_t extra) {
  if (c->budget && c->ws_bytes + c->user_bytes + extra > c->budget)
    return fail(TCX_ERR_BUDGET, "HBM budget exceeded: need " +
                                    std::to_string(c->ws_bytes + c->user_bytes + extra) +
                                    " bytes, budget " + std::to_string(c->budget));
  return TCX_OK;
}

template <typename T>
int ws(tcx_ctx* c, Slot s, size_t count, T** out) {
  size_t bytes = sizeof(T) * (count ? count : 1);
  bytes = (bytes + 255) & ~(size_t)255;
  if (c->slot_bytes[s] < bytes) {
    if (c->slot_ptr[s]) {
      TCX_HIP(hipStreamSynchronize(c->stream));
      TCX_HIP(hipFree(c->slot_ptr[s]));
      c->ws_bytes -= c->slot_bytes[s];
      c->slot_ptr[s] = nullptr;
      c->slot_bytes[s] = 0;
    }
    TCX_TRY(check_budget(c, bytes));
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess)
      return fail(TCX_ERR_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    c->slot_ptr[s] = p;
    c->slot_bytes[s] = bytes;
    c->ws_bytes += bytes;
  }
  *out = reinterpret_cast<T*>(c->slot_ptr[s]);
  return TCX_OK;
}

int twiddles(tcx_ctx* c, int F, const float2** out) {
  auto it = c->twiddles.find(F);
  if (it != c->twiddles.end()) {
    *out = it->second;
    return TCX_OK;
  }
  std::vector<float2> h(F, make_float2(1.f, 0.f));
  tcx::fft_twiddle_table(F, h.data());
  TCX_TRY(check_budget(c, sizeof(float2) * F));
  float2* d = nullptr;
  if (hipMalloc(&d, sizeof(float2) * F) != hipSuccess)
    return fail(TCX_ERR_NOMEM, "twiddle table allocation failed");
  TCX_HIP(hipMemcpy(d, h.data(), sizeof(float2) * F, hipMemcpyHostToDevice));
  c->twiddles[F] = d;
  c->ws_bytes += sizeof(float2) * F;
  *out = d;
  return TCX_OK;
}

int check_series(int n, int m) {
  if (n < 1) return fail(TCX_ERR_ARG, "need at least one series");
  if (m < 2) return fail(TCX_ERR_ARG, "series length must be >= 2");
  if (m > tcx::kMaxFft / 2) return fail(TCX_ERR_UNSUPPORTED, "series length > 4096 is not supported");
  return TCX_OK;
}

// Spectra, spectrum magnitudes, norms and all-zero flags of k rows.
int spectra(tcx_ctx* c, const float* d_rows, int rows, int m, float2** spec, float** norm,
            int** zero, uint16_t* bmag = nullptr) {
  int F = fft_size(m);
  const float2* tw;
  float* mag;
  TCX_TRY(twiddles(c, F, &tw));
  TCX_TRY(ws(c, kSlotCspec, (size_t)rows * F, spec));
  TCX_TRY(ws(c, kSlotCmag, (size_t)rows * F, &mag));
  TCX_TRY(ws(c, kSlotCnorm, (size_t)rows, norm));
  TCX_TRY(ws(c, kSlotCzero, (size_t)rows, zero));
  TCX_HIP(tcx::launch_spectra(F, d_rows, rows, m, tw, *spec, *norm, *zero, mag, bmag,
                              c->stream));
  return TCX_OK;
}

// Assignment with the twiddle table tw of F = fft_size(m) and the
// centroid magnitudes cmag given.  Without a distance matrix the exact
// pruned kernel runs, pairing the rows of `order` (npairs pairs, -1 = hole)
// or rows 2p, 2p+1.
int assign_core_tw(tcx_ctx* c, const float2* tw, const float* d_x, long n, int m, int k,
                   const float2* cspec, const float* cnorm, const float* cmag, int rev, int* label,
                   float* dist, int* shift, float* dmat, int* changed,
                   const int* order = nullptr, long npairs = 0,
                   unsigned long long* evals = nullptr, const float2* xs = nullptr,
                   const float* xn = nullptr, const float* ub = nullptr, int ub_ld = 0) {
  const int F = fft_size(m);
  if (!dmat && k <= tcx::prune_max_k(F)) {
    tcx::PruneArgs p;
    p.x = d_x; p.n = n; p.m = m; p.k = k; p.cspec = cspec; p.cnorm = cnorm; p.tw = tw;
    p.cmag = cmag;
    p.rev = rev; p.label = label; p.dist = dist; p.shift = shift; p.changed = changed;
    p.order = order;
    p.npairs = order ? npairs : (n + 1) / 2;
    p.evals = evals;
    p.xs = xs;
    p.xn = xn;
    p.ub = ub;
    p.ub_ld = ub_ld;
    TCX_HIP(tcx::launch_assign_prune(F, p, c->stream));
    return TCX_OK;
  }
  tcx::SbdArgs a;
  a.x = d_x; a.n = n; a.m = m; a.k = k; a.cspec = cspec; a.cnorm = cnorm; a.tw = tw;
  a.rev = rev; a.label = label; a.dist = dist; a.shift = shift; a.dmat = dmat;
  a.changed = changed;
  TCX_HIP(tcx::launch_sbd(F, tcx::kModeAssign, a, c->stream));
  return TCX_OK;
}

// The same with the context's twiddle table and, without cmag, the
// magnitudes of the context's last spectra() call.
int assign_core(tcx_ctx* c, const float* d_x, long n, int m, int k, const float2* cspec,
                const float* cnorm, int rev, int* label, float* dist, int* shift, float* dmat,
                int* changed, const int* order = nullptr, long npairs = 0,
                unsigned long long* evals = nullptr, const float* cmag = nullptr,
                const float2* xs = nullptr, const float* xn = nullptr,
                const float* ub = nullptr, int ub_ld = 0) {
  const float2* tw;
  TCX_TRY(twiddles(c, fft_size(m), &tw));
  if (!cmag) cmag = reinterpret_cast<const float*>(c->slot_ptr[kSlotCmag]);
  return assign_core_tw(c, tw, d_x, n, m, k, cspec, cnorm, cmag, rev, label, dist, shift, dmat,
                        changed, order, npairs, evals, xs, xn, ub, ub_ld);
}

// ------------------------------------------------ allocation-free assignment
// Device scratch of tcx_assign_noalloc, one copy per GPU for the life of the
// process: the twiddle tables of every FFT length F = 4 .. 8192 (table F at
// entry F - 4, written once per device at set-up), then the centroid
// spectra [k][F], magnitudes [k][F], norms [k] and zero flags [k] of the
// current call, for k <= kNaMaxK and k * F <= kNaMaxKF (3.3 MB in all).
constexpr long kNaMaxKF = 1L << 18;
constexpr int kNaMaxK = 1024;
constexpr size_t kNaTwEntries = 16384;  // 4 + 8 + ... + 8192 = 16380, rounded up
constexpr size_t kNaSpecOff = sizeof(float2) * kNaTwEntries;
constexpr size_t kNaMagOff = kNaSpecOff + sizeof(float2) * kNaMaxKF;
constexpr size_t kNaNormOff = kNaMagOff + sizeof(float) * kNaMaxKF;
constexpr size_t kNaZeroOff = kNaNormOff + sizeof(float) * kNaMaxK;
constexpr size_t kNaBytes = kNaZeroOff + sizeof(int) * kNaMaxK;
constexpr int kNaMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNaBytes];

std::mutex g_na_mu;
unsigned char* g_na_base[kNaMaxDevices];  // device address of each GPU's scratch
float2 g_na_tw_host[kNaTwEntries];        // staging of the twiddle tables

// Address of the current device's scratch (the device is current).  The
// first call on a device looks the address up and writes the twiddle
// tables; no device memory is allocated.
int noalloc_scratch(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kNaMaxDevices) return fail(TCX_ERR_UNSUPPORTED, "device index >= 64");
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_base[dev]) {
    void* p = nullptr;
    TCX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    for (int F = 4; F <= tcx::kMaxFft; F *= 2) {
      float2* t = g_na_tw_host + (F - 4);
      for (int i = 0; i < F; ++i) t[i] = make_float2(1.f, 0.f);
      tcx::fft_twiddle_table(F, t);
    }
    TCX_HIP(hipMemcpy(p, g_na_tw_host, sizeof(g_na_tw_host), hipMemcpyHostToDevice));
    g_na_base[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_na_base[dev];
  return TCX_OK;
}

// Work list of S = Y^T Y for k_syrk: for each cluster with seg_len[j] > 0
// packed columns starting at off[j], the upper triangle of ceil(m/128)^2
// blocks of 128 x 128 (one wave each) over every K chunk of the cluster.
// Cluster j is split into nc_j chunks of nearly equal length (whole
// 16-column steps, at least 16 of them): the chunk bound L is the shortest
// for which every (cluster, chunk) group fits one round of the device (one
// wave per SIMD: the kernel's registers fill a SIMD) with each group's
// waves on one XCD, and nc_j = ceil(steps_j / L).
// The partials of block u of cluster j are consecutive over its chunks
// (summed in chunk order); returns their total count.
long plan_syrk(int m, int k, const long* off, const long* seg_len, int cus,
               std::vector<tcx::SyrkGroup>& groups, std::vector<tcx::SyrkCluster>& clusters) {
  constexpr long kXcds = 8, kMinSteps = 16;
  const long T = (m + tcx::kSyrkSub - 1) / tcx::kSyrkSub, U = T * (T + 1) / 2;  // waves per group
  groups.clear();
  clusters.clear();
  std::vector<long> steps(k);
  long max_steps = 0;
  for (int j = 0; j < k; ++j) {
    steps[j] = (seg_len[j] + 15) / 16;
    max_steps = std::max(max_steps, steps[j]);
  }
  if (max_steps == 0) return 0;
  const long slots = 4L * std::max(cus, 1);  // one wave per SIMD
  const long per_xcd = std::max(slots / kXcds, 1L);
  const long gmax = U <= per_xcd ? kXcds * (per_xcd / U) : std::max(slots / U, 1L);
  auto groups_for = [&](long L) {
    long g = 0;
    for (int j = 0; j < k; ++j) g += (steps[j] + L - 1) / L;
    return g;
  };
  long lo = std::min(kMinSteps, max_steps), hi = max_steps;
  while (lo < hi) {  // shortest chunk bound whose groups fit (else one chunk per cluster)
    const long mid = (lo + hi) / 2;
    if (groups_for(mid) <= gmax) hi = mid; else lo = mid + 1;
  }
  const long L = lo;
  long part = 0;
  for (int j = 0; j < k; ++j) {
    if (steps[j] == 0) continue;
    const long nc = (steps[j] + L - 1) / L;
    clusters.push_back(tcx::SyrkCluster{j, (int)part, (int)nc});
    for (long c = 0; c < nc; ++c) {
      const long s0 = steps[j] * c / nc, s1 = steps[j] * (c + 1) / nc;
      groups.push_back(tcx::SyrkGroup{(int)(part + c), (int)nc, off[j] + 16 * s0,
                                      off[j] + std::min(seg_len[j], 16 * s1)});
    }
    part += U * nc;
  }
  return part;
}

struct RefineResult {
  long n_pad = 0;
  int sum_class = 0;
  int steps_max = 0;
  int fallbacks = 0;
  double check_err = 0.0;  // with a BF16 check: max relative L2 centroid difference
};

// Multi-block Lanczos launches wait for each other: the members of one
// launch wait for each other on the device, so two such launches that were
// each partly resident could wait on each other for good.  Each launch is
// ordered after the previous one on the same device (any context, any
// stream) by a stream wait on that launch's event.
std::mutex g_coop_mu;
std::map<int, hipEvent_t> g_coop_last;

// M^I v0 of every non-empty cluster by Lanczos.  From m = kLzCoopMinM, when
// all k*G blocks fit on the device at once, one launch with G blocks per
// cluster; else for m <= 768 one fused launch (one block per cluster, S_j
// read from L2 every step); else per step one grid-wide matvec launch and
// one per-cluster update launch, and every 8 steps from 16 on the host
// reads the clusters' done flags and stops once all are done.
// Returns on device the vectors v, S v, the row sums r = S 1, the Lanczos
// step counts and the per-cluster acceptance flags.
int power_lanczos(tcx_ctx* c, const float* S, int m, int k, const int* counts, uint64_t seed,
                  int iter, int steps, int mode, const int* mask, const double** vout,
                  const double** svout, const double** rout, const int** steps_out,
                  const int** conv_out) {
  hipStream_t st = c->stream;
  double *Q, *v, *sv, *r, *w;
  int *st_out, *cv_out;
  tcx::LzState* state;
  TCX_TRY(ws(c, kSlotWbuf, (size_t)k * (tcx::kLanczosMax + 1) * m, &Q));
  TCX_TRY(ws(c, kSlotVout, (size_t)k * m, &v));
  TCX_TRY(ws(c, kSlotSvout, (size_t)k * m, &sv));
  TCX_TRY(ws(c, kSlotRbuf, (size_t)k * m, &r));
  TCX_TRY(ws(c, kSlotLzW, (size_t)k * m, &w));
  TCX_TRY(ws(c, kSlotSteps, (size_t)k, &st_out));
  TCX_TRY(ws(c, kSlotConv, (size_t)k, &cv_out));
  TCX_TRY(ws(c, kSlotLzState, (size_t)k, &state));
  tcx::LzCoopPlan cp;
  if (mode == 0 && !mask && tcx::plan_lz_coop(m, k, c->cus, &cp)) {
    // exchange words and per-cluster failure words; their contents only
    // ever match the current launch generation, so they are zeroed only
    // when (re)allocated or when the generation wraps
    uint64_t* xch;
    const size_t words = (size_t)k * cp.xch_len + ((size_t)k + 1) / 2;
    TCX_TRY(ws(c, kSlotLzXch, words, &xch));
    uint32_t* fails = reinterpret_cast<uint32_t*>(xch + (size_t)k * cp.xch_len);
    if (xch != c->lz_xch || c->slot_bytes[kSlotLzXch] != c->lz_xch_bytes ||
        c->lz_gen + 1 >= tcx::kLzCoopMaxGen) {
      TCX_HIP(tcx::launch_fill_int(reinterpret_cast<int*>(xch),
                                   (long)(c->slot_bytes[kSlotLzXch] / sizeof(int)), 0, st));
      c->lz_xch = xch;
      c->lz_xch_bytes = c->slot_bytes[kSlotLzXch];
      c->lz_gen = 0;
    }
    const uint32_t gen = ++c->lz_gen;
    if (!c->coop_ev) TCX_HIP(hipEventCreateWithFlags(&c->coop_ev, hipEventDisableTiming));
    {
      std::lock_guard<std::mutex> lk(g_coop_mu);
      auto prev = g_coop_last.find(c->device);
      if (prev != g_coop_last.end() && prev->second != c->coop_ev)
        TCX_HIP(hipStreamWaitEvent(st, prev->second, 0));
      // every wait of a member gives up after 2 s (TCX_LZ_COOP_TIMEOUT_US
      // sets another limit, a negative one giving up at the first wait, for
      // tests of that path: the cluster then takes the stepwise power
      // iteration)
      long long timeout = c->wall_khz * 1000LL * 2;
      if (const char* e = std::getenv("TCX_LZ_COOP_TIMEOUT_US")) {
        const long long us = std::atoll(e);
        timeout = us < 0 ? -1 : c->wall_khz * us / 1000;
      }
      TCX_HIP(tcx::launch_lz_coop(cp, S, m, k, counts, seed, iter, steps, tcx::kLzPowerTol, xch,
                                  fails, gen, timeout, v, r, st_out, cv_out, st));
      TCX_HIP(hipEventRecord(c->coop_ev, st));
      g_coop_last[c->device] = c->coop_ev;
    }
    TCX_HIP(tcx::launch_cluster_matvec(S, m, k, v, m, sv, nullptr, st));  // S v
    *vout = v;
    *svout = sv;
    *rout = r;
    *steps_out = st_out;
    *conv_out = cv_out;
    return TCX_OK;
  }
  if (mode == 0 && !mask && m <= tcx::kFusedLanczosMaxM) {
    // S_j fits in L2: one block per cluster runs every step in one launch
    TCX_HIP(tcx::launch_power_lanczos(S, m, k, counts, seed, iter, steps, tcx::kLzPowerTol, Q, v,
                                      r, st_out, cv_out, st));
    TCX_HIP(tcx::launch_cluster_matvec(S, m, k, v, m, sv, nullptr, st));  // S v
    *vout = v;
    *svout = sv;
    *rout = r;
    *steps_out = st_out;
    *conv_out = cv_out;
    return TCX_OK;
  }
  TCX_HIP(tcx::launch_cluster_matvec(S, m, k, nullptr, 0, r, nullptr, st));  // r = S 1
  TCX_HIP(tcx::launch_lz_setup(m, k, counts, mask, seed, iter, r, Q, state, st));
  const int jmax = m < tcx::kLanczosMax ? m : tcx::kLanczosMax;
  const long qstride = (long)(tcx::kLanczosMax + 1) * m;
  std::vector<int> done(k);
  for (int it = 0; it < jmax; ++it) {
    TCX_HIP(tcx::launch_cluster_matvec(S, m, k, Q + (long)it * m, qstride, w, state, st));
    TCX_HIP(tcx::launch_lz_step(m, k, mode, steps, mode == 0 ? tcx::kLzPowerTol : 1e-12, r, w, Q,
                                state,
                                st));
    const int len = it + 1;
    const bool check = mode == 0 ? tcx::lz_power_check(len) : (len >= 8 && (len & 7) == 0);
    if (len == jmax || check) {
      TCX_HIP(hipMemcpy2DAsync(done.data(), sizeof(int), &state[0].done, sizeof(tcx::LzState),
                               sizeof(int), k, hipMemcpyDeviceToHost, st));
      TCX_HIP(hipStreamSynchronize(st));
      bool all = true;
      for (int x : done) all = all && x;
      if (all) break;
    }
  }
  TCX_HIP(tcx::launch_lz_finish(m, k, counts, Q, state, v, st_out, cv_out, st));
  TCX_HIP(tcx::launch_cluster_matvec(S, m, k, v, m, sv, nullptr, st));  // S v
  *vout = v;
  *svout = sv;
  *rout = r;
  *steps_out = st_out;
  *conv_out = cv_out;
  return TCX_OK;
}

// One refinement: group, align + pack, S, power iteration.  Reads the
// labels, shifts and old-centroid zero flags; writes k*m new centroids.
int refine(tcx_ctx* c, const float* d_x, long n, int m, int k, const int* d_label,
           const int* d_shift, const int* d_czero, float* d_cout, int iter, uint64_t seed,
           int split, double tol, int cal_select, const double2* row_stats, RefineResult* res,
           const float* d_dist = nullptr, float* d_check = nullptr) {
  hipStream_t st = c->stream;
  if (!row_stats && tcx::align_pack_uses_row_stats(m)) {
    double2* rs;
    TCX_TRY(ws(c, kSlotRowStats, (size_t)n, &rs));
    TCX_HIP(tcx::launch_row_stats(d_x, n, m, rs, st));
    row_stats = rs;
  }
  c->fit.have_order = false;  // the row order is rewritten below
  // rows grouped by (label, bin of the previous distance): clusters stay
  // contiguous, and inside a cluster series of similar distance are
  // adjacent, so the next assignment pairs them (fewer evaluated centroids
  // per pair than pairs by row index)
  const int B = tcx::group_bins(n, k, d_dist != nullptr);
  const long K = (long)k * B;
  long nblk = (n + tcx::kGroupChunk - 1) / tcx::kGroupChunk;
  int *blk_count, *blk_base, *counts, *kcounts, *minrow, *perm, *bad;
  long *offsets, *koff;
  TCX_TRY(ws(c, kSlotBlkCount, (size_t)(nblk * K), &blk_count));
  TCX_TRY(ws(c, kSlotBlkBase, (size_t)(nblk * K), &blk_base));
  // cluster offsets (k + 1), counts (k) and the bad-label flag in one
  // buffer, in the layout of the pinned readback below (one copy)
  const size_t ginfo_bytes = sizeof(long) * (k + 1) + sizeof(int) * (k + 2);
  unsigned char* ginfo;
  TCX_TRY(ws(c, kSlotGroupInfo, ginfo_bytes, &ginfo));
  offsets = reinterpret_cast<long*>(ginfo);
  counts = reinterpret_cast<int*>(offsets + k + 1);
  bad = counts + k;
  TCX_TRY(ws(c, kSlotKeyCounts, (size_t)K, &kcounts));
  TCX_TRY(ws(c, kSlotKeyOff, (size_t)K, &koff));
  TCX_TRY(ws(c, kSlotMinRow, (size_t)k, &minrow));
  TCX_HIP(tcx::launch_fill_int(bad, 1, 0, st));
  TCX_HIP(tcx::launch_fill_int(minrow, k, 0x7fffffff, st));
  TCX_HIP(tcx::launch_group_count(d_label, n, k, d_dist, B, blk_count, minrow, bad, st));
  TCX_HIP(tcx::launch_group_scan((int)nblk, k, B, blk_count, blk_base, kcounts, counts, offsets,
                                 koff, st));
  // Placement and align + pack are launched before the host reads the
  // cluster sizes, on buffers sized for the largest padded count
  // (n + k * (kPackPad - 1)); their blocks past the padded end offsets[k]
  // exit at once.  The host then plans S = Y^T Y while they run.
  const long n_pad_max = std::max<long>(
      (n + (long)k * (tcx::kPackPad - 1)) / tcx::kPackPad * tcx::kPackPad, tcx::kPackPad);
  TCX_TRY(ws(c, kSlotPerm, (size_t)n_pad_max, &perm));
  TCX_HIP(tcx::launch_fill_int(perm, n_pad_max, -1, st));
  TCX_HIP(tcx::launch_group_place(d_label, n, k, d_dist, B, blk_base, koff, perm, st));
  uint16_t *yhi, *ylo;
  float* a0;
  const size_t panel = (size_t)tcx::panel_blocks(m) * 32 * (size_t)n_pad_max;
  TCX_TRY(ws(c, kSlotYhi, panel, &yhi));
  ylo = nullptr;  // a BF16 S needs the hi part only
  if (split) TCX_TRY(ws(c, kSlotYlo, panel, &ylo));
  TCX_TRY(ws(c, kSlotA0, (size_t)k * m, &a0));
  TCX_HIP(tcx::launch_align_pack(d_x, m, perm, n_pad_max, offsets, minrow, k, d_shift, d_czero,
                                 row_stats, yhi, ylo, a0, st));
  void* pin;
  TCX_TRY(pinned_area(c, ginfo_bytes, &pin));
  long* p_off = static_cast<long*>(pin);
  int* p_counts = reinterpret_cast<int*>(p_off + k + 1);
  int* p_bad = p_counts + k;
  TCX_HIP(hipMemcpyAsync(pin, ginfo, ginfo_bytes, hipMemcpyDeviceToHost, st));
  TCX_HIP(hipStreamSynchronize(st));
  // out-of-range labels were left out of the grouping above
  if (*p_bad) return fail(TCX_ERR_ARG, "label outside [0, k)");
  const std::vector<int> h_counts(p_counts, p_counts + k);
  const std::vector<long> h_off(p_off, p_off + k + 1);
  const long n_pad = h_off[k];
  int nonempty = 0;
  for (int j = 0; j < k; ++j) nonempty += h_counts[j] > 0;
  std::vector<long> seg_len(k);
  for (int j = 0; j < k; ++j) seg_len[j] = h_counts[j] > 0 ? h_off[j + 1] - h_off[j] : 0;
  const long nparts =
      plan_syrk(m, k, h_off.data(), seg_len.data(), c->cus, c->sgroups, c->sclusters);
  // groups and clusters in one buffer, uploaded by one copy
  const size_t groups_bytes = (sizeof(tcx::SyrkGroup) * c->sgroups.size() + 255) & ~(size_t)255;
  const size_t plan_bytes = groups_bytes + sizeof(tcx::SyrkCluster) * c->sclusters.size();
  unsigned char* d_plan;
  TCX_TRY(ws(c, kSlotPlan, plan_bytes, &d_plan));
  tcx::SyrkGroup* d_groups = reinterpret_cast<tcx::SyrkGroup*>(d_plan);
  tcx::SyrkCluster* d_clusters = reinterpret_cast<tcx::SyrkCluster*>(d_plan + groups_bytes);
  float *part, *S;
  tcx::PowerInfo* info;
  TCX_TRY(ws(c, kSlotPart, (size_t)std::max(nparts, 1L) * tcx::kSyrkSub * tcx::kSyrkSub, &part));
  TCX_TRY(ws(c, kSlotS, (size_t)k * m * m, &S));
  TCX_TRY(ws(c, kSlotInfo, (size_t)k, &info));
  const int ngroups = (int)c->sgroups.size(), nclusters = (int)c->sclusters.size();
  if (ngroups > 0) {
    // staged in the pinned area (idle here: its readback above was waited
    // for and consumed; its next use, the power stage's readback, is ordered
    // after this copy on the same stream and is sized in now, so the area is
    // not regrown while the copy reads it), so the copy does not block the
    // host
    void* stage;
    TCX_TRY(pinned_area(c, std::max(plan_bytes, sizeof(tcx::PowerInfo) * k), &stage));
    unsigned char* h_plan = static_cast<unsigned char*>(stage);
    std::memcpy(h_plan, c->sgroups.data(), sizeof(tcx::SyrkGroup) * ngroups);
    std::memcpy(h_plan + groups_bytes, c->sclusters.data(), sizeof(tcx::SyrkCluster) * nclusters);
    TCX_HIP(hipMemcpyAsync(d_plan, h_plan, plan_bytes, hipMemcpyHostToDevice, st));
  }
  TCX_HIP(tcx::launch_syrk(d_groups, ngroups, yhi, ylo, m, split, part, st));
  TCX_HIP(tcx::launch_syrk_reduce(d_clusters, nclusters, part, m, S, st));
  const int max_steps = m > 1000 ? 1000 : m;  // test_timesC.py:152-156
  const double* vin = nullptr;
  const double* svin = nullptr;
  const double* rsum = nullptr;
  const int* steps_in = nullptr;
  const int* conv_in = nullptr;
  // cal_select = 1: the top eigenvector of every M_j (eigh, shapeExtract.cuh:366-468 and
  // test_timesC.py:184-192); 0: the power iterate, with the clusters that fail the
  // power result's tests rerun through the eigenvector path.
  const int mode = cal_select == 1 ? 1 : 0;
  if (nonempty > 0)
    TCX_TRY(power_lanczos(c, S, m, k, counts, seed, iter, max_steps, mode, nullptr, &vin, &svin,
                          &rsum, &steps_in, &conv_in));
  TCX_HIP(tcx::launch_power(S, m, k, counts, a0, seed, iter, max_steps, tol, vin, svin, rsum,
                            steps_in, conv_in, mode, nullptr, 1, d_cout, info, st));
  TCX_TRY(pinned_area(c, sizeof(tcx::PowerInfo) * k, &pin));
  tcx::PowerInfo* p_info = static_cast<tcx::PowerInfo*>(pin);
  TCX_HIP(hipMemcpyAsync(p_info, info, sizeof(tcx::PowerInfo) * k, hipMemcpyDeviceToHost, st));
  TCX_HIP(hipStreamSynchronize(st));
  std::vector<tcx::PowerInfo> h_info(p_info, p_info + k);
  if (mode == 0) {
    std::vector<int> redo(k, 0);
    bool any = false;
    for (int j = 0; j < k; ++j)
      if (h_info[j].fallback && !h_info[j].empty && !h_info[j].zero) {
        redo[j] = 1;
        any = true;
      }
    if (any) {
      int* d_mask;
      TCX_TRY(ws(c, kSlotMask, (size_t)k, &d_mask));
      TCX_HIP(hipMemcpyAsync(d_mask, redo.data(), sizeof(int) * k, hipMemcpyHostToDevice, st));
      TCX_TRY(power_lanczos(c, S, m, k, counts, seed, iter, max_steps, 1, d_mask, &vin, &svin,
                            &rsum, &steps_in, &conv_in));
      TCX_HIP(tcx::launch_power(S, m, k, counts, a0, seed, iter, max_steps, tol, vin, svin, rsum,
                                steps_in, conv_in, 1, d_mask, 0, d_cout, info, st));
      TCX_HIP(hipMemcpyAsync(p_info, info, sizeof(tcx::PowerInfo) * k, hipMemcpyDeviceToHost,
                             st));
      TCX_HIP(hipStreamSynchronize(st));
      h_info.assign(p_info, p_info + k);
    }
  }
  if (d_check && nonempty > 0) {
    // the budget check of TCX_S_AUTO: the same S from one BF16 product (the
    // hi panel only) through the same power stage; the largest relative L2
    // difference between its centroids and the split ones above
    TCX_HIP(tcx::launch_syrk(d_groups, ngroups, yhi, ylo, m, 0, part, st));
    TCX_HIP(tcx::launch_syrk_reduce(d_clusters, nclusters, part, m, S, st));
    TCX_TRY(power_lanczos(c, S, m, k, counts, seed, iter, max_steps, 0, nullptr, &vin, &svin,
                          &rsum, &steps_in, &conv_in));
    tcx::PowerInfo* cinfo;
    TCX_TRY(ws(c, kSlotCheckInfo, (size_t)k, &cinfo));
    TCX_HIP(tcx::launch_power(S, m, k, counts, a0, seed, iter, max_steps, tol, vin, svin, rsum,
                              steps_in, conv_in, 0, nullptr, 1, d_check, cinfo, st));
    const size_t cb = sizeof(float) * (size_t)k * m;
    TCX_TRY(pinned_area(c, 2 * cb + sizeof(tcx::PowerInfo) * k, &pin));
    float* h_split = static_cast<float*>(pin);
    float* h_bf = h_split + (size_t)k * m;
    tcx::PowerInfo* h_cinfo = reinterpret_cast<tcx::PowerInfo*>(h_bf + (size_t)k * m);
    TCX_HIP(hipMemcpyAsync(h_split, d_cout, cb, hipMemcpyDeviceToHost, st));
    TCX_HIP(hipMemcpyAsync(h_bf, d_check, cb, hipMemcpyDeviceToHost, st));
    TCX_HIP(hipMemcpyAsync(h_cinfo, cinfo, sizeof(tcx::PowerInfo) * k, hipMemcpyDeviceToHost, st));
    TCX_HIP(hipStreamSynchronize(st));
    double err = 0.0;
    for (int j = 0; j < k; ++j) {
      if (h_info[j].empty || h_info[j].zero) continue;
      if (h_cinfo[j].fallback || h_cinfo[j].zero || h_info[j].fallback) {
        err = HUGE_VAL;  // the BF16 result would take another branch
        break;
      }
      double dd = 0.0, nn = 0.0;
      for (int i = 0; i < m; ++i) {
        const double a = h_bf[(size_t)j * m + i], b = h_split[(size_t)j * m + i];
        dd += (a - b) * (a - b);
        nn += b * b;
      }
      err = std::max(err, nn > 0.0 ? std::sqrt(dd / nn) : std::sqrt(dd));
    }
    if (res) res->check_err = err;
  }
  if (res) {
    res->n_pad = n_pad;
    res->sum_class = nonempty;
    for (int j = 0; j < k; ++j) {
      res->steps_max = std::max(res->steps_max, h_info[j].steps);
      res->fallbacks += h_info[j].fallback;
    }
  }
  return TCX_OK;
}

}  // namespace

// ============================================================== internal
namespace tcx {

int fft_length(int m) { return fft_size(m); }

int model_spectra(tcx_ctx* c, const float* d_cent, int k, int m, void* spec, float* mag,
                  float* norm, int* zero) {
  TCX_TRY(set_dev(c));
  TCX_TRY(check_series(k, m));
  const int F = fft_size(m);
  const float2* tw;
  TCX_TRY(twiddles(c, F, &tw));
  TCX_HIP(launch_spectra(F, d_cent, k, m, tw, static_cast<float2*>(spec), norm, zero, mag, nullptr,
                         c->stream));
  return TCX_OK;
}

int assign_resident(tcx_ctx* c, const float* d_x, long n, int m, int k, const void* spec,
                    const float* mag, const float* norm, int convention, int* label, float* dist,
                    int* shift) {
  TCX_TRY(set_dev(c));
  TCX_TRY(check_series(n, m));
  return assign_core(c, d_x, n, m, k, static_cast<const float2*>(spec), norm, convention, label,
                     dist, shift, nullptr, nullptr, nullptr, 0, nullptr, mag);
}

int stream_sync(tcx_ctx* c) {
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TCX_OK;
}

size_t held_bytes(const tcx_ctx* c) { return c->ws_bytes + c->user_bytes; }

namespace {
hipError_t release_ctx_buffer(tcx_ctx* c, int id) {
  hipError_t e = hipSuccess;
  if (c->cb_ptr[id]) {
    e = c->cb_pinned[id] ? hipHostFree(c->cb_ptr[id]) : hipFree(c->cb_ptr[id]);
    if (!c->cb_pinned[id]) c->ws_bytes -= c->cb_bytes[id];
  }
  c->cb_ptr[id] = nullptr;
  c->cb_bytes[id] = 0;
  return e;
}
}  // namespace

int ctx_buffer(tcx_ctx* c, int id, size_t bytes, bool pinned, void** out) {
  if (id < 0 || id >= kCtxBuffers) return fail(TCX_ERR_ARG, "context buffer: bad id");
  bytes = std::max<size_t>(bytes, 1);
  if (c->cb_ptr[id] && c->cb_bytes[id] >= bytes && c->cb_pinned[id] == pinned) {
    *out = c->cb_ptr[id];
    return TCX_OK;
  }
  TCX_HIP(release_ctx_buffer(c, id));
  void* p = nullptr;
  const hipError_t e = pinned ? hipHostMalloc(&p, bytes, hipHostMallocDefault) : hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(TCX_ERR_NOMEM, "context buffer allocation failed");
  }
  c->cb_ptr[id] = p;
  c->cb_bytes[id] = bytes;
  c->cb_pinned[id] = pinned;
  if (!pinned) c->ws_bytes += bytes;
  *out = p;
  return TCX_OK;
}

size_t ctx_buffer_device_bytes(const tcx_ctx* c) {
  size_t b = 0;
  for (int i = 0; i < kCtxBuffers; ++i)
    if (!c->cb_pinned[i]) b += c->cb_bytes[i];
  return b;
}

int ctx_aux_stream(tcx_ctx* c, void** out) {
  if (!c->aux_stream) TCX_HIP(hipStreamCreateWithFlags(&c->aux_stream, hipStreamNonBlocking));
  *out = c->aux_stream;
  return TCX_OK;
}

size_t device_free_bytes(tcx_ctx* c) {
  if (set_dev(c) != TCX_OK) return 0;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 0;
  return free_b;
}

void* stream_of(tcx_ctx* c) { return c->stream; }
int device_of(const tcx_ctx* c) { return c->device; }
int cus_of(const tcx_ctx* c) { return c->cus; }
int twiddle_table(tcx_ctx* c, int F, const void** tw) {
  TCX_TRY(set_dev(c));
  const float2* t;
  TCX_TRY(twiddles(c, F, &t));
  *tw = t;
  return TCX_OK;
}

int trim_workspace(tcx_ctx* c) {
  TCX_TRY(set_dev(c));
  if (c->fit.active) return fail(TCX_ERR_ARG, "trim: a fit is in progress");
  TCX_HIP(hipStreamSynchronize(c->stream));
  for (int s = 0; s < kNumSlots; ++s)
    if (c->slot_ptr[s]) {
      TCX_HIP(hipFree(c->slot_ptr[s]));
      c->ws_bytes -= c->slot_bytes[s];
      c->slot_ptr[s] = nullptr;
      c->slot_bytes[s] = 0;
    }
  for (auto& kv : c->twiddles) {
    TCX_HIP(hipFree(kv.second));
    c->ws_bytes -= sizeof(float2) * kv.first;
  }
  c->twiddles.clear();
  if (c->aux_stream) TCX_HIP(hipStreamSynchronize(c->aux_stream));
  for (int i = 0; i < kCtxBuffers; ++i) TCX_HIP(release_ctx_buffer(c, i));
  // a new exchange buffer must be zeroed before its first use, even at the
  // same address
  c->lz_xch = nullptr;
  c->lz_xch_bytes = 0;
  c->fit = FitState();
  return TCX_OK;
}

// Upper bound of the workspace bytes one fit of n series of length m into k
// clusters allocates (every slot that tcx_fit_begin, the refinements and the
// assignments grow, each rounded up to 256 bytes, plus the twiddle table).
size_t fit_workspace_bound(long n, int m, int k, int cus) {
  auto r = [](size_t b) { return (std::max<size_t>(b, 1) + 255) & ~(size_t)255; };
  const size_t F = (size_t)fft_size(m), N = (size_t)n, Kc = (size_t)k, Mm = (size_t)m;
  const size_t kp = (Kc + 15) & ~(size_t)15;
  size_t b = 0;
  // labels, shifts, distances; current and new centroids; centroid spectra
  b += 3 * r(4 * N) + 2 * r(4 * Kc * Mm) + r(8 * Kc * F) + r(4 * Kc * F) + 2 * r(4 * Kc);
  // series half spectra, norms, bound rows, bounds, centroid bound rows, row sums
  b += r(8 * N * (F / 2 + 1)) + r(4 * N) + r(2 * (size_t)bound_rows(n) * bound_kpad((int)F)) +
       r(4 * N * kp) + r(2 * (size_t)bound_rows(k) * bound_kpad((int)F)) + r(16 * N);
  // grouping (with distances: the most bins)
  const size_t B = (size_t)group_bins(n, k, true), keys = Kc * B;
  const size_t nblk = (N + kGroupChunk - 1) / kGroupChunk;
  b += 2 * r(4 * nblk * keys) + r(8 * (Kc + 1) + 4 * (Kc + 2)) + r(4 * keys) + r(8 * keys) +
       r(4 * Kc);
  // row order, Y panels (hi, lo), sign-fix rows
  const size_t npad = std::max<size_t>((N + Kc * (kPackPad - 1)) / kPackPad * kPackPad, kPackPad);
  b += r(4 * npad) + 2 * r(2 * (size_t)panel_blocks(m) * 32 * npad) + r(4 * Kc * Mm);
  // S plan and partials: plan_syrk keeps its groups (U waves and partials
  // each) within one round of the device or makes one group per cluster,
  // and a cluster's chunks are at least 256 columns but for one (so at most
  // npad / 256 + Kc groups)
  const size_t T = (Mm + kSyrkSub - 1) / kSyrkSub, U = T * (T + 1) / 2;
  const size_t groups =
      std::min(std::max((size_t)4 * std::max(cus, 1) / U + 8, Kc), npad / 256 + Kc);
  const size_t nparts = groups * U;
  b += r(sizeof(SyrkGroup) * groups + 256 + sizeof(SyrkCluster) * Kc);
  b += r(4 * nparts * kSyrkSub * kSyrkSub) + r(4 * Kc * Mm * Mm) + r(sizeof(PowerInfo) * Kc);
  // power stage: basis, v, S v, r, w, counters, states, exchange words, mask
  const size_t xch = Kc * (2 * (2 * (size_t)kLzCoopMaxG * (kLanczosMax + 2) + Mm) + kLzCoopMaxG);
  b += r(8 * Kc * (kLanczosMax + 1) * Mm) + 4 * r(8 * Kc * Mm) + 2 * r(4 * Kc) +
       r(sizeof(LzState) * Kc) + r(8 * xch + 4 * (Kc + 1)) + r(4 * Kc);
  // evaluation count and changed flag; twiddles
  b += r(8 * 17) + 8 * F;
  return b;
}

}  // namespace tcx

// ======================================================================= ABI
extern "C" {

const char* tcx_last_error(void) { return tcx::g_err.c_str(); }

int tcx_create(int device, size_t hbm_budget, tcx_ctx** out) {
  if (!out) return fail(TCX_ERR_ARG, "null output pointer");
  int ndev = 0;
  TCX_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return fail(TCX_ERR_ARG, "device " + std::to_string(device) + " not present (" +
                                 std::to_string(ndev) + " devices)");
  TCX_HIP(hipSetDevice(device));
  // Raise the dynamic-LDS limit of the large-F kernels; if the device
  // refuses, those launches report the error themselves.
  (void)tcx::configure_sbd_kernels();
  (void)tcx::configure_extract_kernels();
  (void)hipGetLastError();
  tcx_ctx* c = new tcx_ctx();
  c->device = device;
  int cus = 0, wall = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      hipDeviceGetAttribute(&wall, hipDeviceAttributeWallClockRate, device) != hipSuccess) {
    delete c;
    return fail(TCX_ERR_HIP, "device attribute query failed");
  }
  c->cus = cus;
  c->wall_khz = wall > 0 ? wall : 100000;
  c->budget = hbm_budget;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(TCX_ERR_HIP, "stream creation failed");
  }
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      tcx_destroy(c);
      return fail(TCX_ERR_HIP, "event creation failed");
    }
  *out = c;
  tcx::clear_error();
  return TCX_OK;
}

int tcx_destroy(tcx_ctx* c) {
  if (!c) return TCX_OK;
  // teardown: report the first failure but release everything
  hipError_t e = hipSetDevice(c->device);
  auto keep = [&e](hipError_t r) { if (e == hipSuccess) e = r; };
  if (c->stream) keep(hipStreamSynchronize(c->stream));
  for (int s = 0; s < kNumSlots; ++s)
    if (c->slot_ptr[s]) keep(hipFree(c->slot_ptr[s]));
  for (auto& kv : c->twiddles) keep(hipFree(kv.second));
  for (auto& kv : c->user_allocs) keep(hipFree(kv.first));
  for (auto& ev : c->ev)
    if (ev) keep(hipEventDestroy(ev));
  if (c->coop_ev) {
    {
      std::lock_guard<std::mutex> lk(g_coop_mu);
      auto it = g_coop_last.find(c->device);
      if (it != g_coop_last.end() && it->second == c->coop_ev) g_coop_last.erase(it);
    }
    keep(hipEventDestroy(c->coop_ev));
  }
  if (c->pinned) keep(hipHostFree(c->pinned));
  if (c->aux_stream) keep(hipStreamSynchronize(c->aux_stream));
  for (int i = 0; i < tcx::kCtxBuffers; ++i) keep(tcx::release_ctx_buffer(c, i));
  if (c->aux_stream) keep(hipStreamDestroy(c->aux_stream));
  if (c->stream) keep(hipStreamDestroy(c->stream));
  delete c;
  if (e != hipSuccess) return fail(TCX_ERR_HIP, std::string("destroy: ") + hipGetErrorString(e));
  return TCX_OK;
}

size_t tcx_workspace_bytes(const tcx_ctx* c) { return c ? c->ws_bytes : 0; }

size_t tcx_fit_workspace_bytes(const tcx_ctx* c, int n, int m, int k) {
  if (!c || n < 1 || m < 2 || k < 1) return 0;
  return tcx::fit_workspace_bound(n, m, k, c->cus);
}

int tcx_malloc(tcx_ctx* c, void** p, size_t bytes) {
  TCX_TRY(set_dev(c));
  if (!p) return fail(TCX_ERR_ARG, "null pointer");
  TCX_TRY(check_budget(c, bytes));
  if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess)
    return fail(TCX_ERR_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
  c->user_allocs[*p] = bytes;
  c->user_bytes += bytes;
  return TCX_OK;
}

int tcx_free(tcx_ctx* c, void* p) {
  TCX_TRY(set_dev(c));
  auto it = c->user_allocs.find(p);
  if (it == c->user_allocs.end()) return fail(TCX_ERR_ARG, "pointer not allocated by tcx_malloc");
  TCX_HIP(hipStreamSynchronize(c->stream));
  TCX_HIP(hipFree(p));
  c->user_bytes -= it->second;
  c->user_allocs.erase(it);
  return TCX_OK;
}

int tcx_memcpy_h2d(tcx_ctx* c, void* d, const void* h, size_t bytes) {
  TCX_TRY(set_dev(c));
  TCX_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TCX_OK;
}

int tcx_memcpy_d2h(tcx_ctx* c, void* h, const void* d, size_t bytes) {
  TCX_TRY(set_dev(c));
  TCX_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TCX_OK;
}

int tcx_memset(tcx_ctx* c, void* d, int value, size_t bytes) {
  TCX_TRY(set_dev(c));
  TCX_HIP(hipMemsetAsync(d, value, bytes, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TCX_OK;
}

int tcx_synchronize(tcx_ctx* c) {
  TCX_TRY(set_dev(c));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TCX_OK;
}

int tcx_znorm(tcx_ctx* c, float* d_x, int rows, int cols, int ddof) {
  TCX_TRY(set_dev(c));
  if (!d_x || rows < 0 || cols < 1 || ddof < 0 || ddof >= cols + 1)
    return fail(TCX_ERR_ARG, "znorm: bad arguments");
  TCX_HIP(tcx::launch_znorm(d_x, rows, cols, ddof, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TCX_OK;
}

int tcx_ncc(tcx_ctx* c, const float* d_x, const float* d_y, float* d_out, int x_rows,
            int y_rows, int cols) {
  TCX_TRY(set_dev(c));
  TCX_TRY(check_series(x_rows, cols));
  if (!d_x || !d_y || !d_out || y_rows < 1) return fail(TCX_ERR_ARG, "ncc: bad arguments");
  float2* spec;
  float* norm;
  int* zero;
  TCX_TRY(spectra(c, d_y, y_rows, cols, &spec, &norm, &zero));
  int F = fft_size(cols);
  const float2* tw;
  TCX_TRY(twiddles(c, F, &tw));
  tcx::SbdArgs a;
  a.x = d_x; a.n = x_rows; a.m = cols; a.k = y_rows; a.cspec = spec; a.cnorm = norm;
  a.tw = tw; a.ncc_out = d_out;
  TCX_HIP(tcx::launch_sbd(F, tcx::kModeNcc, a, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TCX_OK;
}

int tcx_sbd_align(tcx_ctx* c, float* d_members, const float* d_center, int rows, int cols,
                  int convention) {
  TCX_TRY(set_dev(c));
  TCX_TRY(check_series(rows, cols));
  if (!d_members || !d_center || (convention != TCX_CONV_CPU && convention != TCX_CONV_KSHAPE))
    return fail(TCX_ERR_ARG, "sbd_align: bad arguments");
  float2* spec;
  float* norm;
  int* zero;
  TCX_TRY(spectra(c, d_center, 1, cols, &spec, &norm, &zero));
  int *lab, *sh;
  TCX_TRY(ws(c, kSlotDummy, (size_t)rows, &lab));
  TCX_TRY(ws(c, kSlotShift, (size_t)rows, &sh));
  TCX_TRY(assign_core(c, d_members, rows, cols, 1, spec, norm, convention, lab, nullptr, sh,
                      nullptr, nullptr));
  TCX_HIP(tcx::launch_apply_shift(d_members, rows, cols, sh, zero, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TCX_OK;
}

int tcx_assign(tcx_ctx* c, const float* d_x, const float* d_centroids, int n, int m, int k,
               int* d_label, float* d_dist, int* d_shift, float* d_dmat) {
  TCX_TRY(set_dev(c));
  TCX_TRY(check_series(n, m));
  if (!d_x || !d_centroids || !d_label || k < 1) return fail(TCX_ERR_ARG, "assign: bad arguments");
  float2* spec;
  float* norm;
  int* zero;
  TCX_TRY(spectra(c, d_centroids, k, m, &spec, &norm, &zero));
  TCX_TRY(assign_core(c, d_x, n, m, k, spec, norm, 0, d_label, d_dist, d_shift, d_dmat, nullptr));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TCX_OK;
}

int tcx_assign_noalloc(tcx_ctx* c, const float* d_x, const float* d_centroids, int n, int m,
                       int k, int* d_label, float* d_dist, int* d_shift, float* d_dmat) {
  TCX_TRY(set_dev(c));
  TCX_TRY(check_series(n, m));
  if (!d_x || !d_centroids || !d_label || k < 1)
    return fail(TCX_ERR_ARG, "assign_noalloc: bad arguments");
  const int F = fft_size(m);
  if (k > kNaMaxK || (long)k * F > kNaMaxKF)
    return fail(TCX_ERR_ARG, "assign_noalloc: k = " + std::to_string(k) + " and F = " +
                                 std::to_string(F) + " exceed the static scratch (k <= " +
                                 std::to_string(kNaMaxK) + ", k * F <= " +
                                 std::to_string(kNaMaxKF) + ")");
  unsigned char* base;
  TCX_TRY(noalloc_scratch(c->device, &base));
  const float2* tw = reinterpret_cast<const float2*>(base) + (F - 4);
  float2* spec = reinterpret_cast<float2*>(base + kNaSpecOff);
  float* mag = reinterpret_cast<float*>(base + kNaMagOff);
  float* norm = reinterpret_cast<float*>(base + kNaNormOff);
  int* zero = reinterpret_cast<int*>(base + kNaZeroOff);
  TCX_HIP(tcx::launch_spectra(F, d_centroids, k, m, tw, spec, norm, zero, mag, nullptr,
                              c->stream));
  TCX_TRY(assign_core_tw(c, tw, d_x, n, m, k, spec, norm, mag, 0, d_label, d_dist, d_shift,
                         d_dmat, nullptr));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TCX_OK;
}

int tcx_extract_shape(tcx_ctx* c, const int* d_label, const float* d_x, const int* d_shift,
                      float* d_centroids, int n, int m, int k, int iter, int cal_select,
                      int* sum_class) {
  TCX_TRY(set_dev(c));
  TCX_TRY(check_series(n, m));
  if (!d_label || !d_x || !d_centroids || k < 1)
    return fail(TCX_ERR_ARG, "extract_shape: bad arguments");
  if (cal_select != 0 && cal_select != 1)
    return fail(TCX_ERR_ARG, "extract_shape: cal_select must be 0 or 1");
  float2* spec;
  float* norm;
  int* zero;
  TCX_TRY(spectra(c, d_centroids, k, m, &spec, &norm, &zero));  // old-centroid zero flags
  if (!d_shift) {
    // shift of each series against its own cluster's old centroid
    int *lab, *sh;
    TCX_TRY(ws(c, kSlotDummy, (size_t)n, &lab));
    TCX_TRY(ws(c, kSlotShift, (size_t)n, &sh));
    int F = fft_size(m);
    const float2* tw;
    TCX_TRY(twiddles(c, F, &tw));
    tcx::SbdArgs a;
    a.x = d_x; a.n = n; a.m = m; a.k = k; a.cspec = spec; a.cnorm = norm; a.tw = tw;
    a.label = lab; a.shift = sh; a.fixed_label = d_label;
    TCX_HIP(tcx::launch_sbd(F, tcx::kModeAssign, a, c->stream));
    d_shift = sh;
  }
  float* cnew;
  TCX_TRY(ws(c, kSlotCentNew, (size_t)k * m, &cnew));
  RefineResult rr;
  TCX_TRY(refine(c, d_x, n, m, k, d_label, d_shift, zero, cnew, iter, 0, 1, 1e-12, cal_select,
                 nullptr, &rr));
  TCX_HIP(hipMemcpyAsync(d_centroids, cnew, sizeof(float) * k * m, hipMemcpyDeviceToDevice,
                         c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  if (sum_class) *sum_class = rr.sum_class;
  return TCX_OK;
}

void tcx_fit_opts_default(tcx_fit_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->max_iter = 10;
  o->init_mode = 0;
  o->seed = 0;
  o->convention = TCX_CONV_CPU;
  o->s_precision = TCX_S_SPLIT_BF16;
  o->cal_select = 0;
  o->power_tol = 1e-12;
}

int tcx_fit_begin(tcx_ctx* c, const float* d_x, int n, int m, int k, const tcx_fit_opts* opts) {
  TCX_TRY(set_dev(c));
  TCX_TRY(check_series(n, m));
  FitState& f = c->fit;
  f = FitState();
  f.t_start = std::chrono::steady_clock::now();
  tcx_fit_opts o;
  if (opts) o = *opts; else tcx_fit_opts_default(&o);
  if (!d_x) return fail(TCX_ERR_ARG, "fit: null data pointer");
  if (k < 1) return fail(TCX_ERR_ARG, "fit: k must be >= 1");
  if (o.max_iter < 1) return fail(TCX_ERR_ARG, "fit: max_iter must be >= 1");
  if (o.init_mode != 0 && o.init_mode != 1) return fail(TCX_ERR_ARG, "fit: init_mode must be 0 or 1");
  if (o.init_mode == 1 && k > n) return fail(TCX_ERR_ARG, "fit: k > n with centre initialisation");
  if (o.convention != TCX_CONV_CPU && o.convention != TCX_CONV_KSHAPE)
    return fail(TCX_ERR_ARG, "fit: unknown convention");
  if (o.s_precision != TCX_S_SPLIT_BF16 && o.s_precision != TCX_S_BF16 &&
      o.s_precision != TCX_S_AUTO)
    return fail(TCX_ERR_ARG, "fit: unknown s_precision");
  if (!(o.s_budget >= 0.0)) return fail(TCX_ERR_ARG, "fit: s_budget must be >= 0");
  if (o.cal_select != 0 && o.cal_select != 1)
    return fail(TCX_ERR_ARG, "fit: cal_select must be 0 (power iteration) or 1 (eigenvector)");
  hipStream_t st = c->stream;
  TCX_TRY(ws(c, kSlotLabel, (size_t)n, &f.label));
  TCX_TRY(ws(c, kSlotShift, (size_t)n, &f.shift));
  TCX_TRY(ws(c, kSlotDist, (size_t)n, &f.dist));
  f.have_dist = false;
  // The series do not change during the fit: transform them once for the
  // pruned assignment (n * (F/2 + 1) complex values) and keep their bound
  // rows, so an assignment forms all n * k bounds as one product, when the
  // budget allows.
  f.xs = nullptr;
  f.xn = nullptr;
  f.xm = nullptr;
  f.ub = nullptr;
  f.cm = nullptr;
  {
    const int F = fft_size(m);
    const int kp = (k + 15) & ~15;
    const size_t need = sizeof(float2) * (size_t)n * (F / 2 + 1) + sizeof(float) * (size_t)n +
                        sizeof(uint16_t) * (size_t)tcx::bound_rows(n) * tcx::bound_kpad(F) +
                        sizeof(float) * (size_t)n * kp;
    const bool fits = !c->budget || c->ws_bytes + c->user_bytes + need <= c->budget;
    if (k <= tcx::prune_max_k(F) && fits) {
      const float2* tw;
      TCX_TRY(twiddles(c, F, &tw));
      TCX_TRY(ws(c, kSlotXs, (size_t)n * (F / 2 + 1), &f.xs));
      TCX_TRY(ws(c, kSlotXn, (size_t)n, &f.xn));
      TCX_TRY(ws(c, kSlotXm, (size_t)tcx::bound_rows(n) * tcx::bound_kpad(F), &f.xm));
      TCX_TRY(ws(c, kSlotUb, (size_t)n * kp, &f.ub));
      TCX_TRY(ws(c, kSlotCm, (size_t)tcx::bound_rows(k) * tcx::bound_kpad(F), &f.cm));
      f.ub_ld = kp;
      TCX_HIP(tcx::launch_series_spec(F, d_x, n, m, tw, f.xs, f.xn, f.xm, st));
    }
  }
  TCX_TRY(ws(c, kSlotCent, (size_t)k * m, &f.cent));
  std::vector<int> init;
  if (o.init_mode == 0) {
    init.resize(n);
    if (o.init_labels) {
      for (int i = 0; i < n; ++i) {
        if (o.init_labels[i] < 0 || o.init_labels[i] >= k)
          return fail(TCX_ERR_ARG, "fit: initial label outside [0, k)");
        init[i] = o.init_labels[i];
      }
    } else {
      tcx_init_labels(o.seed, n, k, init.data());
    }
    TCX_HIP(hipMemcpyAsync(f.label, init.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    TCX_HIP(hipMemsetAsync(f.cent, 0, sizeof(float) * k * m, st));
    TCX_HIP(hipMemsetAsync(f.shift, 0, sizeof(int) * n, st));
  } else {
    init.resize(k);
    if (o.init_rows) {
      for (int j = 0; j < k; ++j) {
        if (o.init_rows[j] < 0 || o.init_rows[j] >= n)
          return fail(TCX_ERR_ARG, "fit: initial centre row outside [0, n)");
        init[j] = o.init_rows[j];
      }
    } else {
      tcx_init_rows(o.seed, n, k, init.data());
    }
    for (int j = 0; j < k; ++j)
      TCX_HIP(hipMemcpyAsync(f.cent + (size_t)j * m, d_x + (size_t)init[j] * m,
                             sizeof(float) * m, hipMemcpyDeviceToDevice, st));
    TCX_HIP(hipMemsetAsync(f.label, 0, sizeof(int) * n, st));
  }
  TCX_TRY(spectra(c, f.cent, k, m, &f.spec, &f.norm, &f.zero, f.cm));
  if (o.init_mode == 0) TCX_HIP(tcx::launch_fill_int(f.zero, k, 1, st));  // centroids start at 0
  f.row_stats = nullptr;
  if (tcx::align_pack_uses_row_stats(m)) {
    double2* rs;
    TCX_TRY(ws(c, kSlotFitRowStats, (size_t)n, &rs));
    TCX_HIP(tcx::launch_row_stats(d_x, n, m, rs, st));
    f.row_stats = rs;
  }
  // the host copy of the initial labels must outlive the async upload
  TCX_HIP(hipStreamSynchronize(st));
  f.o = o;
  f.o.init_labels = nullptr;
  f.o.init_rows = nullptr;
  f.tol = o.power_tol > 0 ? o.power_tol : 1e-12;
  f.split = o.s_precision != TCX_S_BF16;  // TCX_S_AUTO starts split
  f.s_bf16_from = o.s_precision == TCX_S_BF16 ? 0 : -1;
  f.s_budget = o.s_budget > 0 ? o.s_budget : 1e-4;
  f.d_x = d_x;
  f.n = n;
  f.m = m;
  f.k = k;
  f.active = true;
  return TCX_OK;
}

namespace {

int fit_refine(tcx_ctx* c, FitState& f) {
  hipStream_t st = c->stream;
  float* cnew;
  TCX_TRY(ws(c, kSlotCentNew, (size_t)f.k * f.m, &cnew));
  // TCX_S_AUTO: at iterations 2..4, until one passes, the refinement also
  // forms the BF16 centroids and measures their difference from the split
  // ones; within the budget, S runs in BF16 from the next iteration on
  float* check = nullptr;
  if (f.o.s_precision == TCX_S_AUTO && f.s_bf16_from < 0 && f.it >= 1 && f.s_checks < 3 &&
      f.o.cal_select == 0)
    TCX_TRY(ws(c, kSlotCheck, (size_t)f.k * f.m, &check));
  TCX_HIP(hipEventRecord(c->ev[0], st));
  RefineResult rr;
  TCX_TRY(refine(c, f.d_x, f.n, f.m, f.k, f.label, f.shift, f.zero, cnew, f.it, f.o.seed,
                 f.split, f.tol, f.o.cal_select, f.row_stats, &rr,
                 f.have_dist ? f.dist : nullptr, check));
  if (check) {
    ++f.s_checks;
    f.s_check_err = rr.check_err;
    if (rr.check_err <= f.s_budget) {
      f.split = 0;
      f.s_bf16_from = f.it + 1;
    }
  }
  // the new centroids become the current ones: the two buffers swap roles
  std::swap(c->slot_ptr[kSlotCent], c->slot_ptr[kSlotCentNew]);
  std::swap(c->slot_bytes[kSlotCent], c->slot_bytes[kSlotCentNew]);
  f.cent = cnew;
  TCX_TRY(spectra(c, f.cent, f.k, f.m, &f.spec, &f.norm, &f.zero, f.cm));
  TCX_HIP(hipEventRecord(c->ev[1], st));
  f.ref_timed = true;  // read at the next synchronization (fit_assign)
  f.sum_class = rr.sum_class;
  f.n_pad = rr.n_pad;
  f.have_order = true;
  f.steps_max = std::max(f.steps_max, rr.steps_max);
  f.fallbacks += rr.fallbacks;
  return TCX_OK;
}

int fit_assign(tcx_ctx* c, FitState& f, int* h_changed) {
  hipStream_t st = c->stream;
  // the evaluation count and the labels-changed flag in one buffer (one
  // fill, one readback), on different 128-byte lines: the count takes an
  // atomic add per block, the flag plain stores
  constexpr int kChangedAt = 128 / sizeof(unsigned long long);
  unsigned long long* evals;
  TCX_TRY(ws(c, kSlotEvals, kChangedAt + 1, &evals));
  int* changed = reinterpret_cast<int*>(evals + kChangedAt);
  TCX_HIP(tcx::launch_fill_int(reinterpret_cast<int*>(evals), 2 * (kChangedAt + 1), 0, st));
  TCX_HIP(hipEventRecord(c->ev[2], st));
  const int* order = f.have_order ? reinterpret_cast<const int*>(c->slot_ptr[kSlotPerm]) : nullptr;
  if (f.xs) TCX_HIP(tcx::launch_bounds(fft_size(f.m), f.xm, f.n, f.cm, f.k, f.ub, f.ub_ld, st));
  TCX_HIP(hipEventRecord(c->ev[4], st));
  TCX_TRY(assign_core(c, f.d_x, f.n, f.m, f.k, f.spec, f.norm, f.o.convention, f.label, f.dist,
                      f.shift, nullptr, changed, order, f.n_pad / 2, evals, nullptr, f.xs,
                      f.xn, f.ub, f.ub_ld));
  f.have_dist = true;
  TCX_HIP(hipEventRecord(c->ev[3], st));
  void* pin;
  TCX_TRY(pinned_area(c, sizeof(unsigned long long) * (kChangedAt + 1), &pin));
  unsigned long long* p_evals = static_cast<unsigned long long*>(pin);
  int* p_changed = reinterpret_cast<int*>(p_evals + kChangedAt);
  TCX_HIP(hipMemcpyAsync(p_evals, evals, sizeof(*p_evals) * kChangedAt + sizeof(int),
                         hipMemcpyDeviceToHost, st));
  if (f.o.labels_per_iter && f.it < f.o.max_iter)
    TCX_HIP(hipMemcpyAsync(f.o.labels_per_iter + (size_t)f.it * f.n, f.label,
                           sizeof(int) * f.n, hipMemcpyDeviceToHost, st));
  TCX_HIP(hipStreamSynchronize(st));
  *h_changed = *p_changed;
  const unsigned long long h_evals = *p_evals;
  float ms = 0.f;
  if (f.ref_timed) {
    TCX_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    f.ms_ref += ms;
    f.ref_timed = false;
  }
  TCX_HIP(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
  f.ms_asg += ms;
  f.ms_asg_last = ms;
  TCX_HIP(hipEventElapsedTime(&ms, c->ev[2], c->ev[4]));
  f.ms_bnd += ms;
  f.evals += (long long)h_evals;
  f.pairs += (long long)f.n * f.k;
  return TCX_OK;
}

}  // namespace

int tcx_fit_step(tcx_ctx* c, int* changed) {
  TCX_TRY(set_dev(c));
  FitState& f = c->fit;
  if (!f.active) return fail(TCX_ERR_ARG, "fit_step: no fit in progress (call tcx_fit_begin)");
  if (f.collapsed) return fail(TCX_ERR_ARG, "fit_step: the fit stopped on collapse");
  int h_changed = 0;
  // the GPU path's stop after a refinement that leaves too few non-empty
  // clusters (kshape.cuh:248-249, 454-455)
  auto collapse = [&f]() {
    return f.o.stop_on_collapse && (f.sum_class == 1 || f.sum_class <= f.k / 2);
  };
  if (f.o.init_mode == 0) {
    // test_timesC.py:223-264: refine, assign, stop when labels repeat
    TCX_TRY(fit_refine(c, f));
    if (collapse()) {
      // no assignment: the labels stay those of the previous one
      f.collapsed = 1;
      if (changed) *changed = 0;
      return TCX_OK;
    }
    TCX_TRY(fit_assign(c, f, &h_changed));
    if (!h_changed) f.converged = 1;
  } else {
    // kshape.cuh:376-461 order: assign, stop test, refine
    TCX_TRY(fit_assign(c, f, &h_changed));
    if (f.it > 0 && !h_changed)
      f.converged = 1;
    else
      TCX_TRY(fit_refine(c, f));
    if (!f.converged && collapse()) f.collapsed = 1;
  }
  ++f.it;
  if (changed) *changed = f.it == 1 && f.o.init_mode == 1 ? 1 : h_changed;
  return TCX_OK;
}

// Adds a finished refine's event time to ms_ref (waits for it).
static int flush_refine_time(tcx_ctx* c) {
  FitState& f = c->fit;
  if (!f.ref_timed) return TCX_OK;
  TCX_HIP(hipEventSynchronize(c->ev[1]));
  float ms = 0.f;
  TCX_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  f.ms_ref += ms;
  f.ref_timed = false;
  return TCX_OK;
}

int tcx_fit_query(tcx_ctx* c, tcx_fit_stats* stats) {
  if (!c || !stats) return fail(TCX_ERR_ARG, "fit_query: null argument");
  const FitState& f = c->fit;
  if (!f.active) return fail(TCX_ERR_ARG, "fit_query: no fit in progress");
  TCX_TRY(set_dev(c));
  TCX_TRY(flush_refine_time(c));
  stats->iterations = f.it;
  stats->converged = f.converged;
  stats->sum_class = f.sum_class;
  stats->power_steps_max = f.steps_max;
  stats->fallbacks = f.fallbacks;
  stats->ms_refine = f.ms_ref;
  stats->ms_assign = f.ms_asg;
  stats->ms_assign_bounds = f.ms_bnd;
  stats->sbd_evals = f.evals;
  stats->sbd_pairs = f.pairs;
  stats->collapsed = f.collapsed;
  stats->s_bf16_from = f.s_bf16_from;
  stats->s_check_err = f.s_check_err;
  stats->ms_total = std::chrono::duration<double, std::milli>(
                        std::chrono::steady_clock::now() - f.t_start).count();
  return TCX_OK;
}

int tcx_fit_snapshot(tcx_ctx* c, int* h_labels, float* h_dist, int* h_shift, float* h_centroids) {
  TCX_TRY(set_dev(c));
  const FitState& f = c->fit;
  if (!f.active) return fail(TCX_ERR_ARG, "fit_snapshot: no fit in progress");
  if (h_dist && !f.have_dist) return fail(TCX_ERR_ARG, "fit_snapshot: no assignment yet");
  hipStream_t st = c->stream;
  if (h_labels)
    TCX_HIP(hipMemcpyAsync(h_labels, f.label, sizeof(int) * f.n, hipMemcpyDeviceToHost, st));
  if (h_dist)
    TCX_HIP(hipMemcpyAsync(h_dist, f.dist, sizeof(float) * f.n, hipMemcpyDeviceToHost, st));
  if (h_shift)
    TCX_HIP(hipMemcpyAsync(h_shift, f.shift, sizeof(int) * f.n, hipMemcpyDeviceToHost, st));
  if (h_centroids)
    TCX_HIP(hipMemcpyAsync(h_centroids, f.cent, sizeof(float) * f.k * f.m, hipMemcpyDeviceToHost,
                           st));
  TCX_HIP(hipStreamSynchronize(st));
  return TCX_OK;
}

int tcx_fit_end(tcx_ctx* c, int* h_labels, float* h_centroids, tcx_fit_stats* stats) {
  TCX_TRY(set_dev(c));
  FitState& f = c->fit;
  if (!f.active) return fail(TCX_ERR_ARG, "fit_end: no fit in progress");
  hipStream_t st = c->stream;
  if (h_labels)
    TCX_HIP(hipMemcpyAsync(h_labels, f.label, sizeof(int) * f.n, hipMemcpyDeviceToHost, st));
  if (h_centroids)
    TCX_HIP(hipMemcpyAsync(h_centroids, f.cent, sizeof(float) * f.k * f.m,
                           hipMemcpyDeviceToHost, st));
  TCX_HIP(hipStreamSynchronize(st));
  TCX_TRY(flush_refine_time(c));
  if (stats) {
    stats->iterations = f.it;
    stats->converged = f.converged;
    stats->sum_class = f.sum_class;
    stats->power_steps_max = f.steps_max;
    stats->fallbacks = f.fallbacks;
    stats->ms_refine = f.ms_ref;
    stats->ms_assign = f.ms_asg;
    stats->ms_assign_bounds = f.ms_bnd;
    stats->sbd_evals = f.evals;
    stats->sbd_pairs = f.pairs;
    stats->collapsed = f.collapsed;
    stats->s_bf16_from = f.s_bf16_from;
    stats->s_check_err = f.s_check_err;
    stats->ms_total = std::chrono::duration<double, std::milli>(
                          std::chrono::steady_clock::now() - f.t_start).count();
  }
  f.active = false;
  return TCX_OK;
}

int tcx_kshape_fit(tcx_ctx* c, const float* d_x, int n, int m, int k, const tcx_fit_opts* opts,
                   int* h_labels, float* h_centroids, tcx_fit_stats* stats) {
  if (!h_labels) return fail(TCX_ERR_ARG, "fit: null label buffer");
  TCX_TRY(tcx_fit_begin(c, d_x, n, m, k, opts));
  FitState& f = c->fit;
  while (f.it < f.o.max_iter && !f.converged && !f.collapsed) {
    int rc = tcx_fit_step(c, nullptr);
    if (rc != TCX_OK) {
      f.active = false;
      return rc;
    }
  }
  return tcx_fit_end(c, h_labels, h_centroids, stats);
}

int tcx_kshape(tcx_ctx* c, const float* d_x, int* h_idx, int n, int m, int k, int* iter_num,
               int* sum_class, int cal_select) {
  if (!iter_num || !h_idx) return fail(TCX_ERR_ARG, "kshape: null pointer");
  tcx_fit_opts o;
  tcx_fit_opts_default(&o);
  o.max_iter = *iter_num;
  o.cal_select = cal_select;
  o.stop_on_collapse = 1;  // kshape.cuh:248-249
  tcx_fit_stats s;
  TCX_TRY(tcx_kshape_fit(c, d_x, n, m, k, &o, h_idx, nullptr, &s));
  for (int i = 0; i < n; ++i) h_idx[i] += 1;  // one-based, kshape.cuh:91
  // the reference sets iter_num only when the labels stop changing
  // (kshape.cuh:297); after a collapse it keeps the cap
  if (!s.collapsed) *iter_num = s.iterations;
  if (sum_class) *sum_class = s.sum_class;
  return TCX_OK;
}

int tcx_kshape_tsl(tcx_ctx* c, const float* d_x, const float* h_x, int* h_idx, int n, int m,
                   int k, int* iter_num, int* sum_class, int cal_select) {
  (void)h_x;
  if (!iter_num || !h_idx) return fail(TCX_ERR_ARG, "kshape_tsl: null pointer");
  tcx_fit_opts o;
  tcx_fit_opts_default(&o);
  o.max_iter = *iter_num;
  o.init_mode = 1;
  o.cal_select = cal_select;
  o.stop_on_collapse = 1;  // kshape.cuh:454-455
  tcx_fit_stats s;
  TCX_TRY(tcx_kshape_fit(c, d_x, n, m, k, &o, h_idx, nullptr, &s));
  for (int i = 0; i < n; ++i) h_idx[i] += 1;
  if (!s.collapsed) *iter_num = s.iterations;
  if (sum_class) *sum_class = s.sum_class;
  return TCX_OK;
}

}  // extern "C"

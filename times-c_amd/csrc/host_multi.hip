// Grouped executor of a batch of independent k-Shape fits.
//
// Every phase of an iteration runs as ONE grouped launch over all jobs
// that are in progress (tcx_kernels.hpp, GroupTable): the grouping, the
// align + pack, S = Y^T Y, the power stage, the centroid spectra, the
// bound product and the pruned assignment.  The host synchronizes three
// times per iteration for the whole batch (cluster sizes that plan S, the
// power stage's status, the assignments' changed flags), instead of three
// times per job, and a launch holds the blocks of every job, so the many
// small fits of a batch fill the device together.  Jobs of different FFT
// lengths (and kernel forms) take one launch per length (form).
//
// A job's results do not depend on the other jobs of the batch: its blocks
// run the same kernel bodies as a launch of its own.  The power stage is
// the fused one-block-per-cluster Lanczos for m <= 768 and the two-launch
// form (grid-wide matvec + per-cluster step) above, for every job (the
// single-fit path uses the multi-block form from m = 256 on), and S is
// one K chunk per tile, so the results equal those of a batch holding the
// job alone, and the single-fit engine's to rounding.
//
// Jobs are admitted while their device memory (one arena each) fits the
// budget; a finished job's arena is released and waiting jobs join at the
// next iteration.  When every arena fits the budget at once they are carved
// from one allocation.  No reference counterpart (the reference runs one fit per
// process, Code/test/test_kshape.cu:13-80).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tcx_host.hpp"
#include "tcx_internal.hpp"
#include "tcx_kernels.hpp"
#include "timesc.h"

using Clock = std::chrono::steady_clock;

namespace tcx {
namespace {

#define GX_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(TCX_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
  } while (0)

#define GX_TRY(call)               \
  do {                             \
    int rc_ = (call);              \
    if (rc_ != TCX_OK) return rc_; \
  } while (0)

constexpr size_t kAlign = 256;
size_t align_up(size_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }

int fft_len(int m) {
  unsigned v = (unsigned)(2 * m - 1);
  int bl = 0;
  while (v) {
    ++bl;
    v >>= 1;
  }
  return 1 << bl;
}

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// Readback layout of a job's grouping (offsets, counts, bad-label flag).
size_t ginfo_bytes(int k) { return align_up(sizeof(long) * (k + 1) + sizeof(int) * (k + 2)); }
constexpr size_t kEvcBytes = 256;  // evaluation count (8 B), changed flag at +128

struct GJob {
  tcx_job* job = nullptr;
  long n = 0;
  int m = 0, k = 0, F = 0, kp = 0;
  tcx_fit_opts o{};
  double tol = 1e-12;
  int split = 1;
  long nblk = 0, n_pad_max = 0;
  int Bmax = 1;
  // state
  bool admitted = false, done = false;
  int it = 0, converged = 0, collapsed = 0, sum_class = 0, steps_max = 0, fallbacks = 0;
  bool have_dist = false;
  long n_pad = 0;
  int nonempty = 0;
  long long evals = 0, pairs = 0;
  double ms_refine = 0, ms_assign = 0;
  Clock::time_point t_admit;
  size_t arena_bytes = 0, slab_off = 0;
  unsigned char* arena = nullptr;
  // pools (offsets into the batch's readback pools)
  size_t ginfo_off = 0, info_off = 0, evc_off = 0;
  // device buffers (inside the arena)
  int *label = nullptr, *shift = nullptr, *zero = nullptr;
  float *dist = nullptr, *cent = nullptr, *mag = nullptr, *norm = nullptr;
  float2* spec = nullptr;
  uint16_t* cm = nullptr;
  float2* xs = nullptr;
  float* xn = nullptr;
  uint16_t* xm = nullptr;
  float* ub = nullptr;
  double2* row_stats = nullptr;
  int *blk_count = nullptr, *blk_base = nullptr, *kcounts = nullptr, *minrow = nullptr;
  int* perm = nullptr;
  long* koff = nullptr;
  uint16_t *yhi = nullptr, *ylo = nullptr;
  float *a0 = nullptr, *part = nullptr, *S = nullptr;
  double *Q = nullptr, *v = nullptr, *sv = nullptr, *r = nullptr, *w = nullptr;
  int *steps = nullptr, *conv = nullptr, *mask = nullptr;
  LzState* lzs = nullptr;
  int* scount = nullptr;  // counts of the S-form clusters, 0 elsewhere
  std::vector<int> mask_h;
  std::vector<char> fat, thin, mask_sel;  // S form, Gram form, eigenvector rerun
  int nfat = 0;
};

// Carves a job's buffers out of its arena (base NULL: sizes only); returns
// the arena's size.
size_t layout(GJob& j, unsigned char* base) {
  size_t off = 0;
  auto take = [&](auto*& p, size_t count) {
    using T = std::remove_reference_t<decltype(*p)>;
    p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += align_up(sizeof(T) * std::max<size_t>(count, 1));
  };
  const size_t n = (size_t)j.n, k = (size_t)j.k, m = (size_t)j.m, F = (size_t)j.F;
  take(j.label, n);
  take(j.shift, n);
  take(j.dist, n);
  take(j.cent, k * m);
  take(j.spec, k * F);
  take(j.mag, k * F);
  take(j.norm, k);
  take(j.zero, k);
  take(j.cm, (size_t)bound_rows(j.k) * bound_kpad(j.F));
  take(j.xs, n * (F / 2 + 1));
  take(j.xn, n);
  take(j.xm, (size_t)bound_rows(j.n) * bound_kpad(j.F));
  take(j.ub, n * (size_t)j.kp);
  if (align_pack_uses_row_stats(j.m)) take(j.row_stats, n);
  const size_t K = k * (size_t)j.Bmax;
  take(j.blk_count, (size_t)j.nblk * K);
  take(j.blk_base, (size_t)j.nblk * K);
  take(j.kcounts, K);
  take(j.koff, K);
  take(j.minrow, k);
  take(j.perm, (size_t)j.n_pad_max);
  take(j.yhi, (size_t)panel_blocks(j.m) * 32 * (size_t)j.n_pad_max);
  take(j.ylo, (size_t)panel_blocks(j.m) * 32 * (size_t)j.n_pad_max);
  take(j.a0, k * m);
  const size_t T = (m + kSyrkSub - 1) / kSyrkSub, U = T * (T + 1) / 2;
  take(j.part, k * U * kSyrkSub * kSyrkSub);
  take(j.S, k * m * m);
  take(j.Q, k * (kLanczosMax + 1) * m);
  take(j.v, k * m);
  take(j.sv, k * m);
  take(j.r, k * m);
  take(j.w, k * m);
  take(j.steps, k);
  take(j.conv, k);
  take(j.mask, k);
  take(j.lzs, k);
  take(j.scount, k);
  return off;
}

// Host-assembled job tables uploaded by one copy (through a pinned area)
// into a grow-only device buffer.  Reused only after the stream has
// consumed the previous upload (every upload follows a synchronization).
// The buffers are the context's (ctx_buffer ids dev_id, pin_id), kept
// between calls.
class Uploader {
 public:
  ~Uploader() {
    if (copied_) (void)hipEventSynchronize(copied_);
    if (copied_) (void)hipEventDestroy(copied_);
  }
  void bind(tcx_ctx* c, int dev_id, int pin_id) {
    ctx_ = c;
    dev_id_ = dev_id;
    pin_id_ = pin_id;
  }
  void reset() { buf_.clear(); }
  size_t put(const void* p, size_t bytes) {
    const size_t off = align_up(buf_.size());
    buf_.resize(off + std::max<size_t>(bytes, 1));
    if (bytes) std::memcpy(buf_.data() + off, p, bytes);
    return off;
  }
  template <class T>
  size_t put_vec(const std::vector<T>& v) {
    return put(v.data(), sizeof(T) * v.size());
  }
  int flush(hipStream_t st) {
    const size_t n = buf_.size();
    if (n == 0) return TCX_OK;
    if (n > dev_cap_ || n > pin_cap_) {
      // growing frees the old buffers (hipFree waits for the device)
      GX_HIP(hipStreamSynchronize(st));
      if (copied_) GX_HIP(hipEventSynchronize(copied_));
      const size_t cap = std::max(n * 2, (size_t)1 << 20);
      if (n > dev_cap_) {
        dev_ = nullptr;
        dev_cap_ = 0;
        GX_TRY(ctx_buffer(ctx_, dev_id_, cap, false, &dev_));
        dev_cap_ = cap;
      }
      if (n > pin_cap_) {
        pin_ = nullptr;
        pin_cap_ = 0;
        GX_TRY(ctx_buffer(ctx_, pin_id_, cap, true, &pin_));
        pin_cap_ = cap;
      }
    }
    // the previous upload must have left the pinned area before it is
    // rewritten (the device buffer is rewritten in stream order)
    if (!copied_) GX_HIP(hipEventCreateWithFlags(&copied_, hipEventDisableTiming));
    GX_HIP(hipEventSynchronize(copied_));
    std::memcpy(pin_, buf_.data(), n);
    GX_HIP(hipMemcpyAsync(dev_, pin_, n, hipMemcpyHostToDevice, st));
    GX_HIP(hipEventRecord(copied_, st));
    return TCX_OK;
  }
  const unsigned char* dev(size_t off) const { return static_cast<const unsigned char*>(dev_) + off; }

 private:
  std::vector<unsigned char> buf_;
  tcx_ctx* ctx_ = nullptr;
  int dev_id_ = 0, pin_id_ = 0;
  void* dev_ = nullptr;
  void* pin_ = nullptr;
  hipEvent_t copied_ = nullptr;
  size_t dev_cap_ = 0, pin_cap_ = 0;
};

// A grouped launch being assembled: job structs and per-job block counts.
template <class T>
struct Launch {
  std::vector<T> jobs;
  std::vector<int> first{0};
  size_t lds = 0;
  size_t jobs_off = 0, first_off = 0;
  void add(const T& a, long blocks, size_t l = 0) {
    if (blocks <= 0) return;
    jobs.push_back(a);
    first.push_back(first.back() + (int)blocks);
    lds = std::max(lds, l);
  }
  bool empty() const { return jobs.empty(); }
  void stage(Uploader& up) {
    if (jobs.empty()) return;
    jobs_off = up.put_vec(jobs);
    first_off = up.put_vec(first);
  }
  GroupTable table(const Uploader& up) const {
    GroupTable g;
    if (jobs.empty()) return g;
    g.jobs = up.dev(jobs_off);
    g.first = reinterpret_cast<const int*>(up.dev(first_off));
    g.njobs = (int)jobs.size();
    g.blocks = first.back();
    g.lds = lds;
    return g;
  }
};

// Jobs split by a key (FFT length, kernel form): one launch per key.
template <class T>
struct Keyed {
  std::vector<std::pair<long, Launch<T>>> v;
  Launch<T>& at(long key) {
    for (auto& e : v)
      if (e.first == key) return e.second;
    v.emplace_back(key, Launch<T>());
    return v.back().second;
  }
  void stage(Uploader& up) {
    for (auto& e : v) e.second.stage(up);
  }
};

// Context buffers (ctx_buffer ids) of the executor.
enum CtxBuf {
  kCbSlab, kCbGinfo, kCbInfo, kCbEvc, kCbGinfoH, kCbInfoH, kCbEvcH, kCbDoneD, kCbDoneH,
  kCbOutH, kCbInitH, kCbUpDev, kCbUpPin, kCbAsgDev, kCbAsgPin, kCbCount
};
static_assert(kCbCount <= kCtxBuffers, "context buffers");

struct Batch {
  tcx_ctx* ctx = nullptr;
  hipStream_t st = nullptr;
  // second stream: the block-form Gram Lanczos runs beside the one-wave
  // form (default priority; joined back into st by an event)
  hipStream_t st2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int cus = 0;
  std::vector<GJob> jobs;
  Uploader up;
  // readback pools (device + pinned host), sized for every job of the run;
  // these, the done flags, the slab and the pinned staging are the
  // context's buffers (kept between calls)
  unsigned char *ginfo_d = nullptr, *info_d = nullptr, *evc_d = nullptr;
  unsigned char *ginfo_h = nullptr, *info_h = nullptr, *evc_h = nullptr;
  size_t ginfo_bytes = 0, info_bytes = 0, evc_bytes = 0;
  int* done_h = nullptr;  // Lanczos done flags of the two-launch form
  int* done_d = nullptr;
  size_t done_cap = 0;
  // every job's arena in one allocation when they all fit the budget (no
  // allocation or free per job: hipFree waits for the device)
  unsigned char* slab = nullptr;
  unsigned char* out_h = nullptr;  // pinned staging of finished jobs' results
  size_t out_cap = 0;
  // the assignment phase as a graph, for the set of jobs it was captured
  // with, and that phase's own table buffer
  hipGraphExec_t asg_exec = nullptr;
  std::vector<GJob*> asg_set;
  Uploader up_asg;
  std::vector<GroupSyrkItem> items;
  std::vector<GroupSyrkTile> tiles;

  ~Batch() {
    if (st) (void)hipStreamSynchronize(st);
    if (st2) (void)hipStreamSynchronize(st2);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    for (GJob& j : jobs)
      if (j.arena && !slab) (void)hipFree(j.arena);
    if (asg_exec) (void)hipGraphExecDestroy(asg_exec);
  }

  long* offsets(GJob& j) { return reinterpret_cast<long*>(ginfo_d + j.ginfo_off); }
  int* counts(GJob& j) { return reinterpret_cast<int*>(offsets(j) + j.k + 1); }
  int* bad(GJob& j) { return counts(j) + j.k; }
  PowerInfo* info(GJob& j) { return reinterpret_cast<PowerInfo*>(info_d + j.info_off); }
  unsigned long long* evals(GJob& j) {
    return reinterpret_cast<unsigned long long*>(evc_d + j.evc_off);
  }
  int* changed(GJob& j) { return reinterpret_cast<int*>(evc_d + j.evc_off + 128); }
};

// ----------------------------------------------------------------- setup
// z-normalization (when asked), series transforms, row sums, initial
// labels, zero centroids of the newly admitted jobs.
int setup_jobs(Batch& b, const std::vector<GJob*>& fresh) {
  Keyed<ZnormJob> zn;
  Keyed<SeriesSpecJob> ss;
  Launch<RowStatsJob> rs;
  Launch<FillSeg> fills;
  long fill_max = 0;
  for (GJob* j : fresh) {
    if (j->job->znorm_ddof >= 0) {
      const int nv = znorm_vectors(j->job->d_x, j->m);
      zn.at(nv).add(ZnormJob{j->job->d_x, j->n, j->m, j->job->znorm_ddof}, znorm_blocks(j->n));
    }
    ss.at(j->F).add(SeriesSpecJob{j->job->d_x, j->n, j->m, j->xs, j->xn, j->xm},
                    series_spec_blocks(j->F, j->n));
    if (j->row_stats)
      rs.add(RowStatsJob{j->job->d_x, j->n, j->m, j->row_stats}, znorm_blocks(j->n));
    fills.add(FillSeg{reinterpret_cast<int*>(j->cent), (long)j->k * j->m, 0}, 1);
    fills.add(FillSeg{j->shift, j->n, 0}, 1);
    fills.add(FillSeg{j->zero, j->k, 1}, 1);  // the initial centroids are zero
    fill_max = std::max(fill_max, std::max(j->n, (long)j->k * j->m));
  }
  b.up.reset();
  zn.stage(b.up);
  ss.stage(b.up);
  rs.stage(b.up);
  const size_t fill_off = fills.jobs.empty() ? 0 : b.up.put_vec(fills.jobs);
  GX_TRY(b.up.flush(b.st));
  for (auto& e : zn.v) GX_HIP(launch_znorm_group(e.second.table(b.up), (int)e.first, b.st));
  for (auto& e : ss.v) {
    const void* tw;
    GX_TRY(twiddle_table(b.ctx, (int)e.first, &tw));
    GX_HIP(launch_series_spec_group((int)e.first, e.second.table(b.up),
                                    static_cast<const float2*>(tw), b.st));
  }
  GX_HIP(launch_row_stats_group(rs.table(b.up), b.st));
  if (!fills.jobs.empty())
    GX_HIP(launch_fill_segs(reinterpret_cast<const FillSeg*>(b.up.dev(fill_off)),
                            (int)fills.jobs.size(), fill_max, b.st));
  // initial labels (the caller's or the seeded ones), staged in one pinned
  // area, one copy per job, one synchronization
  long total = 0;
  for (GJob* j : fresh) total += j->n;
  int* pin = nullptr;
  {
    void* p = nullptr;
    GX_TRY(ctx_buffer(b.ctx, kCbInitH, sizeof(int) * std::max(total, 1L), true, &p));
    pin = static_cast<int*>(p);
  }
  long at = 0;
  int rc = TCX_OK;
  for (GJob* j : fresh) {
    if (j->o.init_labels)
      std::memcpy(pin + at, j->o.init_labels, sizeof(int) * j->n);
    else
      tcx_init_labels(j->o.seed, j->n, j->k, pin + at);
    if (hipMemcpyAsync(j->label, pin + at, sizeof(int) * j->n, hipMemcpyHostToDevice, b.st) !=
        hipSuccess)
      rc = fail(TCX_ERR_HIP, "grouped batch: initial label upload failed");
    at += j->n;
  }
  if (hipStreamSynchronize(b.st) != hipSuccess && rc == TCX_OK)
    rc = fail(TCX_ERR_HIP, "grouped batch: initial label upload failed");
  return rc;
}

// --------------------------------------------------------------- refine
// Grouping and align + pack of every job, then the cluster sizes read back.
int refine_group_pack(Batch& b, const std::vector<GJob*>& act) {
  Launch<FillSeg> fills;
  Launch<GroupCountJob> gc;
  Launch<GroupScanJob> gs;
  Launch<GroupOffsetsJob> go;
  Launch<GroupPlaceJob> gp;
  Keyed<AlignPackJob> ap;
  long fill_max = 1;
  for (GJob* j : act) {
    const int B = j->have_dist ? j->Bmax : 1;
    const float* dist = j->have_dist ? j->dist : nullptr;
    fills.add(FillSeg{b.bad(*j), 1, 0}, 1);
    fills.add(FillSeg{j->minrow, j->k, 0x7fffffff}, 1);
    fills.add(FillSeg{j->perm, j->n_pad_max, -1}, 1);
    fill_max = std::max(fill_max, j->n_pad_max);
    int lds_min = 0;
    const size_t lds = group_count_lds(j->k, B, &lds_min);
    gc.add(GroupCountJob{j->label, j->n, j->k, dist, B, j->blk_count, j->minrow, lds_min,
                         b.bad(*j)},
           j->nblk, lds);
    gs.add(GroupScanJob{(int)j->nblk, j->k * B, j->blk_count, j->blk_base, j->kcounts},
           (long)j->k * B);
    go.add(GroupOffsetsJob{j->k, B, j->kcounts, b.counts(*j), b.offsets(*j), j->koff}, 1);
    gp.add(GroupPlaceJob{j->label, j->n, j->k, dist, B, j->blk_base, j->koff, j->perm}, j->nblk,
           group_place_lds(j->k, B));
    const int form = align_pack_form(j->m);
    ap.at(form).add(AlignPackJob{j->job->d_x, j->m, j->perm, j->n_pad_max, b.offsets(*j),
                                 j->minrow, j->k, j->shift, j->zero, j->row_stats, j->yhi,
                                 j->ylo, j->a0},
                    j->n_pad_max / align_pack_cols(form), align_pack_lds(form, j->m));
  }
  b.up.reset();
  const size_t fill_off = b.up.put_vec(fills.jobs);
  gc.stage(b.up);
  gs.stage(b.up);
  go.stage(b.up);
  gp.stage(b.up);
  ap.stage(b.up);
  GX_TRY(b.up.flush(b.st));
  GX_HIP(launch_fill_segs(reinterpret_cast<const FillSeg*>(b.up.dev(fill_off)),
                          (int)fills.jobs.size(), fill_max, b.st));
  GX_HIP(launch_group_count_group(gc.table(b.up), b.st));
  GX_HIP(launch_group_scan_group(gs.table(b.up), b.st));
  GX_HIP(launch_group_offsets_group(go.table(b.up), b.st));
  GX_HIP(launch_group_place_group(gp.table(b.up), b.st));
  for (auto& e : ap.v) GX_HIP(launch_align_pack_group((int)e.first, e.second.table(b.up), b.st));
  GX_HIP(hipMemcpyAsync(b.ginfo_h, b.ginfo_d, b.ginfo_bytes, hipMemcpyDeviceToHost, b.st));
  GX_HIP(hipStreamSynchronize(b.st));
  return TCX_OK;
}

// The two-launch Lanczos (grid-wide matvec + per-cluster step) of the jobs
// `set` in one grouped launch per step; mode 0: M^I v0 of the clusters
// with sel[c] (their count in lzc), mode 1: top eigenvector of the masked
// clusters (sel = the mask).  A job takes steps while s < min(m,
// kLanczosMax) and some selected cluster of it is not done, as the
// single-fit loop does.  Ends with v, S v, r, the step counts and
// acceptance flags of the selected clusters.
int power_split(Batch& b, const std::vector<GJob*>& set, int mode) {
  if (set.empty()) return TCX_OK;
  int jmax_all = 0;
  long ndone = 0;
  for (GJob* j : set) {
    jmax_all = std::max(jmax_all, std::min(j->m, kLanczosMax));
    ndone += j->k;
  }
  if ((size_t)ndone > b.done_cap) {
    // growing frees the old buffers: nothing in flight may use them
    GX_HIP(hipStreamSynchronize(b.st));
    GX_HIP(hipStreamSynchronize(b.st2));
    b.done_h = nullptr;
    b.done_d = nullptr;
    b.done_cap = 0;
    void *ph = nullptr, *pd = nullptr;
    GX_TRY(ctx_buffer(b.ctx, kCbDoneH, sizeof(int) * ndone, true, &ph));
    GX_TRY(ctx_buffer(b.ctx, kCbDoneD, sizeof(int) * ndone, false, &pd));
    b.done_h = static_cast<int*>(ph);
    b.done_d = static_cast<int*>(pd);
    b.done_cap = (size_t)ndone;
  }
  Launch<ClusterMatvecJob> mv_r, mv_v;
  std::vector<Launch<ClusterMatvecJob>> mv_w(jmax_all);
  std::vector<Launch<LzJob>> lz_s(jmax_all);
  Launch<LzJob> lz;
  long at = 0;
  for (GJob* j : set) {
    const long blocks = cluster_matvec_blocks(j->m, j->k);
    const long blocks1 = cluster_matvec_blocks(j->m, 1);
    const size_t lds = sizeof(double) * j->m;
    const long qs = (long)(kLanczosMax + 1) * j->m;
    const long mm = (long)j->m * j->m;
    const int jmax = std::min(j->m, kLanczosMax);
    const int steps = j->m > 1000 ? 1000 : j->m;
    const std::vector<char>& sel = mode ? j->mask_sel : j->fat;
    const LzJob lj{j->m, mode ? b.counts(*j) : j->scount, mode ? j->mask : nullptr, j->o.seed,
                   j->it, steps, j->r, j->w, j->Q, j->lzs, j->v, j->steps, j->conv,
                   b.done_d + at};
    for (int c = 0; c < j->k; ++c) {
      if (!sel[c]) continue;
      mv_r.add(ClusterMatvecJob{j->S + c * mm, j->m, 1, nullptr, 0, j->r + (long)c * j->m,
                                nullptr},
               blocks1, lds);
      mv_v.add(ClusterMatvecJob{j->S + c * mm, j->m, 1, j->v + (long)c * j->m, j->m,
                                j->sv + (long)c * j->m, nullptr},
               blocks1, lds);
    }
    for (int s = 0; s < jmax; ++s) {
      mv_w[s].add(ClusterMatvecJob{j->S, j->m, j->k, j->Q + (long)s * j->m, qs, j->w, j->lzs},
                  blocks, lds);
      lz_s[s].add(lj, j->k, lz_step_lds(j->m));
    }
    lz.add(lj, j->k);
    at += j->k;
  }
  b.up.reset();
  mv_r.stage(b.up);
  for (auto& l : mv_w) l.stage(b.up);
  for (auto& l : lz_s) l.stage(b.up);
  mv_v.stage(b.up);
  lz.stage(b.up);
  GX_TRY(b.up.flush(b.st));
  const double tol = mode == 0 ? kLzPowerTol : 1e-12;
  GX_HIP(launch_cluster_matvec_group(mv_r.table(b.up), b.st));  // r = S 1
  GX_HIP(launch_lz_setup_group(lz.table(b.up), b.st));
  for (int s = 0; s < jmax_all; ++s) {
    GX_HIP(launch_cluster_matvec_group(mv_w[s].table(b.up), b.st));
    GX_HIP(launch_lz_step_group(lz_s[s].table(b.up), mode, tol, b.st));
    const int len = s + 1;
    bool check = mode == 0 ? lz_power_check(len) : (len >= 8 && (len & 7) == 0);
    for (GJob* j : set) check |= len == std::min(j->m, kLanczosMax);
    if (!check) continue;
    // stop once every job has reached its step limit or has every cluster
    // done (one readback of all the done flags)
    GX_HIP(hipMemcpyAsync(b.done_h, b.done_d, sizeof(int) * ndone, hipMemcpyDeviceToHost, b.st));
    GX_HIP(hipStreamSynchronize(b.st));
    bool all = true;
    at = 0;
    for (GJob* j : set) {
      bool jd = len >= std::min(j->m, kLanczosMax);  // step limit reached
      if (!jd) {
        jd = true;
        for (int c = 0; c < j->k; ++c) jd = jd && b.done_h[at + c] != 0;
      }
      all = all && jd;
      at += j->k;
    }
    if (all) break;
  }
  GX_HIP(launch_lz_finish_group(lz.table(b.up), b.st));
  GX_HIP(launch_cluster_matvec_group(mv_v.table(b.up), b.st));  // S v
  return TCX_OK;
}

// A cluster takes the Gram form when its padded segment nd is shorter than
// m: every Lanczos step then moves nd^2 < m^2 values.  Above order
// kGramBlockMaxND the Gram Lanczos is one block streaming nd^2 values per
// step, slower than the S form's two-launch steps over the whole GPU (on
// the 256-job uneven batch: no limit 190 ms, 1024 182 ms, 768 185 ms, 512
// 191 ms, 256 229 ms).
constexpr long kGramBlockMaxND = 1024;
bool gram_form(int m, long nd, int count) {
  return count > 0 && m <= kGramMaxM && nd > 0 && nd < m && nd <= kGramMaxND &&
         nd <= kGramBlockMaxND;
}

// S = Y^T Y items of the clusters with sel[c]: one per kSyrkSub block of
// the upper triangle (one wave each), over the cluster's whole column range.
void add_syrk_items(Batch& b, int q, const GJob& j, const long* off, const std::vector<char>& sel) {
  const int T = (j.m + kSyrkSub - 1) / kSyrkSub;
  int part = 0;
  for (int c = 0; c < j.k; ++c) {
    if (!sel[c]) continue;
    for (int ti = 0; ti < T; ++ti)
      for (int tj = ti; tj < T; ++tj) {
        b.items.push_back(GroupSyrkItem{q, SyrkItem{c, ti, tj, part, off[c], off[c + 1]}});
        b.tiles.push_back(GroupSyrkTile{q, SyrkTile{c, ti, tj, part, 1}});
        ++part;
      }
  }
}

// S of the selected clusters: syrk by precision, then the reduction.
int launch_syrk_items(Batch& b, const std::vector<GJob*>& jobs, const std::vector<SyrkJob>& sjobs,
                      size_t* sj_off, size_t by_off[2], size_t* tl_off,
                      std::vector<GroupSyrkItem> by[2]) {
  by[0].clear();
  by[1].clear();
  for (const GroupSyrkItem& gi : b.items) by[jobs[gi.job]->split].push_back(gi);
  *sj_off = b.up.put_vec(sjobs);
  by_off[0] = b.up.put_vec(by[0]);
  by_off[1] = b.up.put_vec(by[1]);
  *tl_off = b.up.put_vec(b.tiles);
  return TCX_OK;
}

int run_syrk(Batch& b, size_t sj_off, const size_t by_off[2], size_t tl_off,
             const std::vector<GroupSyrkItem> by[2]) {
  const SyrkJob* dsj = reinterpret_cast<const SyrkJob*>(b.up.dev(sj_off));
  for (int sp = 0; sp < 2; ++sp)
    GX_HIP(launch_syrk_group(dsj, reinterpret_cast<const GroupSyrkItem*>(b.up.dev(by_off[sp])),
                             (int)by[sp].size(), sp, b.st));
  GX_HIP(launch_syrk_reduce_group(dsj, reinterpret_cast<const GroupSyrkTile*>(b.up.dev(tl_off)),
                                  (int)b.tiles.size(), b.st));
  return TCX_OK;
}

// The power stage of every job: S (fat clusters) or the Gram matrix (thin
// clusters), the Lanczos evaluation of M^I v0, the centroids; then the
// power stage's status read back.  Clusters whose power result fails the
// sign / residual tests rerun through the eigenvector path on S (computed
// then for thin ones), as the single-fit refine does.
int refine_power(Batch& b, const std::vector<GJob*>& act, double* ms_power) {
  b.items.clear();
  b.tiles.clear();
  std::vector<SyrkJob> sjobs;
  std::vector<GramItem> gitems;
  std::vector<GramVecJob> gmap;
  std::vector<GramLzJob> glz;
  std::vector<LzWaveJob> wlz;
  size_t lds_map = 0, lds_lz = 0;
  Launch<FillSeg> fills;
  for (size_t q = 0; q < act.size(); ++q) {
    GJob* j = act[q];
    const long* off = reinterpret_cast<const long*>(b.ginfo_h + j->ginfo_off);
    const int* cnt = reinterpret_cast<const int*>(off + j->k + 1);
    const int bad = cnt[j->k];
    if (bad) return fail(TCX_ERR_ARG, "grouped batch: label outside [0, k)");
    j->n_pad = off[j->k];
    j->nonempty = 0;
    j->fat.assign(j->k, 0);
    j->thin.assign(j->k, 0);
    j->nfat = 0;
    const long mm = (long)j->m * j->m;
    const int steps = j->m > 1000 ? 1000 : j->m;
    for (int c = 0; c < j->k; ++c) {
      j->nonempty += cnt[c] > 0;
      const long nd = off[c + 1] - off[c];
      j->thin[c] = gram_form(j->m, nd, cnt[c]);
      j->fat[c] = cnt[c] > 0 && !j->thin[c];
      // fat clusters of short series: one wave each; the others by block
      const bool fat_wave = j->fat[c] && j->m <= kLzWaveMaxD;
      j->nfat += j->fat[c] && !fat_wave;
      fills.add(FillSeg{j->scount + c, 1, j->fat[c] && !fat_wave ? cnt[c] : 0}, 1);
      if (fat_wave)
        wlz.push_back(LzWaveJob{j->S + c * mm, j->m, 0, steps, j->o.seed, j->it, c, nullptr,
                                j->Q + (long)c * (kLanczosMax + 1) * j->m, j->v + (long)c * j->m,
                                j->r + (long)c * j->m, j->sv + (long)c * j->m, j->steps + c,
                                j->conv + c});
      if (!j->thin[c]) continue;
      // Gram order: the members rounded up to 4 (the matvec's float4 rows)
      const int gn = (cnt[c] + 3) & ~3;
      const int T = (gn + 63) / 64;
      double* rc = j->r + (long)c * j->m;
      double* wc = j->w + (long)c * j->m;
      // the diagonal tiles also form the start vector u0 (into r)
      for (int ti = 0; ti < T; ++ti)
        for (int tj = ti; tj < T; ++tj)
          gitems.push_back(GramItem{j->yhi, j->ylo, j->S + c * mm, off[c], j->m, (int)nd, gn, ti,
                                    tj, ti == tj ? rc : nullptr, j->o.seed, j->it, c});
      double* svc = j->sv + (long)c * j->m;
      double* Qc = j->Q + (long)c * (kLanczosMax + 1) * j->m;
      if (gn <= kLzWaveMaxD) {
        wlz.push_back(LzWaveJob{j->S + c * mm, gn, 1, steps - 1, 0, 0, 0, rc, Qc, wc, nullptr,
                                svc, j->steps + c, j->conv + c});
      } else {
        glz.push_back(GramLzJob{j->S + c * mm, gn, steps - 1, rc, Qc, wc, svc, j->steps + c,
                                j->conv + c});
        lds_lz = std::max(lds_lz, gram_lanczos_lds(gn));
      }
      gmap.push_back(
          GramVecJob{j->yhi, j->ylo, off[c], j->m, (int)nd, gn, wc, svc, j->v + (long)c * j->m,
                     svc, rc});
      lds_map = std::max(lds_map, gram_map_lds(j->m, (int)nd));
    }
    sjobs.push_back(SyrkJob{j->yhi, j->ylo, j->m, j->part, j->S});
    add_syrk_items(b, (int)q, *j, off, j->fat);
  }
  const auto t0 = Clock::now();
  Launch<PowerJob> pw;
  std::vector<GJob*> split_set;
  for (GJob* j : act) {
    const int steps = j->m > 1000 ? 1000 : j->m;
    // fat clusters of m > kLzWaveMaxD take the two-launch form: each Lanczos
    // step of all of them is one grid-wide matvec and one step kernel,
    // instead of one latency-bound block per cluster
    if (j->nfat > 0) split_set.push_back(j);
    pw.add(PowerJob{j->S, j->m, b.counts(*j), j->a0, j->o.seed, j->it, steps, j->tol, j->v, j->sv,
                    j->r, j->steps, j->conv, 0, nullptr, 1, j->cent, b.info(*j)},
           j->k, power_lds(j->m));
  }
  b.up.reset();
  size_t sj_off, by_off[2], tl_off;
  std::vector<GroupSyrkItem> by[2];
  GX_TRY(launch_syrk_items(b, act, sjobs, &sj_off, by_off, &tl_off, by));
  const size_t fill_off = b.up.put_vec(fills.jobs);
  const size_t gi_off = b.up.put_vec(gitems);
  GX_TRY(b.up.flush(b.st));
  if (!fills.jobs.empty())
    GX_HIP(launch_fill_segs(reinterpret_cast<const FillSeg*>(b.up.dev(fill_off)),
                            (int)fills.jobs.size(), 1, b.st));
  GX_TRY(run_syrk(b, sj_off, by_off, tl_off, by));
  GX_HIP(launch_gram_tiles(reinterpret_cast<const GramItem*>(b.up.dev(gi_off)), (int)gitems.size(),
                           b.st));
  // the two-launch form first: its finishing pass writes v, steps and conv
  // of every cluster of its jobs, the Gram form's included
  GX_TRY(power_split(b, split_set, 0));  // uploads its own tables
  // largest clusters first: a launch lasts as long as its slowest block,
  // so those start in the first wave of blocks (each cluster's result is
  // independent of its position)
  std::stable_sort(wlz.begin(), wlz.end(),
                   [](const LzWaveJob& a, const LzWaveJob& c) { return a.d > c.d; });
  // orders above kLzWaveSplitD (at the front) take four waves per cluster
  size_t nw4 = 0;
  while (nw4 < wlz.size() && wlz[nw4].d > kLzWaveSplitD) ++nw4;
  std::stable_sort(glz.begin(), glz.end(),
                   [](const GramLzJob& a, const GramLzJob& c) { return a.nd > c.nd; });
  b.up.reset();
  const size_t gl_off = b.up.put_vec(glz), gm_off = b.up.put_vec(gmap);
  const size_t wl_off = b.up.put_vec(wlz);
  pw.stage(b.up);
  GX_TRY(b.up.flush(b.st));
  if (!glz.empty()) {  // the block form on the second stream, after the upload
    GX_HIP(hipEventRecord(b.ev_fork, b.st));
    GX_HIP(hipStreamWaitEvent(b.st2, b.ev_fork, 0));
    GX_HIP(launch_gram_lanczos(reinterpret_cast<const GramLzJob*>(b.up.dev(gl_off)),
                               (int)glz.size(), lds_lz, kLzPowerTol, b.st2));
    GX_HIP(hipEventRecord(b.ev_join, b.st2));
  }
  const LzWaveJob* dwl = reinterpret_cast<const LzWaveJob*>(b.up.dev(wl_off));
  GX_HIP(launch_lanczos_wave(dwl, (int)nw4, 4, kLzPowerTol, b.st));
  GX_HIP(launch_lanczos_wave(dwl + nw4, (int)(wlz.size() - nw4), 1, kLzPowerTol, b.st));
  if (!glz.empty()) GX_HIP(hipStreamWaitEvent(b.st, b.ev_join, 0));
  GX_HIP(launch_gram_map(reinterpret_cast<const GramVecJob*>(b.up.dev(gm_off)), (int)gmap.size(),
                         lds_map, b.st));
  GX_HIP(launch_power_group(pw.table(b.up), b.st));
  GX_HIP(hipMemcpyAsync(b.info_h, b.info_d, b.info_bytes, hipMemcpyDeviceToHost, b.st));
  GX_HIP(hipStreamSynchronize(b.st));
  // clusters whose power result fails the tests: the eigenvector path
  std::vector<GJob*> redo;
  for (GJob* j : act) {
    const PowerInfo* h = reinterpret_cast<const PowerInfo*>(b.info_h + j->info_off);
    bool any = false;
    j->mask_h.assign(j->k, 0);
    j->mask_sel.assign(j->k, 0);
    for (int c = 0; c < j->k; ++c) {
      j->fallbacks += h[c].fallback;
      if (h[c].fallback && !h[c].empty && !h[c].zero) {
        j->mask_h[c] = 1;
        j->mask_sel[c] = 1;
        any = true;
      }
    }
    if (any)
      redo.push_back(j);
    else
      for (int c = 0; c < j->k; ++c) j->steps_max = std::max(j->steps_max, h[c].steps);
  }
  if (!redo.empty()) {
    // S of the masked thin clusters (their S was never formed)
    b.items.clear();
    b.tiles.clear();
    std::vector<SyrkJob> rjobs;
    for (size_t q = 0; q < redo.size(); ++q) {
      GJob* j = redo[q];
      const long* off = reinterpret_cast<const long*>(b.ginfo_h + j->ginfo_off);
      std::vector<char> need(j->k, 0);
      for (int c = 0; c < j->k; ++c) need[c] = j->mask_sel[c] && j->thin[c];
      rjobs.push_back(SyrkJob{j->yhi, j->ylo, j->m, j->part, j->S});
      add_syrk_items(b, (int)q, *j, off, need);
    }
    b.up.reset();
    GX_TRY(launch_syrk_items(b, redo, rjobs, &sj_off, by_off, &tl_off, by));
    GX_TRY(b.up.flush(b.st));
    GX_TRY(run_syrk(b, sj_off, by_off, tl_off, by));
    for (GJob* j : redo)
      GX_HIP(hipMemcpyAsync(j->mask, j->mask_h.data(), sizeof(int) * j->k, hipMemcpyHostToDevice,
                            b.st));
    GX_HIP(hipStreamSynchronize(b.st));
    GX_TRY(power_split(b, redo, 1));
    Launch<PowerJob> pe;
    for (GJob* j : redo) {
      const int steps = j->m > 1000 ? 1000 : j->m;
      pe.add(PowerJob{j->S, j->m, b.counts(*j), j->a0, j->o.seed, j->it, steps, j->tol, j->v,
                      j->sv, j->r, j->steps, j->conv, 1, j->mask, 0, j->cent, b.info(*j)},
             j->k, power_lds(j->m));
    }
    b.up.reset();
    pe.stage(b.up);
    GX_TRY(b.up.flush(b.st));
    GX_HIP(launch_power_group(pe.table(b.up), b.st));
    GX_HIP(hipMemcpyAsync(b.info_h, b.info_d, b.info_bytes, hipMemcpyDeviceToHost, b.st));
    GX_HIP(hipStreamSynchronize(b.st));
    for (GJob* j : redo) {
      const PowerInfo* h = reinterpret_cast<const PowerInfo*>(b.info_h + j->info_off);
      for (int c = 0; c < j->k; ++c) j->steps_max = std::max(j->steps_max, h[c].steps);
    }
  }
  *ms_power = ms_between(t0, Clock::now());
  // the new centroids were written in place (nothing in the refinement
  // reads the old ones: the alignment takes the assignment's shifts and
  // zero flags), so the assignment's pointers stay fixed
  for (GJob* j : act) j->sum_class = j->nonempty;
  return TCX_OK;
}

// --------------------------------------------------------------- assign
// Spectra of the new centroids, the bound products and the pruned
// assignments of every job; then the changed flags and evaluation counts.
// The phase reads only pointers fixed for a job's life (centroids are
// refined in place, the pair count is read on the device), so while the
// set of jobs in progress is unchanged it is one HIP graph, captured once
// and replayed (no memset nodes: the counters are cleared by a fill
// kernel).  TCX_GROUPED_GRAPH=0 enqueues the launches directly instead.
bool graphs_enabled() {
  const char* e = std::getenv("TCX_GROUPED_GRAPH");
  return !(e && e[0] == '0');
}

int enqueue_assign(Batch& b, const std::vector<GJob*>& act) {
  Keyed<SpectraJob> sp;
  Keyed<BoundsJob> bd;
  Keyed<PruneJob> pr;
  Launch<FillSeg> fills;
  for (GJob* j : act) {
    sp.at(j->F).add(SpectraJob{j->cent, j->k, j->m, j->spec, j->norm, j->zero, j->mag, j->cm},
                    spectra_blocks(j->F, j->k));
    fills.add(FillSeg{reinterpret_cast<int*>(b.evals(*j)), (long)(kEvcBytes / sizeof(int)), 0}, 1);
  }
  // the spectra must exist before the bounds read them: stream order
  // suffices (one stream)
  for (GJob* j : act) {
    const int form = bounds_form(j->F, j->k);
    bd.at((long)j->F * 8 + form)
        .add(BoundsJob{j->xm, j->n, j->cm, j->k, j->kp, j->ub}, bounds_blocks(form, j->n));
    const long npairs_max = j->n_pad_max / 2;
    pr.at(j->F).add(PruneJob{j->job->d_x, j->xs, j->xn, j->ub, j->kp, j->n, npairs_max, j->m, j->k,
                             j->spec, j->mag, j->norm, j->o.convention, j->perm, j->label,
                             j->dist, j->shift, b.changed(*j), b.evals(*j), b.offsets(*j) + j->k},
                    prune_blocks(j->F, npairs_max, j->k, j->n),
                    prune_lds(j->F, j->k));
  }
  std::vector<std::pair<int, const void*>> tws;
  std::vector<long> fs;
  for (auto& e : sp.v) fs.push_back(e.first);
  for (auto& e : pr.v) fs.push_back(e.first);
  for (long F : fs) {
    const void* tw;
    GX_TRY(twiddle_table(b.ctx, (int)F, &tw));  // may allocate: never under capture
    tws.emplace_back((int)F, tw);
  }
  auto tw_of = [&](long F) {
    for (auto& t : tws)
      if (t.first == F) return static_cast<const float2*>(t.second);
    return static_cast<const float2*>(nullptr);
  };
  b.up_asg.reset();
  sp.stage(b.up_asg);
  bd.stage(b.up_asg);
  pr.stage(b.up_asg);
  const size_t fill_off = b.up_asg.put_vec(fills.jobs);
  GX_TRY(b.up_asg.flush(b.st));
  const bool capture = graphs_enabled();
  if (capture) GX_HIP(hipStreamBeginCapture(b.st, hipStreamCaptureModeThreadLocal));
  int rc = TCX_OK;
  auto step = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == TCX_OK)
      rc = fail(TCX_ERR_HIP, std::string("grouped assignment: ") + what + ": " +
                                 hipGetErrorString(e));
  };
  for (auto& e : sp.v)
    step(launch_spectra_group((int)e.first, e.second.table(b.up_asg), tw_of(e.first), b.st),
         "spectra");
  step(launch_fill_segs(reinterpret_cast<const FillSeg*>(b.up_asg.dev(fill_off)),
                        (int)fills.jobs.size(), (long)(kEvcBytes / sizeof(int)), b.st),
       "counter fill");
  for (auto& e : bd.v)
    step(launch_bounds_group((int)(e.first / 8), (int)(e.first % 8), e.second.table(b.up_asg),
                             b.st),
         "bounds");
  for (auto& e : pr.v)
    step(launch_assign_prune_group((int)e.first, e.second.table(b.up_asg), tw_of(e.first), b.st),
         "pruned assignment");
  step(hipMemcpyAsync(b.evc_h, b.evc_d, b.evc_bytes, hipMemcpyDeviceToHost, b.st), "readback");
  if (capture) {
    hipGraph_t g = nullptr;
    step(hipStreamEndCapture(b.st, &g), "end capture");  // ends the capture in every case
    if (rc == TCX_OK) step(hipGraphInstantiate(&b.asg_exec, g, nullptr, nullptr, 0), "instantiate");
    if (g) (void)hipGraphDestroy(g);
  }
  return rc;
}

int assign_all(Batch& b, const std::vector<GJob*>& act) {
  if (!b.asg_exec || act != b.asg_set) {
    if (b.asg_exec) GX_HIP(hipGraphExecDestroy(b.asg_exec));
    b.asg_exec = nullptr;
    b.asg_set.clear();
    GX_TRY(enqueue_assign(b, act));
    if (b.asg_exec) b.asg_set = act;
  }
  if (b.asg_exec) GX_HIP(hipGraphLaunch(b.asg_exec, b.st));
  GX_HIP(hipStreamSynchronize(b.st));
  return TCX_OK;
}

int validate(const tcx_job& jb, const char** why) {
  static const char* kWhy[] = {"null data or label pointer", "bad sizes",
                               "series length > 4096", "init_mode 1 (centre initialisation)",
                               "cal_select 1", "k beyond the pruned assignment's limit",
                               "unknown convention or precision", "max_iter < 1"};
  int w = -1;
  if (!jb.d_x || !jb.h_labels) w = 0;
  else if (jb.n < 1 || jb.m < 2 || jb.k < 1) w = 1;
  else if (jb.m > kMaxFft / 2) w = 2;
  else if (jb.opts.init_mode != 0) w = 3;
  else if (jb.opts.cal_select != 0) w = 4;
  else if (jb.k > prune_max_k(fft_len(jb.m))) w = 5;
  else if ((jb.opts.convention != TCX_CONV_CPU && jb.opts.convention != TCX_CONV_KSHAPE) ||
           (jb.opts.s_precision != TCX_S_SPLIT_BF16 && jb.opts.s_precision != TCX_S_BF16))
    w = 6;
  else if (jb.opts.max_iter < 1) w = 7;
  if (why) *why = w < 0 ? "" : kWhy[w];
  return w;
}

}  // namespace

bool grouped_supported(const tcx_job& j, const char** why) { return validate(j, why) < 0; }

int grouped_fit(tcx_ctx* ctx, tcx_job* jobs, const int* idx, int count, size_t hbm_budget,
                Clock::time_point t0, void* second_stream) {
  Batch b;
  b.ctx = ctx;
  b.st = static_cast<hipStream_t>(stream_of(ctx));
  b.cus = cus_of(ctx);
  GX_HIP(hipSetDevice(device_of(ctx)));
  if (second_stream) {
    b.st2 = static_cast<hipStream_t>(second_stream);  // the caller's (idle) stream
  } else {
    void* st2 = nullptr;
    GX_TRY(ctx_aux_stream(ctx, &st2));
    b.st2 = static_cast<hipStream_t>(st2);
  }
  b.up.bind(ctx, kCbUpDev, kCbUpPin);
  b.up_asg.bind(ctx, kCbAsgDev, kCbAsgPin);
  GX_HIP(hipEventCreateWithFlags(&b.ev_fork, hipEventDisableTiming));
  GX_HIP(hipEventCreateWithFlags(&b.ev_join, hipEventDisableTiming));
  b.jobs.resize(count);
  for (int q = 0; q < count; ++q) {
    tcx_job& jb = jobs[idx[q]];
    GJob& j = b.jobs[q];
    j.job = &jb;
    jb.status = -1;
    jb.message[0] = '\0';
    jb.worker = 0;
    std::memset(&jb.stats, 0, sizeof(jb.stats));
    const char* why = nullptr;
    if (validate(jb, &why) >= 0) {
      jb.status = TCX_ERR_ARG;
      std::snprintf(jb.message, sizeof(jb.message), "grouped batch: %s", why);
      j.done = true;
      continue;
    }
    j.n = jb.n;
    j.m = jb.m;
    j.k = jb.k;
    j.F = fft_len(jb.m);
    j.kp = (jb.k + 15) & ~15;
    j.o = jb.opts;
    j.tol = jb.opts.power_tol > 0 ? jb.opts.power_tol : 1e-12;
    j.split = jb.opts.s_precision == TCX_S_SPLIT_BF16;
    j.nblk = (j.n + kGroupChunk - 1) / kGroupChunk;
    j.Bmax = group_bins(j.n, j.k, true);
    j.n_pad_max = std::max<long>((j.n + (long)j.k * (kPackPad - 1)) / kPackPad * kPackPad,
                                 kPackPad);
    if (jb.opts.init_labels)
      for (long i = 0; i < j.n; ++i)
        if (jb.opts.init_labels[i] < 0 || jb.opts.init_labels[i] >= j.k) {
          jb.status = TCX_ERR_ARG;
          std::snprintf(jb.message, sizeof(jb.message),
                        "grouped batch: initial label outside [0, k)");
          j.done = true;
          break;
        }
    j.arena_bytes = layout(j, nullptr);
    j.ginfo_off = b.ginfo_bytes;
    b.ginfo_bytes += ginfo_bytes(j.k);
    j.info_off = b.info_bytes;
    b.info_bytes += align_up(sizeof(PowerInfo) * j.k);
    j.evc_off = b.evc_bytes;
    b.evc_bytes += kEvcBytes;
  }
  const size_t pool[3] = {std::max<size_t>(b.ginfo_bytes, 1), std::max<size_t>(b.info_bytes, 1),
                          std::max<size_t>(b.evc_bytes, 1)};
  unsigned char** pd[3] = {&b.ginfo_d, &b.info_d, &b.evc_d};
  unsigned char** ph[3] = {&b.ginfo_h, &b.info_h, &b.evc_h};
  for (int i = 0; i < 3; ++i) {
    void *d = nullptr, *h = nullptr;
    GX_TRY(ctx_buffer(ctx, kCbGinfo + i, pool[i], false, &d));
    GX_TRY(ctx_buffer(ctx, kCbGinfoH + i, pool[i], true, &h));
    *pd[i] = static_cast<unsigned char*>(d);
    *ph[i] = static_cast<unsigned char*>(h);
  }
  size_t budget = hbm_budget;
  if (!budget) {  // most of the free memory and of what this context keeps for reuse
    size_t free_b = 0, total_b = 0;
    GX_HIP(hipMemGetInfo(&free_b, &total_b));
    budget = (free_b + ctx_buffer_device_bytes(ctx)) / 10 * 8;
  }
  size_t used = 0;
  {
    size_t total = 0;
    for (GJob& j : b.jobs)
      if (!j.done) {
        j.slab_off = total;
        total += align_up(j.arena_bytes);
      }
    void* p = nullptr;
    if (total > 0 && total <= budget && ctx_buffer(ctx, kCbSlab, total, false, &p) == TCX_OK)
      b.slab = static_cast<unsigned char*>(p);
    else
      clear_error();  // per-job allocations instead
  }
  for (;;) {
    // admit waiting jobs while their arenas fit (largest first, so the
    // long jobs start early)
    std::vector<GJob*> waiting;
    for (GJob& j : b.jobs)
      if (!j.done && !j.admitted) waiting.push_back(&j);
    std::stable_sort(waiting.begin(), waiting.end(),
                     [](const GJob* a, const GJob* c) { return a->arena_bytes > c->arena_bytes; });
    std::vector<GJob*> fresh;
    bool none_active = true;
    for (GJob& j : b.jobs) none_active &= !(j.admitted && !j.done);
    for (GJob* j : waiting) {
      if (used + j->arena_bytes > budget && !(none_active && fresh.empty())) continue;
      void* p = b.slab ? b.slab + j->slab_off : nullptr;
      if (!p && hipMalloc(&p, j->arena_bytes) != hipSuccess) {
        (void)hipGetLastError();
        if (none_active && fresh.empty()) {
          j->job->status = TCX_ERR_NOMEM;
          std::snprintf(j->job->message, sizeof(j->job->message),
                        "grouped batch: device allocation of %zu bytes failed", j->arena_bytes);
          j->done = true;
        }
        continue;
      }
      j->arena = static_cast<unsigned char*>(p);
      layout(*j, j->arena);
      j->admitted = true;
      j->t_admit = Clock::now();
      used += j->arena_bytes;
      fresh.push_back(j);
    }
    if (!fresh.empty()) GX_TRY(setup_jobs(b, fresh));
    std::vector<GJob*> act;
    for (GJob& j : b.jobs)
      if (j.admitted && !j.done) act.push_back(&j);
    if (act.empty()) {
      bool left = false;
      for (GJob& j : b.jobs) left |= !j.done;
      if (!left) break;
      continue;  // waiting jobs admitted next round
    }
    // one iteration of every active job: refine, then assign (the
    // label-initialized loop, test_timesC.py:223-264)
    auto ta = Clock::now();
    GX_TRY(refine_group_pack(b, act));
    double ms_power = 0;
    GX_TRY(refine_power(b, act, &ms_power));
    auto tb = Clock::now();
    std::vector<GJob*> asg;
    for (GJob* j : act) {
      j->ms_refine += ms_between(ta, tb);
      if (j->o.stop_on_collapse && (j->sum_class == 1 || j->sum_class <= j->k / 2)) {
        j->collapsed = 1;  // the labels stay those of the previous assignment
        j->done = true;
      } else {
        asg.push_back(j);
      }
    }
    if (!asg.empty()) GX_TRY(assign_all(b, asg));
    auto tc = Clock::now();
    std::vector<GJob*> finished;
    for (GJob* j : act) {
      if (!j->collapsed) {
        j->ms_assign += ms_between(tb, tc);
        const unsigned long long ev =
            *reinterpret_cast<const unsigned long long*>(b.evc_h + j->evc_off);
        const int ch = *reinterpret_cast<const int*>(b.evc_h + j->evc_off + 128);
        j->evals += (long long)ev;
        j->pairs += (long long)j->n * j->k;
        j->have_dist = true;
        if (j->o.labels_per_iter)
          GX_HIP(hipMemcpyAsync(j->o.labels_per_iter + (size_t)j->it * j->n, j->label,
                                sizeof(int) * j->n, hipMemcpyDeviceToHost, b.st));
        ++j->it;
        if (!ch) {
          j->converged = 1;
          j->done = true;
        } else if (j->it >= j->o.max_iter) {
          j->done = true;
        }
      }
      if (j->done) finished.push_back(j);
    }
    if (!finished.empty()) {
      // results through a pinned staging area (a copy into pageable host
      // memory is staged and synchronized by the runtime, one at a time)
      size_t need = 0;
      for (GJob* j : finished)
        need += align_up(sizeof(int) * j->n) + align_up(sizeof(float) * j->k * j->m);
      if (need > b.out_cap) {
        GX_HIP(hipStreamSynchronize(b.st));
        b.out_h = nullptr;
        b.out_cap = 0;
        void* p = nullptr;
        GX_TRY(ctx_buffer(b.ctx, kCbOutH, need * 2, true, &p));
        b.out_h = static_cast<unsigned char*>(p);
        b.out_cap = need * 2;
      }
      size_t at = 0;
      for (GJob* j : finished) {
        GX_HIP(hipMemcpyAsync(b.out_h + at, j->label, sizeof(int) * j->n, hipMemcpyDeviceToHost,
                              b.st));
        at += align_up(sizeof(int) * j->n);
        GX_HIP(hipMemcpyAsync(b.out_h + at, j->cent, sizeof(float) * j->k * j->m,
                              hipMemcpyDeviceToHost, b.st));
        at += align_up(sizeof(float) * j->k * j->m);
      }
      GX_HIP(hipStreamSynchronize(b.st));
      at = 0;
      for (GJob* j : finished) {
        std::memcpy(j->job->h_labels, b.out_h + at, sizeof(int) * j->n);
        at += align_up(sizeof(int) * j->n);
        if (j->job->h_centroids)
          std::memcpy(j->job->h_centroids, b.out_h + at, sizeof(float) * j->k * j->m);
        at += align_up(sizeof(float) * j->k * j->m);
      }
      const auto te = Clock::now();
      for (GJob* j : finished) {
        tcx_job& jb = *j->job;
        jb.status = TCX_OK;
        jb.stats.iterations = j->it;
        jb.stats.converged = j->converged;
        jb.stats.collapsed = j->collapsed;
        jb.stats.s_bf16_from = j->split ? -1 : 0;
        jb.stats.sum_class = j->sum_class;
        jb.stats.power_steps_max = j->steps_max;
        jb.stats.fallbacks = j->fallbacks;
        jb.stats.sbd_evals = j->evals;
        jb.stats.sbd_pairs = j->pairs;
        jb.stats.ms_refine = j->ms_refine;
        jb.stats.ms_assign = j->ms_assign;
        jb.stats.ms_total = ms_between(j->t_admit, te);
        jb.ms_queued = ms_between(t0, j->t_admit);
        jb.ms_run = ms_between(j->t_admit, te);
        jb.ms_latency = ms_between(t0, te);
        if (!b.slab) GX_HIP(hipFree(j->arena));
        j->arena = nullptr;
        used -= j->arena_bytes;
      }
    }
  }
  return TCX_OK;
}

}  // namespace tcx

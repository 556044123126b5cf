// Internal launchers of the HIP kernels (host-callable, stream-ordered).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tcx {

constexpr int kModeAssign = 0;
constexpr int kModeNcc = 1;

struct SbdArgs {
  const float* x = nullptr;      // [n][m] series
  long n = 0;
  int m = 0;
  int k = 0;                     // centroids (assign) or y rows (ncc)
  const float2* cspec = nullptr; // [k][F] lane-strided spectra
  const float* cnorm = nullptr;  // [k]
  const float2* tw = nullptr;    // [F] exp(-2 pi i q / F)
  int rev = 0;                   // 1: k-Shape lag order
  int* label = nullptr;
  float* dist = nullptr;
  int* shift = nullptr;
  float* dmat = nullptr;
  int* changed = nullptr;
  float* ncc_out = nullptr;
  const int* fixed_label = nullptr;  // shift against this centroid, no argmin
};

// largest supported FFT length (m <= 4096)
constexpr int kMaxFft = 8192;

hipError_t launch_znorm(float* x, long rows, int m, int ddof, hipStream_t st);
// Twiddle table (F entries) of the length-F transform, host side.
void fft_twiddle_table(int F, float2* out);
// bmag (may be NULL): rows (bound_rows(rows) x bound_kpad(F), fragment
// order) of |Y[f]| / |y| for f <= F/2, zero padded, as IEEE binary16 bits
// rounded up (the centroid operand of launch_bounds).
hipError_t launch_spectra(int F, const float* y, int rows, int m, const float2* tw,
                          float2* spec, float* norms, int* zero, float* mag, uint16_t* bmag,
                          hipStream_t st);

// Exact pruned assignment (bound-and-prune on the spectra), k <= kPruneMaxK.
constexpr int kPruneMaxK = 1024;
constexpr int kLdsBytes = 160 * 1024;  // LDS of one workgroup on gfx950
// Largest k the pruned assignment takes at FFT length F (LDS-bound).
int prune_max_k(int F);
// Columns of a series' bound row (launch_series_spec's xm): F/2+1
// magnitudes padded to a multiple of 16 (the bound product's K step).
__host__ __device__ constexpr int bound_kpad(int F) { return ((F / 2 + 1) + 15) & ~15; }
// Bound rows are IEEE binary16 values in MFMA fragment order: the 16 x 16
// block of rows 16 b .. 16 b + 15 and columns 16 s .. 16 s + 15 is 256
// consecutive values, element (r, c) at 4 (r + 16 (c / 4)) + c % 4, so that
// lane l of a wave reads its operand of v_mfma_f32_16x16x16_f16 (row l % 16,
// columns 4 (l / 16) .. + 3) as the l-th 8-byte vector of the block.  Rows
// are padded to a multiple of 16.
__host__ __device__ constexpr long bound_frag(long row, int f, int kpad) {
  return (((row >> 4) * (kpad >> 4) + (f >> 4)) * 64 + (row & 15) + 16 * ((f & 15) >> 2)) * 4 +
         (f & 3);
}
__host__ __device__ constexpr long bound_rows(long rows) { return (rows + 15) & ~15L; }
struct PruneArgs {
  const float* x = nullptr;
  const float2* xs = nullptr;    // per-series half spectra [n][F/2+1] (may be NULL: from x)
  const float* xn = nullptr;     // per-series norms (with xs)
  const float* ub = nullptr;     // per-series bounds [n][ub_ld] of launch_bounds (with xs)
  int ub_ld = 0;
  long n = 0;
  long npairs = 0;               // pairs of order[] (or ceil(n/2) without order)
  int m = 0, k = 0;
  const float2* cspec = nullptr;
  const float* cmag = nullptr;   // |C| in the spectrum layout
  const float* cnorm = nullptr;
  const float2* tw = nullptr;
  int rev = 0;
  const int* order = nullptr;    // 2*npairs row indices, -1 = hole (may be NULL)
  int* label = nullptr;
  float* dist = nullptr;
  int* shift = nullptr;
  int* changed = nullptr;
  unsigned long long* evals = nullptr;  // inverse transforms run (may be NULL)
};
hipError_t launch_assign_prune(int F, const PruneArgs& a, hipStream_t st);
// Half spectra X_r[0 .. F/2] of every row and the row norms (k_series_spec);
// with xm (may be NULL) also the bound rows (bound_rows(n) x bound_kpad(F),
// fragment order, binary16 rounded up): weight_f |X_r[f]| / |x_r|, weight 2
// for 0 < f < F/2, zero padded.
hipError_t launch_series_spec(int F, const float* x, long n, int m, const float2* tw, float2* xs,
                              float* xn, uint16_t* xm, hipStream_t st);
// ub[i][j] = sum_f xm[i][f] cm[j][f] / F (cm: launch_spectra's bmag rows),
// i < n, j < k, row stride ld >= k: the pruned assignment's bounds
// (k_bounds).  Both operands are rounded up, so ub is never below the
// product of the FP32 magnitudes.
hipError_t launch_bounds(int F, const uint16_t* xm, long n, const uint16_t* cm, int k, float* ub,
                         int ld, hipStream_t st);
hipError_t launch_sbd(int F, int mode, const SbdArgs& a, hipStream_t st);
hipError_t launch_apply_shift(float* x, long rows, int m, const int* shift, const int* skip,
                              hipStream_t st);
hipError_t configure_sbd_kernels();

// ------------------------------------------------------------ extraction
constexpr int kGroupChunk = 512;  // rows per grouping block
constexpr int kPackPad = 64;       // cluster segments padded to this many columns
constexpr int kSyrkTile = 256;     // panel rows are padded to a multiple of this
constexpr int kSyrkSub = 128;      // S block edge (one wave)

// The packed panel Y (time t x packed column p, BF16 hi and lo parts) is
// stored in MFMA operand order: chunk (c = p / 16, r = t / 32) holds 512
// values at element offset (c * panel_blocks(m) + r) * 512, value (t, p) at
// ((t % 32) + 32 * ((p / 8) % 2)) * 8 + p % 8 inside it.  Rows run to a
// multiple of kSyrkTile, rows t >= m are zero.
__host__ __device__ constexpr int panel_blocks(int m) {
  return (m + kSyrkTile - 1) / kSyrkTile * (kSyrkTile / 32);
}

// A work item of S = Y^T Y (grouped executor): one kSyrkSub block of S_j
// (one wave) over a packed column range.
struct SyrkItem {
  int cluster;
  int ti, tj;         // block row and column (ti <= tj)
  int part;           // index of this item's partial
  long p0, p1;        // packed column range [p0, p1), multiples of 16
};

// A kSyrkSub block of S_j and its K-chunk partials.
struct SyrkTile {
  int cluster;
  int ti, tj;
  int part0, nparts;  // partials part0 .. part0+nparts-1, summed in order
};

// A (cluster, K chunk) group of the single fit's S = Y^T Y: the upper
// triangle of kSyrkSub blocks of S_j (row-major order) over the packed
// columns [p0, p1); block u's partial is part + u * stride.
struct SyrkGroup {
  int part, stride;
  long p0, p1;        // multiples of 16
};

// A cluster of the single fit with partials: block u of S_j sums the
// nparts partials from part0 + u * nparts on, in order.
struct SyrkCluster {
  int cluster, part0, nparts;
};

struct PowerInfo {
  double lambda;
  int steps;
  int converged;
  int fallback;
  int flipped;
  int zero;           // power iteration hit M v = 0 (NaN path, zero centroid)
  int empty;
};

hipError_t launch_fill_int(int* p, long count, int value, hipStream_t st);
// Grouping by the key label * B + bin(dist) (see k_group_count): bins per
// cluster for n rows and k clusters, 1 without distances or for large k.
constexpr int kGroupBins = 64;
constexpr long kGroupMaxBlkKeys = 32L << 20;  // per-block key counts kept (ints)
int group_bins(long n, int k, bool have_dist);
hipError_t launch_group_count(const int* label, long n, int k, const float* dist, int B,
                              int* blk_count, int* minrow, int* bad, hipStream_t st);
hipError_t launch_group_scan(int nblk, int k, int B, const int* blk_count, int* blk_base,
                             int* kcounts, int* counts, long* offsets, long* koff,
                             hipStream_t st);
hipError_t launch_group_place(const int* label, long n, int k, const float* dist, int B,
                              const int* blk_base, const long* koff, int* perm, hipStream_t st);
// FP64 (sum, sum of squares) of every row.
hipError_t launch_row_stats(const float* x, long rows, int m, double2* stats, hipStream_t st);
// True when launch_align_pack reads per-row sums (launch_row_stats) for
// this m; otherwise `stats` may be NULL.
bool align_pack_uses_row_stats(int m);
hipError_t launch_align_pack(const float* x, int m, const int* perm, long n_pad,
                             const long* offsets, const int* minrow, int k, const int* shift,
                             const int* czero, const double2* stats, uint16_t* yhi,
                             uint16_t* ylo, float* a0, hipStream_t st);
// Single fit: blocks of k_syrk for ngroups groups (each group's waves on
// one XCD), the launch over the groups, and the reduction of the clusters'
// kSyrkSub x kSyrkSub partials into S.
long syrk_blocks(int m, int ngroups);
hipError_t launch_syrk(const SyrkGroup* groups, int ngroups, const uint16_t* yhi,
                       const uint16_t* ylo, int m, int split, float* part, hipStream_t st);
hipError_t launch_syrk_reduce(const SyrkCluster* clusters, int nclusters, const float* part, int m,
                              float* S, hipStream_t st);
hipError_t launch_power(const float* S, int m, int k, const int* counts, const float* a0,
                        uint64_t seed, int iter, int max_steps, double tol, const double* vin,
                        const double* svin, const double* rsum, const int* steps_in,
                        const int* conv_in, int eig_mode, const int* mask, int defer_fallback,
                        float* c_out, PowerInfo* info, hipStream_t st);

// Lanczos evaluation of the power iterate M^I v0 (one block per cluster).
// Lanczos evaluation of the power iterate (one step = two launches).
constexpr int kLanczosMax = 128;
struct LzState {
  double sigma;                   // 1^T S 1
  double alpha[kLanczosMax + 1];  // tridiagonal T: diagonal
  double beta[kLanczosMax + 1];   // off-diagonal, beta[l] = T[l][l-1]
  double y[kLanczosMax];          // T^I e1 / |T^I e1| at the last check
  int it;                         // index of the newest basis vector
  int len;                        // steps done
  int prev_len;                   // len at the last check
  int conv;                       // accepted
  int done;                       // finished (accepted, breakdown or cap)
};
// Fused single-launch form (one block per cluster, for small m).
constexpr int kFusedLanczosMaxM = 768;
// Acceptance of the Lanczos evaluation of M^I v0: y = T_j^I e1 is formed at
// j = kLzFirstCheck and every kLzCheckEvery steps after, and accepted when
// it moved by <= kLzPowerTol since the previous check.  For a geometric
// convergence rate rho per step the accepted y is then within about
// rho^4 / (1 - rho^4) * kLzPowerTol of the limit.  S itself, in FP32 from
// split-BF16 products, leaves the centroids about 2e-6 (relative L2) from
// the FP64 oracle's at n=100000, m=512, k=16; 1e-7 leaves that and the SBD
// error unchanged (1e-9 needed 28 steps there, 1e-7 needs 24).
constexpr int kLzFirstCheck = 12;
constexpr int kLzCheckEvery = 4;
constexpr double kLzPowerTol = 1e-7;
__host__ __device__ constexpr bool lz_power_check(int len) {
  return len >= kLzFirstCheck && len % kLzCheckEvery == 0;
}

// Lanczos with G blocks per cluster in one launch (members exchange through
// device-scope memory and wait for each other, so all k*G blocks must be
// resident at once).  Used from m = kLzCoopMinM on, when the plan fits.
constexpr int kLzCoopMinM = 256;
constexpr int kLzCoopMaxR = 64;   // rows of S per member
constexpr int kLzCoopMaxG = 64;   // members per cluster
constexpr int kLzCoopMaxM = kLzCoopMaxR * kLzCoopMaxG;
constexpr size_t kLzCoopSLdsBytes = 96 * 1024;  // S rows kept in LDS up to this size
constexpr size_t kLzCoopLdsMax = 160 * 1024;    // LDS of one CU
struct LzCoopPlan {
  int G = 0, R = 0, nb = 0;  // members per cluster, rows per member, blocks
  bool s_lds = false;
  size_t lds = 0;            // dynamic LDS bytes per block
  long xch_len = 0;          // exchange words per cluster
};
// False when m, k or the device do not allow every member to be resident.
bool plan_lz_coop(int m, int k, int cus, LzCoopPlan* p);
// xch: k * p.xch_len words and fails: k words, zero when first used; gen:
// launch generation, 1 <= gen < kLzCoopMaxGen, different from every earlier
// launch on the same buffers since they were zeroed.  timeout: wall-clock
// ticks a member waits for the others.
constexpr uint32_t kLzCoopMaxGen = 1u << 22;
hipError_t launch_lz_coop(const LzCoopPlan& p, const float* S, int m, int k, const int* counts,
                          uint64_t seed, int iter, int I, double lz_tol, uint64_t* xch,
                          uint32_t* fails, uint32_t gen, long long timeout, double* vout,
                          double* rout, int* steps_out, int* conv_out, hipStream_t st);
hipError_t launch_power_lanczos(const float* S, int m, int k, const int* counts, uint64_t seed,
                                int iter, int I, double lz_tol, double* Q, double* vout,
                                double* rout, int* steps_out, int* conv_out, hipStream_t st);
hipError_t launch_cluster_matvec(const float* S, int m, int k, const double* x, long xstride,
                                 double* y, const LzState* state, hipStream_t st);
// mask (may be NULL): clusters to run; others are left done.
hipError_t launch_lz_setup(int m, int k, const int* counts, const int* mask, uint64_t seed,
                           int iter, const double* rbuf, double* Q, LzState* state,
                           hipStream_t st);
// mode 0: power iterate M^I v0; mode 1: top eigenvector of M (tol on the
// Ritz residual relative to the Ritz value).
hipError_t launch_lz_step(int m, int k, int mode, int I, double lz_tol, const double* rbuf,
                          const double* wbuf, double* Q, LzState* state, hipStream_t st);
hipError_t launch_lz_finish(int m, int k, const int* counts, const double* Q,
                            const LzState* state, double* vout, int* steps_out, int* conv_out,
                            hipStream_t st);
hipError_t configure_extract_kernels();

// ----------------------------------------------------- grouped launches
// One launch runs a kernel's body for many independent jobs (the fits of a
// batch): job j owns the blocks [first[j], first[j+1]) of the grid and runs
// them as blocks 0 .. first[j+1] - first[j] - 1 of a launch of its own, so a
// job's results do not depend on the other jobs of the launch.  `jobs` and
// `first` are device arrays (njobs entries, njobs + 1 prefix values).
struct GroupTable {
  const void* jobs = nullptr;  // device array of the kernel's job struct
  const int* first = nullptr;  // device prefix of blocks per job
  int njobs = 0;
  int blocks = 0;              // first[njobs]
  size_t lds = 0;              // dynamic LDS of the launch (largest job's)
};

struct FillSeg {  // p[0 .. count) = value
  int* p;
  long count;
  int value;
};
hipError_t launch_fill_segs(const FillSeg* segs, int nseg, long max_count, hipStream_t st);

struct ZnormJob {
  float* x;
  long rows;
  int m, ddof;
};
// Rows of every job z-normalized (jobs with m % 4 == 0 and 16-byte aligned
// rows in groups by vector count nv: 0 = the scalar form, else 2, 4, 8, 16).
hipError_t launch_znorm_group(const GroupTable& g, int nv, hipStream_t st);
int znorm_vectors(const float* x, int m);  // nv of launch_znorm for these rows
long znorm_blocks(long rows);

struct SeriesSpecJob {
  const float* x;
  long n;
  int m;
  float2* xs;
  float* xn;
  uint16_t* xm;
};
long series_spec_blocks(int F, long n);
hipError_t launch_series_spec_group(int F, const GroupTable& g, const float2* tw, hipStream_t st);

struct SpectraJob {
  const float* y;
  int rows, m;
  float2* spec;
  float* norms;
  int* zero;
  float* mag;
  uint16_t* bmag;
};
long spectra_blocks(int F, long rows);
hipError_t launch_spectra_group(int F, const GroupTable& g, const float2* tw, hipStream_t st);

struct BoundsJob {
  const uint16_t* xm;
  long n;
  const uint16_t* cm;
  int k, ld;
  float* ub;
};
// Bound-product kernel form for (F, k): 0..4; blocks of a job of n rows.
int bounds_form(int F, int k);
long bounds_blocks(int form, long n);
hipError_t launch_bounds_group(int F, int form, const GroupTable& g, hipStream_t st);

struct PruneJob {
  const float* x;
  const float2* xs;
  const float* xn;
  const float* ub;
  int ub_ld;
  long n, npairs;
  int m, k;
  const float2* cspec;
  const float* cmag;
  const float* cnorm;
  int rev;
  const int* order;
  int* label;
  float* dist;
  int* shift;
  int* changed;
  unsigned long long* evals;
  const long* npad;  // may be NULL; else npairs = *npad / 2, read on the device
};
long prune_blocks(int F, long npairs, int k, long n);  // grid cap tuned per (F, k, n)
size_t prune_lds(int F, int k);
hipError_t launch_assign_prune_group(int F, const GroupTable& g, const float2* tw, hipStream_t st);

struct GroupCountJob {
  const int* label;
  long n;
  int k;
  const float* dist;
  int B;
  int* blk_count;
  int* minrow;
  int lds_min;
  int* bad;
};
size_t group_count_lds(int k, int B, int* lds_min);
hipError_t launch_group_count_group(const GroupTable& g, hipStream_t st);

struct GroupScanJob {
  int nrowblk, K;
  const int* blk_count;
  int* blk_base;
  int* counts;
};
hipError_t launch_group_scan_group(const GroupTable& g, hipStream_t st);

struct GroupOffsetsJob {
  int k, B;
  const int* kcounts;
  int* counts;
  long* offsets;
  long* koff;
};
hipError_t launch_group_offsets_group(const GroupTable& g, hipStream_t st);

struct GroupPlaceJob {
  const int* label;
  long n;
  int k;
  const float* dist;
  int B;
  const int* blk_base;
  const long* koff;
  int* perm;
};
size_t group_place_lds(int k, int B);
hipError_t launch_group_place_group(const GroupTable& g, hipStream_t st);

struct RowStatsJob {
  const float* x;
  long rows;
  int m;
  double2* stats;
};
hipError_t launch_row_stats_group(const GroupTable& g, hipStream_t st);

struct AlignPackJob {
  const float* x;
  int m;
  const int* perm;
  long n_pad;
  const long* offsets;
  const int* minrow;
  int k;
  const int* shift;
  const int* czero;
  const double2* stats;
  uint16_t* yhi;
  uint16_t* ylo;
  float* a0;
};
// Align + pack form for m: R = 16, 32, 64 (rows staged in LDS) or 0 (from
// HBM with per-row sums, m > 1024); columns per block; dynamic LDS bytes.
int align_pack_form(int m);
long align_pack_cols(int form);
size_t align_pack_lds(int form, int m);
hipError_t launch_align_pack_group(int form, const GroupTable& g, hipStream_t st);

// S = Y^T Y of many jobs: one item (tile) per block, the job's panels,
// length and partials from syrk_jobs[job].
struct SyrkJob {
  const uint16_t* yhi;
  const uint16_t* ylo;
  int m;
  float* part;
  float* S;
};
struct GroupSyrkItem {
  int job;
  SyrkItem it;
};
struct GroupSyrkTile {
  int job;
  SyrkTile t;
};
hipError_t launch_syrk_group(const SyrkJob* jobs, const GroupSyrkItem* items, int nitems,
                             int split, hipStream_t st);
hipError_t launch_syrk_reduce_group(const SyrkJob* jobs, const GroupSyrkTile* tiles,
                                    int ntiles, hipStream_t st);

struct PowerLanczosJob {
  const float* S;
  int m;
  const int* counts;
  uint64_t seed;
  int iter, I;
  double* Q;
  double* vout;
  double* rout;
  int* steps_out;
  int* conv_out;
};
size_t power_lanczos_lds(int m);
hipError_t launch_power_lanczos_group(const GroupTable& g, double lz_tol, hipStream_t st);

// Gram form of thin clusters (segment nd < m, nd <= kGramMaxND, m <=
// kGramMaxM): G = Yc Yc^T in 64 x 64 tiles (the diagonal ones also form
// the start u0 = Yc v0), the Lanczos stage on G, and the map back to
// v = Yc^T g, M v = Yc^T h.  One item or job per block.
constexpr int kGramMaxM = 4096;
constexpr int kGramMaxND = 2048;
struct GramItem {
  const uint16_t* yhi;
  const uint16_t* ylo;
  float* G;  // n x n, row stride n
  long p0;   // first packed column of the cluster (multiple of 64)
  int m, nd, n, ti, tj;  // segment nd (multiple of 64), members n (<= nd)
  // diagonal tiles also form the start u0 = Yc v0 of their 64 members
  // (u0 NULL: not this tile), v0 = power_v0(seed, iter, cluster, .)
  double* u0;
  uint64_t seed;
  int iter, cluster;
};
struct GramVecJob {  // the map back of one thin cluster
  const uint16_t* yhi;
  const uint16_t* ylo;
  long p0;
  int m, nd, n;
  const double* g;    // g (n)
  const double* h;    // h (n; may be sv: read before sv is written)
  double* v;          // m
  double* sv;         // M v (m)
  double* r;          // zeros (m)
};
struct GramLzJob {
  const float* G;
  int nd, I;
  const double* u0;
  double* Q;
  double* g;
  double* h;
  int* steps;
  int* conv;
};
hipError_t launch_gram_tiles(const GramItem* items, int nitems, hipStream_t st);
size_t gram_map_lds(int m, int nd);
size_t gram_lanczos_lds(int nd);
hipError_t launch_gram_lanczos(const GramLzJob* jobs, int njobs, size_t lds, double lz_tol,
                               hipStream_t st);
hipError_t launch_gram_map(const GramVecJob* jobs, int njobs, size_t lds, hipStream_t st);

// One wave per cluster of dimension d <= kLzWaveMaxD: the Lanczos
// evaluation of the power iterate on the cluster's S (S form: centring,
// start power_v0(seed, iter, cluster, .), r = S 1 written) or on its Gram
// matrix (start vstart, no centring, power steps when the check never
// passes); v = the iterate's direction, h = S v (the matrix given).
constexpr int kLzWaveMaxD = 256;
struct LzWaveJob {
  const float* S;  // d x d, row stride d
  int d, gram, I;
  uint64_t seed;
  int iter, cluster;
  const double* vstart;  // Gram form
  double* Q;             // (kLanczosMax + 1) x d
  double* v;
  double* r;             // S form
  double* h;
  int* steps;
  int* conv;
};
// waves_per_cluster 1: four clusters per block; 4: one cluster per block,
// each matvec shared by its four waves (for the larger orders)
hipError_t launch_lanczos_wave(const LzWaveJob* jobs, int njobs, int waves_per_cluster,
                               double lz_tol, hipStream_t st);
constexpr int kLzWaveSplitD = 128;  // orders above this take four waves

struct ClusterMatvecJob {
  const float* S;
  int m, k;
  const double* x;
  long xstride;
  double* y;
  const LzState* state;
};
long cluster_matvec_blocks(int m, int k);
hipError_t launch_cluster_matvec_group(const GroupTable& g, hipStream_t st);

struct LzJob {  // k_lz_setup / k_lz_step / k_lz_finish of one job
  int m;
  const int* counts;
  const int* mask;
  uint64_t seed;
  int iter, I;
  const double* rbuf;
  const double* wbuf;
  double* Q;
  LzState* state;
  double* vout;
  int* steps_out;
  int* conv_out;
  int* done_out;  // may be NULL: per-cluster copy of state.done (one readback)
};
size_t lz_step_lds(int m);
hipError_t launch_lz_setup_group(const GroupTable& g, hipStream_t st);
hipError_t launch_lz_step_group(const GroupTable& g, int mode, double lz_tol, hipStream_t st);
hipError_t launch_lz_finish_group(const GroupTable& g, hipStream_t st);

struct PowerJob {
  const float* S;
  int m;
  const int* counts;
  const float* a0;
  uint64_t seed;
  int iter, max_steps;
  double tol;
  const double* vin;
  const double* svin;
  const double* rsum;
  const int* steps_in;
  const int* conv_in;
  int eig_mode;
  const int* mask;
  int defer_fallback;
  float* c_out;
  PowerInfo* info;
};
hipError_t launch_power_group(const GroupTable& g, hipStream_t st);
size_t power_lds(int m);

}  // namespace tcx

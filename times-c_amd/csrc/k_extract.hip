// Shape-extraction kernels: stable grouping of rows by label, shift-align +
// z-normalize + BF16 (hi, lo) pack into a transposed panel, S_j = Y_j^T Y_j
// on MFMA with FP32 accumulation, and the per-cluster power iteration with
// rank-1 centering, FP64 Rayleigh-quotient test, sign fix and final z-norm.
// Semantics: _preSBD_extract_shape + _extract_shape,
// Code/test/test_timesC.py:127-204 (grouping order :228-243).
#include <cfloat>
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <cmath>

#include "tcx_device.hpp"
#include "tcx_internal.hpp"
#include "tcx_kernels.hpp"
#include "tcx_rng.hpp"

namespace tcx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// --------------------------------------------------------------- helpers
__global__ void k_fill_int(int* p, long count, int value) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long)gridDim.x * blockDim.x)
    p[i] = value;
}

// -------------------------------------------------------------- grouping
// Rows are grouped by the key label * B + bin, bin = the row's previous
// SBD distance (in [0, 2]) in B equal bins (B = 1 without distances).
// Clusters stay contiguous and ordered; inside a cluster rows run by bin,
// then by index, so the assignment pairs series of similar distance.  The
// grouping is stable and deterministic; the cluster's lowest row index
// (the reference's first member, for the sign fix) is tracked separately.
__device__ __forceinline__ int group_key(int l, long row, const float* dist, int B) {
  if (B == 1) return l;
  const float d = dist[row];
  int b = 0;
  if (d == d) {  // not NaN
    const float f = d * (0.5f * (float)B);
    b = f >= (float)(B - 1) ? B - 1 : (f > 0.f ? (int)f : 0);
  }
  return l * B + b;
}

// Per block of kGroupChunk rows: histogram of the K = k * B keys and the
// lowest row of each cluster (atomicMin into minrow, INT_MAX when empty).
__device__ __forceinline__ void group_count_body(const int bid, const int nblk,
    const int* __restrict__ label, long n,
                                                     int k, const float* __restrict__ dist, int B,
                                                     int* __restrict__ blk_count,
                                                     int* __restrict__ minrow, int lds_min,
                                                     int* __restrict__ bad) {
  extern __shared__ int cnt[];  // K counts, then (lds_min) k minima
  const int K = k * B;
  int* mn = cnt + K;
  for (int j = threadIdx.x; j < K; j += blockDim.x) cnt[j] = 0;
  if (lds_min)
    for (int j = threadIdx.x; j < k; j += blockDim.x) mn[j] = 0x7fffffff;
  __syncthreads();
  const long b0 = (long)bid * kGroupChunk;
  const long b1 = b0 + kGroupChunk < n ? b0 + kGroupChunk : n;
  for (long i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
    const int l = label[i];
    if (l >= 0 && l < k) {
      atomicAdd(&cnt[group_key(l, i, dist, B)], 1);
      if (lds_min)
        atomicMin(&mn[l], (int)i);
      else
        atomicMin(&minrow[l], (int)i);
    } else {
      *bad = 1;  // out-of-range label: the row is left out, the host reports it
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < K; j += blockDim.x) blk_count[(long)bid * K + j] = cnt[j];
  if (lds_min)
    for (int j = threadIdx.x; j < k; j += blockDim.x)
      if (mn[j] != 0x7fffffff) atomicMin(&minrow[j], mn[j]);
}

__global__ __launch_bounds__(256) void k_group_count(const int* __restrict__ label, long n,
                                                     int k, const float* __restrict__ dist, int B,
                                                     int* __restrict__ blk_count,
                                                     int* __restrict__ minrow, int lds_min,
                                                     int* __restrict__ bad) {
  group_count_body(blockIdx.x, gridDim.x, label, n, k, dist, B, blk_count, minrow, lds_min, bad);
}

// Exclusive scan over blocks for one key per block (blockIdx.x = key):
// blk_base[b][key] = sum of blk_count[b'][key] for b' < b; counts[key] =
// total.  Each of the 256 threads scans a contiguous run of blocks, then the
// run totals are scanned in LDS.
__device__ __forceinline__ void group_scan_body(const int bid, const int nblk,
    int nrowblk, int k,
                                                    const int* __restrict__ blk_count,
                                                    int* __restrict__ blk_base,
                                                    int* __restrict__ counts) {
  __shared__ int tot[256];
  const int j = bid, t = threadIdx.x;
  const int per = (nrowblk + 255) / 256;
  const int b0 = t * per, b1 = b0 + per < nrowblk ? b0 + per : nrowblk;
  int run = 0;
  for (int b = b0; b < b1; ++b) run += blk_count[(long)b * k + j];
  tot[t] = run;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {  // inclusive Hillis-Steele scan
    const int v = t >= o ? tot[t - o] : 0;
    __syncthreads();
    tot[t] += v;
    __syncthreads();
  }
  run = t > 0 ? tot[t - 1] : 0;
  for (int b = b0; b < b1; ++b) {
    blk_base[(long)b * k + j] = run;
    run += blk_count[(long)b * k + j];
  }
  if (t == 255) counts[j] = tot[255];
}

__global__ __launch_bounds__(256) void k_group_scan(int nrowblk, int k,
                                                    const int* __restrict__ blk_count,
                                                    int* __restrict__ blk_base,
                                                    int* __restrict__ counts) {
  group_scan_body(blockIdx.x, gridDim.x, nrowblk, k, blk_count, blk_base, counts);
}

// One block of 1024 threads: cluster sizes from the key counts, padded
// cluster offsets (each cluster's columns start at a multiple of kPackPad;
// the k sizes are scanned 1024 at a time) and the offset of every key
// inside its cluster.  Phase 1 runs one wave per cluster (lane b holds bin
// b, B <= 64) and keeps the in-cluster exclusive prefix in koff; phase 3
// adds the cluster offsets.
__device__ __forceinline__ void group_offsets_body(int k, int B,
                                                   const int* __restrict__ kcounts,
                                                   int* __restrict__ counts,
                                                   long* __restrict__ offsets,
                                                   long* __restrict__ koff);

__global__ __launch_bounds__(1024) void k_group_offsets(int k, int B,
                                                        const int* __restrict__ kcounts,
                                                        int* __restrict__ counts,
                                                        long* __restrict__ offsets,
                                                        long* __restrict__ koff) {
  group_offsets_body(k, B, kcounts, counts, offsets, koff);
}

__device__ __forceinline__ void group_offsets_body(int k, int B,
                                                   const int* __restrict__ kcounts,
                                                   int* __restrict__ counts,
                                                   long* __restrict__ offsets,
                                                   long* __restrict__ koff) {
  __shared__ long tot[1024];
  __shared__ long carry;
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t >> 6;
  for (int j = w; j < k; j += 1024 / kWave) {
    const int c = lane < B ? kcounts[j * B + lane] : 0;
    int inc = c;  // inclusive scan over the lanes
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const int u = __shfl_up(inc, o);
      if (lane >= o) inc += u;
    }
    if (lane < B) koff[j * B + lane] = inc - c;
    if (lane == kWave - 1) counts[j] = inc;
  }
  if (t == 0) {
    carry = 0;
    offsets[0] = 0;
  }
  __syncthreads();
  for (int j0 = 0; j0 < k; j0 += 1024) {
    const int j = j0 + t;
    const long c = j < k ? counts[j] : 0;
    tot[t] = (c + kPackPad - 1) / kPackPad * kPackPad;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const long v = t >= o ? tot[t - o] : 0;
      __syncthreads();
      tot[t] += v;
      __syncthreads();
    }
    if (j < k) offsets[j + 1] = carry + tot[t];
    __syncthreads();
    if (t == 0) carry += tot[1023];
    __syncthreads();
  }
  for (int e = t; e < k * B; e += 1024) koff[e] += offsets[e / B];
}

// Stable placement: perm[koff[key] + rank] = row, rank = number of earlier
// rows with the same key.  One wave per block of rows; ranks inside a
// 64-row slice come from ballots over the distinct keys of the slice.
// (The fallback for K > kPlaceParMaxK: B is 1 there, the key the label.)
__global__ __launch_bounds__(64) void k_group_place(const int* __restrict__ label, long n,
                                                    int k, const int* __restrict__ blk_base,
                                                    const long* __restrict__ koff,
                                                    int* __restrict__ perm) {
  extern __shared__ int run[];
  const int lane = threadIdx.x;
  for (int j = lane; j < k; j += kWave) run[j] = blk_base[(long)blockIdx.x * k + j];
  wave_lds_sync();
  const long b0 = (long)blockIdx.x * kGroupChunk;
  const long b1 = b0 + kGroupChunk < n ? b0 + kGroupChunk : n;
  const uint64_t below = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  for (long base = b0; base < b1; base += kWave) {
    const long row = base + lane;
    int lab = row < b1 ? label[row] : -1;
    if (lab >= k) lab = -1;
    bool todo = lab >= 0;
    for (;;) {
      uint64_t pending = __ballot(todo);
      if (pending == 0) break;
      const int leader = __ffsll((unsigned long long)pending) - 1;
      const int j = __shfl(lab, leader);
      const uint64_t mask = __ballot(lab == j);
      if (lab == j) {
        const int rank = run[j] + __popcll(mask & below);
        perm[koff[j] + rank] = (int)row;
        todo = false;
      }
      wave_lds_sync();
      if (lane == leader) run[j] += __popcll(mask);
      wave_lds_sync();
    }
  }
}

// The same placement for keys label * B + bin, with the kGroupChunk / 64
// slices of a block handled by one wave each at once.  A lane's peers (the
// lanes of its slice with the same key) come from one ballot per key bit;
// the lowest peer writes the slice's count of that key, the per-slice
// counts are scanned across waves in LDS, then every lane places its row
// at its key's base plus the number of lower peers.  LDS: 8 * K ints.
constexpr int kPlaceWaves = kGroupChunk / kWave;
constexpr int kPlaceParMaxK = 2048;  // 8 * K * 4 bytes <= 64 KB

__device__ __forceinline__ void group_place_par_body(const int bid, const int nblk,
    const int* __restrict__ label, long n, int k, const float* __restrict__ dist, int B,
    const int* __restrict__ blk_base, const long* __restrict__ koff, int* __restrict__ perm) {
  extern __shared__ int cnt[];  // [kPlaceWaves][K]: counts, then bases
  const int K = k * B;
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x >> 6;
  int* mine = cnt + w * K;
  for (int j = lane; j < K; j += kWave) mine[j] = 0;
  const long row = (long)bid * kGroupChunk + threadIdx.x;
  int key = -1;
  if (row < n) {
    const int l = label[row];
    if (l >= 0 && l < k) key = group_key(l, row, dist, B);
  }
  const uint64_t below = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  uint64_t peers = __ballot(key >= 0);
  const int nbits = 32 - __clz(K - 1 > 0 ? K - 1 : 1);
  for (int b = 0; b < nbits; ++b) {
    const bool bit = (key >> b) & 1;
    const uint64_t set = __ballot(bit);
    peers &= bit ? set : ~set;
  }
  wave_lds_sync();
  // 1. slice counts, written by the lowest lane of each key
  if (key >= 0 && (peers & below) == 0) mine[key] = __popcll(peers);
  __syncthreads();
  // 2. exclusive scan over the slices, from the block's base
  for (int j = threadIdx.x; j < K; j += kGroupChunk) {
    int run = blk_base[(long)bid * K + j];
#pragma unroll
    for (int v = 0; v < kPlaceWaves; ++v) {
      const int c = cnt[v * K + j];
      cnt[v * K + j] = run;
      run += c;
    }
  }
  __syncthreads();
  // 3. placement
  if (key >= 0) perm[koff[key] + mine[key] + __popcll(peers & below)] = (int)row;
}

__global__ __launch_bounds__(kGroupChunk) void k_group_place_par(
    const int* __restrict__ label, long n, int k, const float* __restrict__ dist, int B,
    const int* __restrict__ blk_base, const long* __restrict__ koff, int* __restrict__ perm) {
  group_place_par_body(blockIdx.x, gridDim.x, label, n, k, dist, B, blk_base, koff, perm);
}

// ------------------------------------------------------------ align + pack
__device__ __forceinline__ int cluster_of(long p, const long* offsets, int k) {
  int lo = 0, hi = k;  // offsets[lo] <= p < offsets[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (offsets[mid] <= p) lo = mid; else hi = mid;
  }
  return lo;
}

// FP64 sum and sum of squares of every row of x (one wave per row).  The
// shape extraction derives each aligned member's mean and variance from
// these and the few values its shift drops, so it reads x once per refine.
__device__ __forceinline__ void row_stats_body(const int bid, const int nblk,
    const float* __restrict__ x, long rows, int m,
                                                   double2* __restrict__ stats) {
  const int lane = threadIdx.x & (kWave - 1);
  const long row = (long)bid * 4 + (threadIdx.x >> 6);
  if (row >= rows) return;
  const float* r = x + row * (long)m;
  double s = 0.0, q = 0.0;
  for (int t = lane; t < m; t += kWave) {
    const double v = (double)r[t];
    s += v;
    q += v * v;
  }
  s = wave_sum_f64(s);
  q = wave_sum_f64(q);
  if (lane == 0) stats[row] = make_double2(s, q);
}

__global__ __launch_bounds__(256) void k_row_stats(const float* __restrict__ x, long rows, int m,
                                                   double2* __restrict__ stats) {
  row_stats_body(blockIdx.x, gridDim.x, x, rows, m, stats);
}

// Block of 64 packed columns.  For each column (a member of cluster cl):
// a = roll_zeropad(x_row, s) with s = shift[row] unless the old centroid of
// cl is all zero (test_timesC.py:128-132); y = zscore(a, ddof=1) with FP64
// statistics (:142); y -> (hi, lo) BF16 with y ~= hi + lo, written
// transposed: Y[t][p].  The first column of a cluster (its lowest-index
// member) also writes its aligned row to a0 for the sign fix (:163).
// The statistics of a come from the row's sum and sum of squares (stats)
// minus the |s| values the shift drops (or, when most values are dropped,
// from the kept values); where the one-pass form would lose precision
// (variance below 1e-8 of the mean square, or the kept sum of squares below
// 1e-8 of the row's) the row is summed again in two passes.  x is then read
// once, to pack.
__device__ __forceinline__ void align_pack_body(const int bid, const int nblk,
    const float* __restrict__ x, int m, const int* __restrict__ perm, long n_pad,
    const long* __restrict__ offsets, const int* __restrict__ minrow, int k,
    const int* __restrict__ shift,
    const int* __restrict__ czero, const double2* __restrict__ stats,
    uint16_t* __restrict__ yhi, uint16_t* __restrict__ ylo, float* __restrict__ a0) {
  __shared__ double s_mean[64], s_isd[64];
  __shared__ int s_row[64], s_sh[64], s_first[64];
  __shared__ __align__(16) uint16_t t_hi[64][72], t_lo[64][72];  // 144-byte rows
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const long p0 = (long)bid * 64;
  if (p0 >= offsets[k]) return;  // past the padded end (grid sized for the maximum)
  // each column's row, shift and (for a cluster's first member) cluster,
  // resolved by the 64 lanes of wave 0 at once
  if (wid == 0) {
    const long p = p0 + lane;
    const int row = p < n_pad ? perm[p] : -1;
    int sh = 0, first = -1;
    if (row >= 0) {
      const int cl = cluster_of(p, offsets, k);
      sh = czero[cl] ? 0 : shift[row];
      if (row == minrow[cl]) first = cl;
    }
    s_row[lane] = row;
    s_sh[lane] = sh;
    s_first[lane] = first;
  }
  __syncthreads();
  // statistics of 16 columns per wave, 4 at a time so that their loads
  // overlap
  for (int q0 = wid * 16; q0 < wid * 16 + 16; q0 += 4) {
    double ps[4], pq[4], tq[4];
    double2 tot[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int row = s_row[q0 + u], sh = s_sh[q0 + u];
      const float* xr = x + (long)(row >= 0 ? row : 0) * m;
      // kept source range [k0, k1) of roll_zeropad(x, sh); sum the kept
      // values when they are few, else the dropped ones
      const int k0 = sh > 0 ? 0 : (-sh < m ? -sh : m);
      const int k1 = sh > 0 ? (sh < m ? m - sh : 0) : m;
      const bool direct = 2 * (k1 - k0) <= m;
      const int a1 = direct ? k0 : 0, n1 = direct ? k1 - k0 : k0, n2 = direct ? 0 : m - k1;
      const int cnt = row >= 0 ? n1 + n2 : 0;
      double a = 0.0, b = 0.0;
      for (int i = lane; i < cnt; i += kWave) {
        const double v = (double)xr[i < n1 ? a1 + i : k1 + (i - n1)];
        a += v;
        b += v * v;
      }
      ps[u] = a;
      pq[u] = b;
      tot[u] = row >= 0 ? stats[row] : make_double2(0.0, 0.0);
      tq[u] = direct ? 0.0 : 1.0;  // 1: subtract from the row's totals
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = q0 + u;
      const int row = s_row[q], sh = s_sh[q];
      const double a = wave_sum_f64(ps[u]), b = wave_sum_f64(pq[u]);
      const bool sub = tq[u] != 0.0;
      const double sum = sub ? tot[u].x - a : a;
      const double sq = sub ? tot[u].y - b : b;
      const double tsq = sub ? tot[u].y : b;
      double mean = sum / (double)m;
      double ss = sq - (double)m * mean * mean;
      if (row >= 0 && !(ss > 1e-8 * sq && sq > 1e-8 * tsq)) {
        // ill-conditioned, constant or empty aligned row: two passes
        const float* xr = x + (long)row * m;
        double s1 = 0.0;
        for (int t = lane; t < m; t += kWave) {
          const int src = t - sh;
          s1 += (src >= 0 && src < m) ? (double)xr[src] : 0.0;
        }
        mean = wave_sum_f64(s1) / (double)m;
        double s2 = 0.0;
        for (int t = lane; t < m; t += kWave) {
          const int src = t - sh;
          const double d = ((src >= 0 && src < m) ? (double)xr[src] : 0.0) - mean;
          s2 += d * d;
        }
        ss = wave_sum_f64(s2);
      }
      if (lane == 0) {
        s_mean[q] = mean;
        s_isd[q] = 1.0 / sqrt(ss / (double)(m - 1));
      }
      const int first = s_first[q];
      if (first >= 0) {  // the cluster's lowest-index member: its aligned row
        const float* xr = x + (long)row * m;
        float* o = a0 + (long)first * m;
        for (int t = lane; t < m; t += kWave) {
          const int src = t - sh;
          o[t] = (src >= 0 && src < m) ? xr[src] : 0.f;
        }
      }
    }
  }
  __syncthreads();
  const int tm = panel_blocks(m);
  for (int t0 = 0; t0 < tm * 32; t0 += 64) {
#pragma unroll 4
    for (int it = 0; it < 16; ++it) {
      const int idx = it * 256 + tid;
      const int q = idx >> 6, tt = idx & 63;
      const int t = t0 + tt;
      const int row = s_row[q];
      float f = 0.f;
      if (row >= 0 && t < m) {
        const int src = t - s_sh[q];
        const double a = (src >= 0 && src < m) ? (double)x[(long)row * m + src] : 0.0;
        const double y = (a - s_mean[q]) * s_isd[q];
        f = isnan(y) ? 0.f : (isinf(y) ? (y > 0 ? 3.402823466e38f : -3.402823466e38f) : (float)y);
      }
      const uint16_t h = f2bf_rne(f);
      t_hi[tt][q] = h;
      t_lo[tt][q] = f2bf_rne(f - bf2f(h));
    }
    __syncthreads();
    // eight 1 KB fragment chunks (4 column chunks x 2 row chunks) per array,
    // one 16-byte store per lane (tm is a multiple of 8: both row chunks exist)
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
      const int sidx = pass * 256 + tid;
      const int ch = sidx >> 6, ln = sidx & 63;
      const int cc = ch >> 1, rr = ch & 1;
      const int r32 = (t0 >> 5) + rr;
      const int tt = rr * 32 + (ln & 31), q = cc * 16 + 8 * (ln >> 5);
      const uint4 h = *reinterpret_cast<const uint4*>(&t_hi[tt][q]);
      const uint4 l = *reinterpret_cast<const uint4*>(&t_lo[tt][q]);
      const long off = (((p0 >> 4) + cc) * (long)tm + r32) * 512 + ln * 8;
      *reinterpret_cast<uint4*>(yhi + off) = h;
      if (ylo) *reinterpret_cast<uint4*>(ylo + off) = l;  // NULL: BF16 S (hi only)
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_align_pack(
    const float* __restrict__ x, int m, const int* __restrict__ perm, long n_pad,
    const long* __restrict__ offsets, const int* __restrict__ minrow, int k,
    const int* __restrict__ shift,
    const int* __restrict__ czero, const double2* __restrict__ stats,
    uint16_t* __restrict__ yhi, uint16_t* __restrict__ ylo, float* __restrict__ a0) {
  align_pack_body(blockIdx.x, gridDim.x, x, m, perm, n_pad, offsets, minrow, k, shift, czero, stats, yhi, ylo, a0);
}

// The same result as k_align_pack for m <= kPackLdsMaxM, from R whole rows
// staged in LDS (R = 16, 32 or 64 packed columns per block, R * (m + 4)
// floats): every gathered row is read from HBM once, the statistics of the
// aligned row a = roll_zeropad(x_row, s) are two-pass FP64 sums over LDS
// (mean, then the centred sum of squares with the |s| zeros included), and
// the panel chunks are formed from LDS.
constexpr int kPackLdsMaxM = 1024;
constexpr int kPackLdsBytes = 36 * 1024;     // R = 64 or 32 within this: four blocks per CU
constexpr int kPackLdsMaxBytes = 72 * 1024;  // R = 16 up to m = kPackLdsMaxM

__host__ __device__ constexpr int pack_lds_rows(int m) {
  return (m + 4) * 64 * 4 <= kPackLdsBytes ? 64 : ((m + 4) * 32 * 4 <= kPackLdsBytes ? 32 : 16);
}

template <int R>
__device__ __forceinline__ void align_pack_lds_body(const int bid, const int nblk,
    const float* __restrict__ x, int m, const int* __restrict__ perm,
    const long* __restrict__ offsets, const int* __restrict__ minrow, int k,
    const int* __restrict__ shift,
    const int* __restrict__ czero, uint16_t* __restrict__ yhi, uint16_t* __restrict__ ylo,
    float* __restrict__ a0) {
  extern __shared__ __align__(16) float rows_s[];  // R rows of stride ms
  __shared__ double s_mean[64], s_isd[64];
  __shared__ int s_row[64], s_sh[64];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int ms = m + 4;
  const long p0 = (long)bid * R;
  if (p0 >= offsets[k]) return;  // past the padded end (grid sized for the maximum)
  int first = -1;  // lane q < R of wave 0: cluster whose lowest member is column q
  if (wid == 0 && lane < R) {
    const long p = p0 + lane;
    const int row = perm[p];
    int sh = 0;
    if (row >= 0) {
      const int cl = cluster_of(p, offsets, k);
      sh = czero[cl] ? 0 : shift[row];
      if (row == minrow[cl]) first = cl;
    }
    s_row[lane] = row;
    s_sh[lane] = sh;
  }
  __syncthreads();
  // stage the R source rows (unshifted; absent rows as zeros): wave w holds
  // rows w, w+4, ...; the R/4 loads of a column step are issued together
  // (holes read row 0 and are zeroed)
  constexpr int RPW = R / 4;
  if ((m & 3) == 0) {
    const int nv = m >> 2;
    for (int v0 = 0; v0 < nv; v0 += kWave) {
      const int v = v0 + lane < nv ? v0 + lane : nv - 1;
      float4 val[RPW];
#pragma unroll
      for (int r = 0; r < RPW; ++r)
        val[r] = reinterpret_cast<const float4*>(x + (long)max(s_row[wid + 4 * r], 0) * m)[v];
      if (v0 + lane < nv) {
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
          const int q = wid + 4 * r;
          *reinterpret_cast<float4*>(rows_s + q * ms + 4 * v) =
              s_row[q] >= 0 ? val[r] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
      }
    }
  } else {
    for (int t0 = 0; t0 < m; t0 += kWave) {
      const int t = t0 + lane < m ? t0 + lane : m - 1;
      float val[RPW];
#pragma unroll
      for (int r = 0; r < RPW; ++r) val[r] = x[(long)max(s_row[wid + 4 * r], 0) * m + t];
      if (t0 + lane < m) {
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
          const int q = wid + 4 * r;
          rows_s[q * ms + t] = s_row[q] >= 0 ? val[r] : 0.f;
        }
      }
    }
  }
  // two-pass statistics of the aligned rows, one wave per row
  for (int q = wid; q < R; q += 4) {
    const int sh = s_sh[q];
    const float* r = rows_s + q * ms;
    const int src0 = sh > 0 ? 0 : (-sh < m ? -sh : m);    // kept source range
    const int src1 = sh > 0 ? (sh < m ? m - sh : 0) : m;
    double a = 0.0;
    for (int i = src0 + lane; i < src1; i += kWave) a += (double)r[i];
    const double mean = wave_sum_f64(a) / (double)m;
    double b = 0.0;
    for (int i = src0 + lane; i < src1; i += kWave) {
      const double d = (double)r[i] - mean;
      b += d * d;
    }
    const double ss = wave_sum_f64(b) + (double)(m - (src1 - src0)) * mean * mean;
    if (lane == 0) {
      s_mean[q] = mean;
      s_isd[q] = 1.0 / sqrt(ss / (double)(m - 1));
    }
  }
  __syncthreads();
  // the lowest-index member of a cluster writes its aligned row to a0 (after
  // the barrier: its row may have been staged by another wave)
  if (wid == 0) {
    unsigned long long fm = __ballot(first >= 0);
    while (fm) {
      const int q = __builtin_ctzll(fm);
      fm &= fm - 1;
      const int cl = __shfl(first, q);
      const int sh = s_sh[q];
      const float* r = rows_s + q * ms;
      float* o = a0 + (long)cl * m;
      for (int t = lane; t < m; t += kWave) {
        const int src = t - sh;
        o[t] = (src >= 0 && src < m) ? r[src] : 0.f;
      }
    }
  }
  // panel chunks (R/16 column chunks x tm row chunks), 16 bytes per lane:
  // lane ln packs columns qb..qb+7 (their parameters held in registers) at
  // t = 32 * r32 + ln % 32 for the row chunks r32 = w, w+4, ...
  // y = ((a - mean_hi) - mean_lo) * isd in FP32: |y| <= sqrt(m - 1) for a
  // row z-normalized with ddof 1, so nothing overflows; rows whose mean or
  // 1/sd is not finite (constant rows, NaN input) get mean = isd = 0 and
  // pack as zeros (nan_to_num of their zscore).  The BF16 halves are
  // rounded to nearest even by the hardware conversion.
  const int tm = panel_blocks(m);
#pragma unroll 1
  for (int cc = 0; cc < R / 16; ++cc) {
    const int qb = cc * 16 + 8 * (lane >> 5);
    int c_sh[8];
    float c_mh[8], c_ml[8], c_isd[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const double mean = s_mean[qb + v], isd = s_isd[qb + v];
      const bool ok = s_row[qb + v] >= 0 && isfinite(mean) && isfinite(isd);
      c_sh[v] = s_sh[qb + v];
      c_mh[v] = ok ? (float)mean : 0.f;
      c_ml[v] = ok ? (float)(mean - (double)c_mh[v]) : 0.f;
      c_isd[v] = ok ? (float)isd : 0.f;
    }
    for (int r32 = wid; r32 < tm; r32 += 4) {
      const int t = r32 * 32 + (lane & 31);
      const bool tin = t < m;
      float y[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int src = t - c_sh[c];
        const bool in = (unsigned)src < (unsigned)m;  // else the zero fill
        const float a = rows_s[(qb + c) * ms + (in ? src : 0)];
        const float d = ((in ? a : 0.f) - c_mh[c]) - c_ml[c];
        y[c] = tin ? d * c_isd[c] : 0.f;
      }
      uint32_t h[4], l[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const __hip_bfloat162 hb = __float22bfloat162_rn(make_float2(y[2 * u], y[2 * u + 1]));
        const float2 hf = __bfloat1622float2(hb);
        const __hip_bfloat162 lb =
            __float22bfloat162_rn(make_float2(y[2 * u] - hf.x, y[2 * u + 1] - hf.y));
        h[u] = *reinterpret_cast<const uint32_t*>(&hb);
        l[u] = *reinterpret_cast<const uint32_t*>(&lb);
      }
      const long off = (((p0 >> 4) + cc) * (long)tm + r32) * 512 + lane * 8;
      *reinterpret_cast<uint4*>(yhi + off) = make_uint4(h[0], h[1], h[2], h[3]);
      if (ylo)  // NULL: BF16 S (hi only)
        *reinterpret_cast<uint4*>(ylo + off) = make_uint4(l[0], l[1], l[2], l[3]);
    }
  }
}

template <int R>
__global__ __launch_bounds__(256) void k_align_pack_lds(
    const float* __restrict__ x, int m, const int* __restrict__ perm,
    const long* __restrict__ offsets, const int* __restrict__ minrow, int k,
    const int* __restrict__ shift,
    const int* __restrict__ czero, uint16_t* __restrict__ yhi, uint16_t* __restrict__ ylo,
    float* __restrict__ a0) {
  align_pack_lds_body<R>(blockIdx.x, gridDim.x, x, m, perm, offsets, minrow, k, shift, czero, yhi, ylo, a0);
}

// --------------------------------------------------------------- S = Y^T Y
// One wave's 128x128 block of S_j over the packed column chunks [c0, c1)
// (16 columns each): panel rows of the 32-row blocks ra0..ra0+3 against
// rb0..rb0+3, as 4x4 MFMA 32x32x16 BF16 blocks with FP32 accumulation;
// kSplit adds hi*lo + lo*hi (split-BF16).  The panel is in fragment order
// (panel_blocks), so every operand is one coalesced 1 KB load per wave, 16
// bytes per lane; the next 16 columns' operands load while the current ones
// multiply.  The block is stored at o with row stride ostride.
template <bool kSplit>
__device__ __forceinline__ void syrk_wave(const int ra0, const int rb0, const long c0,
                                          const long c1, const int tm,
                                          const uint16_t* __restrict__ yhi,
                                          const uint16_t* __restrict__ ylo,
                                          float* __restrict__ o, const int ostride) {
  const int lane = threadIdx.x & 63;
  const bf16x8* H = reinterpret_cast<const bf16x8*>(yhi) + lane;
  const bf16x8* Lo = reinterpret_cast<const bf16x8*>(ylo) + lane;
  f32x16 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.f;
  // two operand sets used in turn (no register copies between steps, so a
  // step waits only for its own set's loads)
  struct Ops {
    bf16x8 ah[4], al[4], bh[4], bl[4];
  };
  auto load = [&](long c, Ops& x) {
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const long oa = (c * tm + ra0 + a) * 64, ob = (c * tm + rb0 + a) * 64;
      x.ah[a] = H[oa];
      x.bh[a] = H[ob];
      if constexpr (kSplit) {
        x.al[a] = Lo[oa];
        x.bl[a] = Lo[ob];
      }
    }
  };
  auto mma = [&](const Ops& x) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x.ah[a], x.bh[b], acc[a][b], 0, 0, 0);
    if constexpr (kSplit) {
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x.ah[a], x.bl[b], acc[a][b], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x.al[a], x.bh[b], acc[a][b], 0, 0, 0);
    }
  };
  Ops X, Y;
  load(c0, X);
  long c = c0;
  // the scheduling barriers keep each set's loads ahead of the other set's
  // multiplies, so the loads have a whole block of MFMAs to land
  for (; c + 1 < c1; c += 2) {
    load(c + 1, Y);
    __builtin_amdgcn_sched_barrier(0);
    mma(X);
    __builtin_amdgcn_sched_barrier(0);
    load(c + 2 < c1 ? c + 2 : c + 1, X);  // past the end: a harmless reload
    __builtin_amdgcn_sched_barrier(0);
    mma(Y);
    __builtin_amdgcn_sched_barrier(0);
  }
  if (c < c1) mma(X);
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int i = a * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
        const int j = b * 32 + (lane & 31);
        o[i * ostride + j] = acc[a][b][r];
      }
}

// Block u (row-major order of the upper triangle) of a T x T block grid:
// its row and column.
__device__ __forceinline__ void syrk_block_of(int u, const int T, int& ti, int& tj) {
  ti = 0;
  while (u >= T - ti) {
    u -= T - ti;
    ++ti;
  }
  tj = ti + u;
}

// Single fit: every wave is its own work item, derived from its slot, so
// diagonal blocks leave no SIMD idle.  Blocks are dealt to the XCDs round
// robin (block b on XCD b % 8); group g takes the U consecutive wave slots
// 4 * (b / 8) + wave of XCD g % 8 from (g / 8) * U on, so the waves that
// read one group's panel rows share an L2.  Wave u of a group computes the
// kSyrkSub block u of S_j over the group's columns into that block's
// partial.  Slots past the last group are holes.
template <bool kSplit>
__global__ __launch_bounds__(256) void k_syrk(const SyrkGroup* __restrict__ groups, int ngroups,
                                              const uint16_t* __restrict__ yhi,
                                              const uint16_t* __restrict__ ylo, int m,
                                              float* __restrict__ part) {
  static_assert(kSyrkSub == 128, "one wave's block");
  const int T = (m + kSyrkSub - 1) / kSyrkSub, U = T * (T + 1) / 2;
  // the slot is uniform per wave: held in scalar registers, like all that
  // is derived from it
  const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x >> 3) * 4 + (int)(threadIdx.x >> 6));
  const int g = (blockIdx.x & 7) + 8 * (w / U);
  if (g >= ngroups) return;
  const int u = w % U;
  int ti, tj;
  syrk_block_of(u, T, ti, tj);
  const SyrkGroup gr = groups[g];
  syrk_wave<kSplit>(ti * 4, tj * 4, gr.p0 >> 4, gr.p1 >> 4, panel_blocks(m), yhi, ylo,
                    part + (long)(gr.part + u * gr.stride) * kSyrkSub * kSyrkSub, kSyrkSub);
}

// S_j from the K-chunk partials of each kSyrkSub block, summed in a fixed
// order.  One thread block per 64x64 sub-block sb (the lower ones of
// diagonal blocks skipped): 16-byte partial loads, four rows per thread in
// flight, the sums staged in LDS; then the entries with i <= j and their
// mirrors are stored as coalesced row segments, so S_j is exactly
// symmetric.
__device__ __forceinline__ void syrk_reduce_body(const int sb, const SyrkTile tl,
                                                 const float* __restrict__ part, int m,
                                                 float* __restrict__ S) {
  constexpr int TS = kSyrkSub, kSub = TS / 64;
  const int si = sb / kSub, sj = sb % kSub;
  if (tl.ti == tl.tj && si > sj) return;
  const int gi0 = tl.ti * TS + si * 64, gj0 = tl.tj * TS + sj * 64;
  if (gi0 >= m || gj0 >= m) return;
  __shared__ float tr[64][65];
  const int c4 = (threadIdx.x & 15) * 4, rr = threadIdx.x >> 4;
  const float* pp = part + (long)tl.part0 * TS * TS + (long)(si * 64 + rr) * TS +
                    sj * 64 + c4;
  float4 acc[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) acc[u] = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll 2
  for (int c = 0; c < tl.nparts; ++c) {
    const float* pc = pp + (long)c * TS * TS;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float4 v = *reinterpret_cast<const float4*>(pc + (long)16 * u * TS);
      acc[u].x += v.x;
      acc[u].y += v.y;
      acc[u].z += v.z;
      acc[u].w += v.w;
    }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    float* t = &tr[rr + 16 * u][c4];
    t[0] = acc[u].x;
    t[1] = acc[u].y;
    t[2] = acc[u].z;
    t[3] = acc[u].w;
  }
  __syncthreads();
  float* Sj = S + (long)tl.cluster * m * m;
  const int col = threadIdx.x & 63, r0 = threadIdx.x >> 6;
#pragma unroll 4
  for (int r = r0; r < 64; r += 4) {
    const int gi = gi0 + r, gj = gj0 + col;
    if (gi < m && gj < m && gi <= gj) Sj[(long)gi * m + gj] = tr[r][col];
    const int hi = gi0 + col, hj = gj0 + r;  // mirror of entry (hi, hj)
    if (hi < m && hj < m && hi <= hj) Sj[(long)hj * m + hi] = tr[col][r];
  }
}

// Single fit: (kSyrkSub / 64)^2 blocks per kSyrkSub block of each cluster
// with partials, clusters[c] covering the blocks from c * T(T+1)/2 on.
__global__ __launch_bounds__(256) void k_syrk_reduce(const SyrkCluster* __restrict__ clusters,
                                                     const float* __restrict__ part, int m,
                                                     float* __restrict__ S) {
  constexpr int kSub = kSyrkSub / 64;
  const int T = (m + kSyrkSub - 1) / kSyrkSub, U = T * (T + 1) / 2;
  const int t = blockIdx.x / (kSub * kSub), u = t % U;
  const SyrkCluster c = clusters[t / U];
  int ti, tj;
  syrk_block_of(u, T, ti, tj);
  syrk_reduce_body(blockIdx.x % (kSub * kSub),
                             SyrkTile{c.cluster, ti, tj, c.part0 + u * c.nparts, c.nparts}, part,
                             m, S);
}

// ---------------------------------------------------------- power iteration
constexpr int kPowerThreads = 1024;
constexpr int kPowerWaves = kPowerThreads / kWave;

__device__ __forceinline__ double block_sum_f64(double v, double* scratch) {
  v = wave_sum_f64(v);
  const int wid = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) scratch[wid] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < kPowerWaves; ++w) s += scratch[w];
  return s;
}

// Two block sums with one pair of barriers (scratch: 2 * kPowerWaves).
__device__ __forceinline__ void block_sum2_f64(double& a, double& b, double* scratch) {
  a = wave_sum_f64(a);
  b = wave_sum_f64(b);
  const int wid = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    scratch[wid] = a;
    scratch[kPowerWaves + wid] = b;
  }
  __syncthreads();
  double sa = 0.0, sb = 0.0;
#pragma unroll
  for (int w = 0; w < kPowerWaves; ++w) {
    sa += scratch[w];
    sb += scratch[kPowerWaves + w];
  }
  a = sa;
  b = sb;
}

__device__ __forceinline__ int block_min_int(int v, int* scratch) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    int ov = __shfl_xor(v, o, 64);
    v = ov < v ? ov : v;
  }
  const int wid = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) scratch[wid] = v;
  __syncthreads();
  int s = scratch[0];
#pragma unroll
  for (int w = 1; w < kPowerWaves; ++w) s = scratch[w] < s ? scratch[w] : s;
  return s;
}

// Work split of a block-wide symmetric matvec: 64-column strips times
// q_parts slices of the c range, q_parts = max(1, waves / strips).
__device__ __forceinline__ int sym_qparts(int m) {
  const int strips = (m + 63) / 64;
  const int q = kPowerWaves / strips;
  return q > 1 ? q : 1;
}

// LDS doubles of `part`: sym_matvec_block needs at most 4096 (4 * 1024
// float4 partials, or sym_qparts(m) * m <= 4096 for m <= 4096), the
// Lanczos projection max(1024, m) <= 4096.
__host__ __device__ constexpr int sym_part_len(int) { return 4096; }

// y = S x for symmetric S (FP32) with FP64 x (LDS) into y (LDS): thread
// <-> output, so every load is a coalesced row segment of S and x[c] a
// broadcast.  m % 4 == 0: 16-byte loads of 4 outputs per thread, the c
// range split over 1024 / (m/4) thread groups.  `part` holds
// sym_part_len(m) doubles of LDS.
__device__ void sym_matvec_block(const float* __restrict__ Sj, int m, const double* x, double* y,
                                 double* part) {
  if (m % 4 == 0) {
    const int nv = m / 4;  // float4 per row, <= 1024
    const int P = kPowerThreads / nv > 0 ? kPowerThreads / nv : 1;
    const int t = threadIdx.x;
    const int g = t % nv, p = t / nv;
    if (p < P) {
      const int per = (m + P - 1) / P;
      const int c0 = p * per, c1 = c0 + per < m ? c0 + per : m;
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
      const float4* col = reinterpret_cast<const float4*>(Sj) + g;
#pragma unroll 8
      for (int c = c0; c < c1; ++c) {
        const float4 sv = col[(long)c * nv];
        const double xc = x[c];
        a0 += (double)sv.x * xc;
        a1 += (double)sv.y * xc;
        a2 += (double)sv.z * xc;
        a3 += (double)sv.w * xc;
      }
      double* pp = part + (long)p * m + 4 * g;
      pp[0] = a0;
      pp[1] = a1;
      pp[2] = a2;
      pp[3] = a3;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += kPowerThreads) {
      double s = 0.0;
      for (int q = 0; q < P; ++q) s += part[(long)q * m + i];
      y[i] = s;
    }
    __syncthreads();
    return;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int strips = (m + 63) / 64;
  const int Q = sym_qparts(m);
  const int per = (m + Q - 1) / Q;
  for (int item = wid; item < strips * Q; item += kPowerWaves) {
    const int st = item % strips, q = item / strips;
    const int i = st * 64 + lane;
    const int c0 = q * per, c1 = c0 + per < m ? c0 + per : m;
    double acc = 0.0;
    if (i < m) {
#pragma unroll 8
      for (int c = c0; c < c1; ++c) acc += (double)Sj[(long)c * m + i] * x[c];
      part[(long)q * m + i] = acc;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < m; i += kPowerThreads) {
    double t = 0.0;
    for (int q = 0; q < Q; ++q) t += part[(long)q * m + i];
    y[i] = t;
  }
  __syncthreads();
}

// w = M v with M = P S P = S - (r 1^T + 1 r^T)/m + (sigma/m^2) 1 1^T
// (test_timesC.py:145-149 in rank-1 form; r = S 1, sigma = 1^T r), FP32 S,
// FP64 v and accumulation.  Returns (|w|^2, v.w).
__device__ void center_matvec(const float* __restrict__ Sj, int m, const double* v,
                              const double* r, double sigma, double* w, double* part,
                              double* scratch, double& nw2, double& vw) {
  double sv = 0.0, rv = 0.0;
  for (int i = threadIdx.x; i < m; i += kPowerThreads) {
    sv += v[i];
    rv += r[i] * v[i];
  }
  block_sum2_f64(sv, rv, scratch);
  sym_matvec_block(Sj, m, v, w, part);
  const double inv_m = 1.0 / (double)m;
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < m; i += kPowerThreads) {
    const double wi = w[i] - (r[i] * sv + rv) * inv_m + sigma * sv * inv_m * inv_m;
    w[i] = wi;
    a += wi * wi;
    b += v[i] * wi;
  }
  block_sum2_f64(a, b, scratch + 2 * kPowerWaves);
  nw2 = a;
  vw = b;
}

// Runs power steps v <- M v / |M v| until the Rayleigh quotient changes by
// <= tol relative (FP64) or `steps_cap` steps.  Returns 0, or 1 if M v = 0.
__device__ int power_steps(const float* Sj, int m, double* v, const double* r, double sigma,
                           double* w, double* part, double* scratch, int steps_cap, double tol,
                           int& steps, int& converged) {
  double lam_prev = 0.0;
  for (int t = 0; t < steps_cap; ++t) {
    double nw2, vw;
    center_matvec(Sj, m, v, r, sigma, w, part, scratch, nw2, vw);
    if (!(nw2 > 0.0)) return 1;
    const double nw = sqrt(nw2);
    for (int i = threadIdx.x; i < m; i += kPowerThreads) v[i] = w[i] / nw;
    __syncthreads();
    ++steps;
    if (t > 0 && fabs(vw - lam_prev) <= tol * fabs(vw)) {
      converged = 1;
      break;
    }
    lam_prev = vw;
  }
  return 0;
}

__device__ void sign_fix(const float* a0, double* c, double* w, int m, double* scratch,
                         int& flipped) {
  double s1 = 0.0, s2 = 0.0;
  for (int i = threadIdx.x; i < m; i += kPowerThreads) {
    double d1 = (double)a0[i] - c[i], d2 = (double)a0[i] + c[i];
    s1 += d1 * d1;
    s2 += d2 * d2;
  }
  s1 = block_sum_f64(s1, scratch);
  s2 = block_sum_f64(s2, scratch + kPowerWaves);
  if (sqrt(s1) >= sqrt(s2)) {  // test_timesC.py:165
    for (int i = threadIdx.x; i < m; i += kPowerThreads) {
      c[i] = -c[i];
      w[i] = -w[i];
    }
    flipped = 1;
  }
  __syncthreads();
}

// One block per cluster.  Steps of test_timesC.py:152-204: power iteration
// (start vector = power_v0(seed, iter, cluster, .)), lambda = v.Mv, sign fix
// by a0, the vecleft/vecright sign test and the signed-sum residual test,
// then zscore(c, ddof=1).  With vin given, v = vin is the already iterated
// vector (from the Lanczos stage) and no power steps run here.
// eig_mode: vin is the top eigenvector of M (the eigh branch, used for
// cal_select = 1 and for clusters whose power result failed the tests):
// sign fix and zscore only.  defer_fallback: a cluster that fails the tests
// only reports fallback = 1 (the caller reruns it in eig_mode); otherwise,
// and when no eigenvector converged, the power method keeps iterating
// towards the top eigenvector.  mask (may be NULL): clusters to process.
__device__ __forceinline__ void power_body(const int bid, const int nblk,
    const float* __restrict__ S, int m, const int* __restrict__ counts,
    const float* __restrict__ a0all, uint64_t seed, int iter, int max_steps, double tol,
    const double* __restrict__ vin, const double* __restrict__ svin,
    const double* __restrict__ rsum, const int* __restrict__ steps_in,
    const int* __restrict__ conv_in, int eig_mode, const int* __restrict__ mask,
    int defer_fallback, float* __restrict__ c_out, PowerInfo* __restrict__ info) {
  extern __shared__ double sh[];
  double* v = sh;
  double* w = sh + m;
  double* r = sh + 2 * m;
  double* scratch = sh + 3 * m;  // 8 * kPowerWaves doubles
  int* iscratch = reinterpret_cast<int*>(scratch + 6 * kPowerWaves);
  double* part = scratch + 8 * kPowerWaves;  // sym_part_len(m) doubles
  const int j = bid;
  if (mask && !mask[j]) return;
  const float* Sj = S + (long)j * m * m;
  const float* a0 = a0all + (long)j * m;
  float* out = c_out + (long)j * m;
  PowerInfo inf = {};
  if (counts[j] == 0) {  // test_timesC.py:139-140
    for (int i = threadIdx.x; i < m; i += kPowerThreads) out[i] = 0.f;
    if (threadIdx.x == 0) {
      inf.empty = 1;
      info[j] = inf;
    }
    return;
  }
  // r = S 1 (column sums = row sums, S symmetric), sigma = 1^T S 1
  if (rsum) {
    for (int i = threadIdx.x; i < m; i += kPowerThreads) r[i] = rsum[(long)j * m + i];
  } else {
    for (int i = threadIdx.x; i < m; i += kPowerThreads) w[i] = 1.0;
    __syncthreads();
    sym_matvec_block(Sj, m, w, r, part);
  }
  const bool use_vin = vin && (!conv_in || conv_in[j]);
  double n0 = 0.0;
  for (int i = threadIdx.x; i < m; i += kPowerThreads) {
    v[i] = use_vin ? vin[(long)j * m + i] : power_v0(seed, iter, j, i);
    n0 += v[i] * v[i];
  }
  __syncthreads();
  double sigma = 0.0;
  for (int i = threadIdx.x; i < m; i += kPowerThreads) sigma += r[i];
  sigma = block_sum_f64(sigma, scratch);
  n0 = sqrt(block_sum_f64(n0, scratch + kPowerWaves));
  const bool vzero = !(n0 > 0.0);
  for (int i = threadIdx.x; i < m; i += kPowerThreads) v[i] /= n0;
  __syncthreads();

  int steps = use_vin ? (steps_in ? steps_in[j] : max_steps) : 0;
  int converged = use_vin ? 1 : 0;
  const bool zero_path =
      vzero || (!use_vin && power_steps(Sj, m, v, r, sigma, w, part, scratch, max_steps, tol,
                                        steps, converged));
  if (zero_path) {
    // M v = 0: NumPy's v / 0 gives NaN, zscore(NaN) -> 0 (nan_to_num)
    for (int i = threadIdx.x; i < m; i += kPowerThreads) out[i] = 0.f;
    if (threadIdx.x == 0) {
      inf.zero = 1;
      inf.steps = steps;
      info[j] = inf;
    }
    return;
  }
  double nw2, vw;
  if (use_vin && svin) {
    // w = M v from S vin (computed grid-wide): S v = S vin / |vin|
    double sv = 0.0, rv = 0.0;
    for (int i = threadIdx.x; i < m; i += kPowerThreads) {
      sv += v[i];
      rv += r[i] * v[i];
    }
    sv = block_sum_f64(sv, scratch);
    rv = block_sum_f64(rv, scratch + kPowerWaves);
    const double inv_m = 1.0 / (double)m;
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < m; i += kPowerThreads) {
      const double wi = svin[(long)j * m + i] / n0 - (r[i] * sv + rv) * inv_m +
                        sigma * sv * inv_m * inv_m;
      w[i] = wi;
      a += wi * wi;
      b += v[i] * wi;
    }
    nw2 = block_sum_f64(a, scratch + 2 * kPowerWaves);
    vw = block_sum_f64(b, scratch + 3 * kPowerWaves);
  } else {
    center_matvec(Sj, m, v, r, sigma, w, part, scratch, nw2, vw);  // w = M v
  }
  double vv = 0.0;
  for (int i = threadIdx.x; i < m; i += kPowerThreads) vv += v[i] * v[i];
  vv = block_sum_f64(vv, scratch + 4 * kPowerWaves);
  double lambda = vw / vv;  // :37
  int flipped = 0, fallback = 0;
  bool eig_done = false;
  if (eig_mode) {
    // v is the top eigenvector of M (eigh's last column): sign fix only
    // (test_timesC.py:184-192)
    fallback = 1;
    if (use_vin) {
      sign_fix(a0, v, w, m, scratch, flipped);
      eig_done = true;
    }
  } else if (lambda == 0.0) {
    fallback = 1;
  } else {
    sign_fix(a0, v, w, m, scratch, flipped);  // :163-166
    // :170-174 first index where vecleft (= M c = w) and vecright (= lambda c)
    // are both non-zero
    int pos = 0x7fffffff;
    for (int i = threadIdx.x; i < m; i += kPowerThreads)
      if (w[i] != 0.0 && lambda * v[i] != 0.0) {
        pos = i;
        break;
      }
    pos = block_min_int(pos, iscratch);
    if (pos == 0x7fffffff) pos = 0;
    bool neg = (w[pos] < 0.0 && lambda * v[pos] > 0.0);  // :175
    if (neg) lambda = -lambda;
    double sd = 0.0;  // :178-179
    for (int i = threadIdx.x; i < m; i += kPowerThreads) {
      double vr = neg ? -w[i] : lambda * v[i];
      sd += vr - w[i];
    }
    sd = block_sum_f64(sd, scratch + 5 * kPowerWaves);
    if (lambda <= 0.0 || fabs(sd) > (double)m * 0.01) fallback = 1;  // :183
  }
  if (fallback && !eig_done && !(defer_fallback && !eig_mode)) {  // :184-192
    int more = 0, conv2 = 0;
    power_steps(Sj, m, v, r, sigma, w, part, scratch, 8 * max_steps, tol, more, conv2);
    steps += more;
    converged = conv2;
    center_matvec(Sj, m, v, r, sigma, w, part, scratch, nw2, vw);
    flipped = 0;
    sign_fix(a0, v, w, m, scratch, flipped);
  }
  // :204 zscore(c, ddof=1)
  double s = 0.0;
  for (int i = threadIdx.x; i < m; i += kPowerThreads) s += v[i];
  s = block_sum_f64(s, scratch);
  const double mean = s / (double)m;
  double ss = 0.0;
  for (int i = threadIdx.x; i < m; i += kPowerThreads) {
    double d = v[i] - mean;
    ss += d * d;
  }
  ss = block_sum_f64(ss, scratch + kPowerWaves);
  const double sdv = sqrt(ss / (double)(m - 1));
  for (int i = threadIdx.x; i < m; i += kPowerThreads) {
    double y = (v[i] - mean) / sdv;
    out[i] = isnan(y) ? 0.f : (float)y;
  }
  if (threadIdx.x == 0) {
    inf.lambda = lambda;
    inf.steps = steps;
    inf.converged = converged;
    inf.fallback = fallback;
    inf.flipped = flipped;
    info[j] = inf;
  }
}

__global__ __launch_bounds__(kPowerThreads) void k_power(
    const float* __restrict__ S, int m, const int* __restrict__ counts,
    const float* __restrict__ a0all, uint64_t seed, int iter, int max_steps, double tol,
    const double* __restrict__ vin, const double* __restrict__ svin,
    const double* __restrict__ rsum, const int* __restrict__ steps_in,
    const int* __restrict__ conv_in, int eig_mode, const int* __restrict__ mask,
    int defer_fallback, float* __restrict__ c_out, PowerInfo* __restrict__ info) {
  power_body(blockIdx.x, gridDim.x, S, m, counts, a0all, seed, iter, max_steps, tol, vin, svin, rsum, steps_in, conv_in, eig_mode, mask, defer_fallback, c_out, info);
}

// ----------------------------------------------------- Lanczos power result
// The power iterate v_I = M^I v0 / |M^I v0| (test_timesC.py:30-38 with
// I = min(m, 1000)) equals, in exact arithmetic, Q_j T_j^I e1 (normalized)
// once the Krylov space K_j(M, v0) holds every eigen-direction whose weight
// (lambda_i / lambda_1)^I is not negligible: Q_j = [q_1..q_j] orthonormal
// Lanczos basis, T_j = Q_j^T M Q_j tridiagonal.  Lanczos with full
// reorthogonalization in FP64 (Q in global memory, steps below);
// y = T_j^I e1 by I steps of the tridiagonal recurrence in one wave (FP64,
// T scaled by its Gershgorin bound, renormalized every 32 steps).  At the
// lengths lz_power_check selects y is compared with the previous check; the
// result is accepted when |y_j - y_prev| <= lz_tol or the Krylov space
// became invariant (breakdown).
// Clusters that do not converge within kLzMax steps get conv_out = 0 and the
// caller runs the plain power iteration for them.
constexpr int kLzMax = kLanczosMax;

// y = T^I e1 / |T^I e1| for T = tridiag(beta[1..len-1], alpha[0..len-1],
// beta[1..len-1]), len <= 128; executed by wave 0 only (lanes hold entries
// lane and lane + 64).  Writes y[0..len) (zeros if T^I e1 = 0).
__device__ void tridiag_power_wave(const double* alpha, const double* beta, int len, int I,
                                   double* y) {
  const int l = threadIdx.x & 63;
  const int e0 = l, e1 = l + 64;
  double g = 0.0;  // Gershgorin bound of T
  for (int e = l; e < len; e += 64) {
    double r = fabs(alpha[e]) + (e > 0 ? fabs(beta[e]) : 0.0) + (e + 1 < len ? fabs(beta[e + 1]) : 0.0);
    g = r > g ? r : g;
  }
  g = wave_max_f64(g);
  const double sc = g > 0.0 ? 1.0 / g : 0.0;
  const double a0 = e0 < len ? alpha[e0] * sc : 0.0;
  const double a1 = e1 < len ? alpha[e1] * sc : 0.0;
  const double bl0 = (e0 > 0 && e0 < len) ? beta[e0] * sc : 0.0;      // T[e0][e0-1]
  const double bl1 = (e1 < len) ? beta[e1] * sc : 0.0;                // T[e1][e1-1]
  const double bu0 = (e0 + 1 < len) ? beta[e0 + 1] * sc : 0.0;        // T[e0][e0+1]
  const double bu1 = (e1 + 1 < len) ? beta[e1 + 1] * sc : 0.0;        // T[e1][e1+1]
  double y0 = (e0 == 0) ? 1.0 : 0.0, y1 = 0.0;
  if (len <= 64) {
    // one entry per lane, two steps at a time with the pentadiagonal T^2
    // (neighbours by DPP lane shifts only): row e of T^2 is
    // [T_e,e-1 T_e-1,e-2,  T_e,e-1 (T_e-1,e-1 + T_e,e),
    //  T_e,e-1 T_e-1,e + T_e,e^2 + T_e,e+1 T_e+1,e,  T_e,e+1 (T_e,e + T_e+1,e+1),
    //  T_e,e+1 T_e+1,e+2]
    const double blm = wave_shr1_f64(bl0, 0.0), am = wave_shr1_f64(a0, 0.0);
    const double bum = wave_shr1_f64(bu0, 0.0);
    const double ap = wave_shl1_f64(a0, 0.0), blp = wave_shl1_f64(bl0, 0.0);
    const double bup = wave_shl1_f64(bu0, 0.0);
    const double c2m = bl0 * blm, c1m = bl0 * (am + a0);
    const double c0 = bl0 * bum + a0 * a0 + bu0 * blp;
    const double c1p = bu0 * (a0 + ap), c2p = bu0 * bup;
    int s = 0;
    for (; s + 1 < I; s += 2) {
      const double p1 = wave_shr1_f64(y0, 0.0), n1 = wave_shl1_f64(y0, 0.0);
      const double p2 = wave_shr1_f64(p1, 0.0), n2 = wave_shl1_f64(n1, 0.0);
      y0 = (c2m * p2 + c1m * p1) + c0 * y0 + (c1p * n1 + c2p * n2);
      if ((s & 31) == 30 || s + 2 >= I) {
        const double mx = wave_max_f64(fabs(y0));
        if (mx > 0.0) y0 *= 1.0 / mx;
      }
    }
    if (s < I) {  // odd I: one more step with T
      const double p0 = wave_shr1_f64(y0, 0.0), n0 = wave_shl1_f64(y0, 0.0);
      y0 = bl0 * p0 + a0 * y0 + bu0 * n0;
      const double mx = wave_max_f64(fabs(y0));
      if (mx > 0.0) y0 *= 1.0 / mx;
    }
  } else {
  for (int s = 0; s < I; ++s) {
    // neighbours through DPP lane shifts: entries l-1, l+1 and l+63, l+65
    const double y63 = readlane_f64(y0, 63), y64 = readlane_f64(y1, 0);
    const double p0 = wave_shr1_f64(y0, 0.0), p1 = wave_shr1_f64(y1, y63);
    const double n0 = wave_shl1_f64(y0, y64), n1 = wave_shl1_f64(y1, 0.0);
    const double z0 = bl0 * p0 + a0 * y0 + bu0 * n0;
    const double z1 = bl1 * p1 + a1 * y1 + bu1 * n1;
    y0 = z0;
    y1 = z1;
    if ((s & 31) == 31 || s == I - 1) {
      const double mx = wave_max_f64(fmax(fabs(y0), fabs(y1)));
      if (mx > 0.0) {
        const double inv = 1.0 / mx;
        y0 *= inv;
        y1 *= inv;
      }
    }
  }
  }
  double nn = y0 * y0 + y1 * y1;
  nn = wave_sum_f64(nn);
  const double inv = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
  if (e0 < len) y[e0] = y0 * inv;
  if (e1 < len) y[e1] = y1 * inv;
}

// y = T^I e1 / |T^I e1| for T = tridiag(beta[1..len-1], alpha[0..len-1],
// beta[1..len-1]), len <= kLzSqMax: binary powering with dense symmetric
// matrices in LDS (wk: kLzSqDoubles), all NT threads of the block, instead of
// I dependent steps of one wave.  T is scaled by its Gershgorin bound;
// every later power is the square of a symmetric matrix, hence positive
// semidefinite with its largest entry on the diagonal, and is scaled by the
// reciprocal of its trace as it is used (entries in [1/len, 1] in size):
// nothing overflows or underflows.  Every thread computes the scale factors
// and the final norm itself (same LDS values, same order), so each product
// costs one barrier.  Ends with a barrier.
constexpr int kLzSqMax = 48;
constexpr int kLzSqLd = kLzSqMax + 1;  // padded row stride
constexpr int kLzSqDoubles = 2 * kLzSqMax * kLzSqLd + 2 * kLzSqMax;
template <int NT>
__device__ void tridiag_power_block(const double* alpha, const double* beta, int len, int I,
                                    double* wk, double* y) {
  const int tid = threadIdx.x, n = len;
  double* A = wk;
  double* B = wk + kLzSqMax * kLzSqLd;
  double* yv = wk + 2 * kLzSqMax * kLzSqLd;  // kLzSqMax
  double* yt = yv + kLzSqMax;                // kLzSqMax
  double g = 0.0;
#pragma unroll 8
  for (int e = 0; e < n; ++e)
    g = fmax(g, fabs(alpha[e]) + (e > 0 ? fabs(beta[e]) : 0.0) +
                    (e + 1 < n ? fabs(beta[e + 1]) : 0.0));
  const double sc = g > 0.0 ? 1.0 / g : 0.0;
  for (int idx = tid; idx < n * n; idx += NT) {
    const int i = idx / n, jj = idx - i * n;
    const double v = i == jj ? alpha[i] : (jj == i + 1 ? beta[jj] : (i == jj + 1 ? beta[i] : 0.0));
    A[i * kLzSqLd + jj] = v * sc;
  }
  for (int i = tid; i < n; i += NT) yv[i] = i == 0 ? 1.0 : 0.0;
  __syncthreads();
  double f = 1.0;  // the current power is f * A
  for (int e = I;;) {
    if (e & 1) {  // yt = f A yv, then swap
      for (int i = tid; i < n; i += NT) {
        const double* ai = A + i * kLzSqLd;
        double s0 = 0.0, s1 = 0.0;
        int c = 0;
#pragma unroll 4
        for (; c + 1 < n; c += 2) {
          s0 += ai[c] * yv[c];
          s1 += ai[c + 1] * yv[c + 1];
        }
        if (c < n) s0 += ai[c] * yv[c];
        yt[i] = (s0 + s1) * f;
      }
      __syncthreads();
      double* t = yv;
      yv = yt;
      yt = t;
    }
    e >>= 1;
    if (e == 0) break;
    // B = (f A)^2; A symmetric, so B[i][j] = f^2 sum_c A[i][c] A[j][c]
    const double f2 = f * f;
    for (int idx = tid; idx < n * n; idx += NT) {
      const int i = idx / n, jj = idx - i * n;
      const double *ai = A + i * kLzSqLd, *aj = A + jj * kLzSqLd;
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      int c = 0;
#pragma unroll 2
      for (; c + 3 < n; c += 4) {
        s0 += ai[c] * aj[c];
        s1 += ai[c + 1] * aj[c + 1];
        s2 += ai[c + 2] * aj[c + 2];
        s3 += ai[c + 3] * aj[c + 3];
      }
      for (; c < n; ++c) s0 += ai[c] * aj[c];
      B[i * kLzSqLd + jj] = ((s0 + s1) + (s2 + s3)) * f2;
    }
    __syncthreads();
    double tr = 0.0;
#pragma unroll 8
    for (int i = 0; i < n; ++i) tr += B[i * kLzSqLd + i];
    f = tr > 0.0 ? 1.0 / tr : 0.0;
    double* t = A;
    A = B;
    B = t;
  }
  double nn = 0.0;
#pragma unroll 8
  for (int i = 0; i < n; ++i) nn += yv[i] * yv[i];
  const double inv = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
  for (int i = tid; i < n; i += NT) y[i] = yv[i] * inv;
  __syncthreads();
}

// Number of eigenvalues of T = tridiag(beta, alpha, beta) (len n) below x:
// sign count of the LDL^T pivots (Sturm sequence), pivots guarded away
// from zero by pivmin.
__device__ int sturm_count(const double* alpha, const double* beta, int n, double x,
                           double pivmin) {
  double q = alpha[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int c = q < 0.0;
  for (int i = 1; i < n; ++i) {
    q = (alpha[i] - x) - beta[i] * beta[i] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    c += q < 0.0;
  }
  return c;
}

// Largest eigenvalue theta of T and its unit eigenvector s (len n <= 128),
// executed by wave 0: multisection on the Sturm count (64 points per
// round) to full FP64 resolution, then inverse iteration with Gaussian
// elimination with partial pivoting on T - theta I (two solves, lane 0).
// wk holds 5 * kLzMax doubles of LDS.  Returns theta.
__device__ double tridiag_top_eig_wave(const double* alpha, const double* beta, int n,
                                       double* wk, double* s) {
  const int l = threadIdx.x & 63;
  double lo = 0.0, hi = 0.0, g = 0.0;
  for (int e = l; e < n; e += 64) {
    const double r = (e > 0 ? fabs(beta[e]) : 0.0) + (e + 1 < n ? fabs(beta[e + 1]) : 0.0);
    const double a = alpha[e];
    lo = (e == l) ? a - r : fmin(lo, a - r);
    hi = (e == l) ? a + r : fmax(hi, a + r);
    g = fmax(g, fabs(a) + r);
  }
  if (l >= n) {  // lanes without entries take lane 0's values below
    lo = INFINITY;
    hi = -INFINITY;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmax(hi, __shfl_xor(hi, o, 64));
    g = fmax(g, __shfl_xor(g, o, 64));
  }
  const double pivmin = fmax(g, 1e-300) * 1e-300;
  double L = lo, H = hi;  // all eigenvalues lie in [L, H]
  for (int round = 0; round < 16; ++round) {
    if (!(H - L > 4.0 * DBL_EPSILON * fmax(fabs(L), fabs(H)))) break;
    const double x = L + (H - L) * (double)(l + 1) / 65.0;
    const bool above = sturm_count(alpha, beta, n, x, pivmin) >= n;  // x above the top
    const unsigned long long mask = __ballot(above);
    const int first = mask ? __builtin_ctzll(mask) : 64;  // first point above the top
    const double nH = first < 64 ? L + (H - L) * (double)(first + 1) / 65.0 : H;
    const double nL = first > 0 ? L + (H - L) * (double)first / 65.0 : L;
    L = nL;
    H = nH;
  }
  const double theta = 0.5 * (L + H);
  if (l == 0) {
    // inverse iteration: (T - theta I) s = rhs, LU with partial pivoting
    double* dl = wk;               // sub-diagonal, then multipliers
    double* d = wk + kLzMax;       // diagonal of U
    double* du = wk + 2 * kLzMax;  // first super-diagonal of U
    double* du2 = wk + 3 * kLzMax; // second super-diagonal of U
    int* piv = reinterpret_cast<int*>(wk + 4 * kLzMax);  // 1: rows i, i+1 swapped
    const double tiny = fmax(g, 1e-300) * DBL_EPSILON;
    for (int i = 0; i < n; ++i) {
      d[i] = alpha[i] - theta;
      if (i + 1 < n) {
        dl[i] = beta[i + 1];
        du[i] = beta[i + 1];
      }
      du2[i] = 0.0;
      piv[i] = 0;
    }
    for (int i = 0; i + 1 < n; ++i) {
      if (fabs(d[i]) >= fabs(dl[i])) {
        if (d[i] == 0.0) d[i] = tiny;
        const double f = dl[i] / d[i];
        dl[i] = f;
        d[i + 1] -= f * du[i];
      } else {
        const double f = d[i] / dl[i];
        d[i] = dl[i];
        dl[i] = f;
        const double t = du[i];
        du[i] = d[i + 1];
        d[i + 1] = t - f * d[i + 1];
        if (i + 2 < n) {
          du2[i] = du[i + 1];
          du[i + 1] = -f * du[i + 1];
        }
        piv[i] = 1;
      }
    }
    if (d[n - 1] == 0.0) d[n - 1] = tiny;
    for (int i = 0; i < n; ++i) s[i] = 1.0;
    for (int rep = 0; rep < 3; ++rep) {
      for (int i = 0; i + 1 < n; ++i) {  // L y = P rhs
        if (piv[i]) {
          const double t = s[i];
          s[i] = s[i + 1];
          s[i + 1] = t - dl[i] * s[i + 1];
        } else {
          s[i + 1] -= dl[i] * s[i];
        }
      }
      for (int i = n - 1; i >= 0; --i) {  // U x = y
        double t = s[i];
        if (i + 1 < n) t -= du[i] * s[i + 1];
        if (i + 2 < n) t -= du2[i] * s[i + 2];
        s[i] = t / d[i];
      }
      double nn = 0.0, mx = 0.0;
      for (int i = 0; i < n; ++i) mx = fmax(mx, fabs(s[i]));
      for (int i = 0; i < n; ++i) {
        s[i] /= mx;
        nn += s[i] * s[i];
      }
      const double inv = 1.0 / sqrt(nn);
      for (int i = 0; i < n; ++i) s[i] *= inv;
    }
  }
  return theta;
}

// Fused form for small m (S_j resident in L2): one block runs every step of
// one cluster, with the symmetric matvec done by the block itself.
//
// Two uses.  vstart NULL: Sj is the cluster's S (FP32, m x m), the operator
// is M = P S P (rank-1 centring from r = S 1, written to rj), the start is
// power_v0(seed, iter, j, .), and vj = M^I v0 / |.| (zero when the check
// never passed: the caller's power method takes over).
// vstart given: Sj is an already centred symmetric matrix (the Gram form of
// a thin cluster), no centring, the start is vstart; vj = Sj^I vstart / |.|
// and hj = Sj vj.  A zero start gives vj = hj = 0; when the check never
// passes, plain power steps on Sj from vstart (at most I of them, stopping
// when the Rayleigh quotient moves by <= 1e-12 relative) give vj.
__device__ __forceinline__ void power_lanczos_cluster(const float* __restrict__ Sj, int m,
    const double* __restrict__ vstart, uint64_t seed, int iter, int j, int I, double lz_tol,
    double* __restrict__ Q, double* __restrict__ vj, double* __restrict__ rj,
    double* __restrict__ hj, int* __restrict__ steps_j, int* __restrict__ conv_j) {
  extern __shared__ double lz_sh[];
  double* q = lz_sh;                  // m
  double* w = q + m;                  // m
  double* r = w + m;                  // m
  double* qp = r + m;                 // m: previous Lanczos vector
  double* scratch = qp + m;           // 8 * kPowerWaves
  double* alpha = scratch + 8 * kPowerWaves;  // kLzMax + 1
  double* beta = alpha + kLzMax + 1;  // kLzMax + 1
  double* cf = beta + kLzMax + 1;     // kLzMax + 1
  double* y = cf + kLzMax + 1;        // kLzMax
  double* yp = y + kLzMax;            // kLzMax
  int* flag = reinterpret_cast<int*>(yp + kLzMax);  // 2 ints (padded to 1 double)
  double* part = yp + kLzMax + 1;     // sym_part_len(m)
  // the acceptance test (kLzSqDoubles) reuses the partials of the matvec and
  // the projection: the two are never live at once
  double* sq = part;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double sigma = 0.0, n0 = 0.0;
  if (vstart) {
    for (int i = threadIdx.x; i < m; i += kPowerThreads) {
      r[i] = 0.0;
      q[i] = vstart[i];
      qp[i] = 0.0;
      n0 += q[i] * q[i];
    }
    n0 = sqrt(block_sum_f64(n0, scratch + kPowerWaves));
    if (!(n0 > 0.0)) {  // every thread sees the same sum
      for (int i = threadIdx.x; i < m; i += kPowerThreads) {
        vj[i] = 0.0;
        hj[i] = 0.0;
      }
      if (threadIdx.x == 0) {
        *steps_j = 0;
        *conv_j = 1;
      }
      return;
    }
  } else {
    for (int i = threadIdx.x; i < m; i += kPowerThreads) w[i] = 1.0;
    __syncthreads();
    sym_matvec_block(Sj, m, w, r, part);  // r = S 1
    double sg = 0.0;
    for (int i = threadIdx.x; i < m; i += kPowerThreads) {
      sg += r[i];
      q[i] = power_v0(seed, iter, j, i);
      qp[i] = 0.0;
      n0 += q[i] * q[i];
      rj[i] = r[i];
    }
    sigma = block_sum_f64(sg, scratch);
    n0 = sqrt(block_sum_f64(n0, scratch + kPowerWaves));
  }
  for (int i = threadIdx.x; i < m; i += kPowerThreads) {
    q[i] /= n0;
    Q[i] = q[i];
  }
  if (threadIdx.x == 0) beta[0] = 0.0;
  __syncthreads();
  const int jmax = m < kLzMax ? m : kLzMax;
  // projection split: H groups of threads share the l range of w -= Q c
  const int H = m <= kPowerThreads ? kPowerThreads / m : 1;
  int len = 0, conv = 0, prev_len = 0;
  for (int it = 0; it < jmax; ++it) {
    double nw2, qw;
    center_matvec(Sj, m, q, r, sigma, w, part, scratch, nw2, qw);  // w = M q_it
    const double bprev = beta[it];
    // three-term recurrence, then full reorthogonalization (classical
    // Gram-Schmidt against q_0..q_it, repeated when it removed more than
    // 1 - 1/sqrt(2) of the norm: Daniel-Gragg-Kaufman-Stewart)
    double w1 = 0.0;
    for (int e = threadIdx.x; e < m; e += kPowerThreads) {
      const double x = w[e] - qw * q[e] - bprev * qp[e];
      w[e] = x;
      w1 += x * x;
    }
    w1 = block_sum_f64(w1, scratch + kPowerWaves);
    double b = sqrt(w1);
    for (int pass = 0; pass < 2; ++pass) {
      for (int l = wid; l <= it; l += kPowerWaves) {
        const double* ql = Q + (long)l * m;
        double a0 = 0.0, a1 = 0.0;
        int e = lane;
        for (; e + 64 < m; e += 128) {
          a0 += ql[e] * w[e];
          a1 += ql[e + 64] * w[e + 64];
        }
        if (e < m) a0 += ql[e] * w[e];
        const double acc = wave_sum_f64(a0 + a1);
        if (lane == 0) cf[l] = acc;
      }
      __syncthreads();
      for (int t = threadIdx.x; t < m * H; t += kPowerThreads) {
        const int e = t % m, g = t / m;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int l = g;
        for (; l + 3 * H <= it; l += 4 * H) {
          a0 += cf[l] * Q[(long)l * m + e];
          a1 += cf[l + H] * Q[(long)(l + H) * m + e];
          a2 += cf[l + 2 * H] * Q[(long)(l + 2 * H) * m + e];
          a3 += cf[l + 3 * H] * Q[(long)(l + 3 * H) * m + e];
        }
        for (; l <= it; l += H) a0 += cf[l] * Q[(long)l * m + e];
        part[(long)g * m + e] = (a0 + a1) + (a2 + a3);
      }
      __syncthreads();
      double w2 = 0.0;
      for (int e = threadIdx.x; e < m; e += kPowerThreads) {
        double s2 = 0.0;
        for (int g = 0; g < H; ++g) s2 += part[(long)g * m + e];
        const double x = w[e] - s2;
        w[e] = x;
        w2 += x * x;
      }
      w2 = block_sum_f64(w2, scratch + 2 * kPowerWaves);
      const bool again = w2 < 0.5 * w1;  // |w| fell below |w1| / sqrt(2)
      b = sqrt(w2);
      w1 = w2;
      if (!again) break;
    }
    if (threadIdx.x == 0) alpha[it] = qw;
    len = it + 1;
    const bool breakdown = !(b > 1e-12 * sqrt(nw2));
    if (breakdown || len == jmax || lz_power_check(len)) {
      __syncthreads();
      if (len <= kLzSqMax) {
        tridiag_power_block<kPowerThreads>(alpha, beta, len, I, sq, y);
      } else {
        if (wid == 0) tridiag_power_wave(alpha, beta, len, I, y);
        __syncthreads();
      }
      if (wid == 0) {
        double d = 0.0;
        for (int e = lane; e < len; e += 64) {
          const double pv = e < prev_len ? yp[e] : 0.0;  // previous y, zero padded
          const double dd = y[e] - pv;
          d += dd * dd;
        }
        d = wave_sum_f64(d);
        if (lane == 0) flag[0] = breakdown || (prev_len > 0 && sqrt(d) <= lz_tol);
        for (int e = lane; e < len; e += 64) yp[e] = y[e];
      }
      prev_len = len;
      __syncthreads();
      conv = flag[0];
      __syncthreads();
      if (conv || breakdown) break;
    }
    if (len == jmax) break;
    if (threadIdx.x == 0) beta[it + 1] = b;
    double* qn = Q + (long)(it + 1) * m;
    const double inv_b = 1.0 / b;
    for (int e = threadIdx.x; e < m; e += kPowerThreads) {
      qp[e] = q[e];
      q[e] = w[e] * inv_b;
      qn[e] = q[e];
    }
    __syncthreads();
  }
  if (!vstart) {
    // v = Q y
    for (int e = threadIdx.x; e < m; e += kPowerThreads) {
      double acc = 0.0;
      for (int l = 0; l < len; ++l) acc += y[l] * Q[(long)l * m + e];
      vj[e] = conv ? acc : 0.0;
    }
    if (threadIdx.x == 0) {
      *steps_j = len;
      *conv_j = conv;
    }
    return;
  }
  // Gram form: v (into q) = Q y, or the power steps when the check never
  // passed; then h = Sj v
  int steps = len;
  if (conv) {
    for (int e = threadIdx.x; e < m; e += kPowerThreads) {
      double acc = 0.0;
      for (int l = 0; l < len; ++l) acc += y[l] * Q[(long)l * m + e];
      q[e] = acc;
    }
  } else {
    for (int e = threadIdx.x; e < m; e += kPowerThreads) {
      q[e] = vstart[e] / n0;
      r[e] = 0.0;
    }
    __syncthreads();
    int more = 0, conv2 = 0;
    if (power_steps(Sj, m, q, r, 0.0, w, part, scratch, I, 1e-12, more, conv2))
      for (int e = threadIdx.x; e < m; e += kPowerThreads) q[e] = 0.0;  // S v = 0
    steps += more;
  }
  __syncthreads();
  sym_matvec_block(Sj, m, q, w, part);  // h = Sj v (ends with a barrier)
  for (int e = threadIdx.x; e < m; e += kPowerThreads) {
    vj[e] = q[e];
    hj[e] = w[e];
  }
  if (threadIdx.x == 0) {
    *steps_j = steps;
    *conv_j = 1;
  }
}

__device__ __forceinline__ void power_lanczos_body(const int bid, const int nblk,
    const float* __restrict__ S, int m, const int* __restrict__ counts, uint64_t seed, int iter,
    int I, double lz_tol, double* __restrict__ Qg, double* __restrict__ vout,
    double* __restrict__ rout, int* __restrict__ steps_out, int* __restrict__ conv_out) {
  const int j = bid;
  if (counts[j] == 0) return;
  power_lanczos_cluster(S + (long)j * m * m, m, nullptr, seed, iter, j, I, lz_tol,
                        Qg + (long)j * (kLzMax + 1) * m, vout + (long)j * m, rout + (long)j * m,
                        nullptr, steps_out + j, conv_out + j);
}

__global__ __launch_bounds__(kPowerThreads) void k_power_lanczos(
    const float* __restrict__ S, int m, const int* __restrict__ counts, uint64_t seed, int iter,
    int I, double lz_tol, double* __restrict__ Qg, double* __restrict__ vout,
    double* __restrict__ rout, int* __restrict__ steps_out, int* __restrict__ conv_out) {
  power_lanczos_body(blockIdx.x, gridDim.x, S, m, counts, seed, iter, I, lz_tol, Qg, vout, rout, steps_out, conv_out);
}

// --------------------------------------------- Lanczos, one step per launch
// The Lanczos stage runs as a host-driven sequence of two launches per step
// so that the symmetric matvec S_j q_j of every cluster spreads over the
// whole GPU instead of one CU per cluster:
//   k_cluster_matvec  y_j = S_j x_j for all active clusters (grid-wide,
//                     one wave per row of S, FP64 accumulation)
//   k_lz_step         one block per cluster: centring, three-term
//                     recurrence, full reorthogonalization, the next basis
//                     vector and, every 8 steps from 16 on, the check of
//                     y = T^I e1 against the previous check.
// k_lz_setup starts every cluster (r = S 1 from a matvec of ones, sigma,
// q_0 = power_v0 / |power_v0|) and k_lz_finish forms v = Q y.  State that
// lives across launches (alpha, beta, y, step counters) is in global
// memory, LzState per cluster.
constexpr int kLzPower = 0;  // mode 0: M^I v0 (power iterate); mode 1: top eigenvector of M
constexpr int kMvThreads = 256;
constexpr int kMvRowsPerWave = 4;
constexpr int kMvRows = kMvRowsPerWave * (kMvThreads / kWave);  // rows per block

__device__ __forceinline__ void cluster_matvec_body(const int bid, const int nblk,
    const float* __restrict__ S, int m, int k, const double* __restrict__ x, long xstride,
    double* __restrict__ y, const LzState* __restrict__ state) {
  const int panels = (m + kMvRows - 1) / kMvRows;
  const int j = bid / panels, p = bid % panels;
  if (j >= k || (state && state[j].done)) return;
  extern __shared__ double xs[];  // m doubles
  const double* xj = x ? x + (long)j * xstride : nullptr;
  for (int i = threadIdx.x; i < m; i += kMvThreads) xs[i] = xj ? xj[i] : 1.0;
  __syncthreads();
  const float* Sj = S + (long)j * m * m;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int r0 = p * kMvRows + wid * kMvRowsPerWave;
  double acc[kMvRowsPerWave];
#pragma unroll
  for (int q = 0; q < kMvRowsPerWave; ++q) acc[q] = 0.0;
  if ((m & 3) == 0) {
    const int nv = m / 4;
    for (int c = lane; c < nv; c += kWave) {
      const double x0 = xs[4 * c], x1 = xs[4 * c + 1], x2 = xs[4 * c + 2], x3 = xs[4 * c + 3];
#pragma unroll
      for (int q = 0; q < kMvRowsPerWave; ++q) {
        const int r = r0 + q < m ? r0 + q : m - 1;  // clamped rows are discarded below
        const float4 sv = reinterpret_cast<const float4*>(Sj + (long)r * m)[c];
        acc[q] += (double)sv.x * x0 + (double)sv.y * x1 + (double)sv.z * x2 + (double)sv.w * x3;
      }
    }
  } else {
    for (int c = lane; c < m; c += kWave) {
      const double xc = xs[c];
#pragma unroll
      for (int q = 0; q < kMvRowsPerWave; ++q) {
        const int r = r0 + q < m ? r0 + q : m - 1;
        acc[q] += (double)Sj[(long)r * m + c] * xc;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kMvRowsPerWave; ++q) {
    const double t = wave_sum_f64(acc[q]);
    if (lane == 0 && r0 + q < m) y[(long)j * m + r0 + q] = t;
  }
}

__global__ __launch_bounds__(kMvThreads) void k_cluster_matvec(
    const float* __restrict__ S, int m, int k, const double* __restrict__ x, long xstride,
    double* __restrict__ y, const LzState* __restrict__ state) {
  cluster_matvec_body(blockIdx.x, gridDim.x, S, m, k, x, xstride, y, state);
}

// Starts the Lanczos process of every cluster from r = S 1 (in rbuf).
__device__ __forceinline__ void lz_setup_body(const int bid, const int nblk,
    int m, const int* __restrict__ counts, const int* __restrict__ mask, uint64_t seed, int iter,
    const double* __restrict__ rbuf, double* __restrict__ Qg, LzState* __restrict__ state) {
  __shared__ double scratch[2 * kPowerWaves];
  const int j = bid;
  double* Q = Qg + (long)j * (kLzMax + 1) * m;
  double sg = 0.0, n0 = 0.0;
  for (int i = threadIdx.x; i < m; i += kPowerThreads) {
    sg += rbuf[(long)j * m + i];
    const double v = power_v0(seed, iter, j, i);
    n0 += v * v;
  }
  const double sigma = block_sum_f64(sg, scratch);
  n0 = sqrt(block_sum_f64(n0, scratch + kPowerWaves));
  for (int i = threadIdx.x; i < m; i += kPowerThreads) Q[i] = power_v0(seed, iter, j, i) / n0;
  if (threadIdx.x == 0) {
    LzState z = {};
    z.sigma = sigma;
    z.done = counts[j] == 0 || (mask && !mask[j]);
    state[j] = z;
  }
}

__global__ __launch_bounds__(kPowerThreads) void k_lz_setup(
    int m, const int* __restrict__ counts, const int* __restrict__ mask, uint64_t seed, int iter,
    const double* __restrict__ rbuf, double* __restrict__ Qg, LzState* __restrict__ state) {
  lz_setup_body(blockIdx.x, gridDim.x, m, counts, mask, seed, iter, rbuf, Qg, state);
}

// One Lanczos step of every active cluster, given w = S q_it in wbuf.
__device__ __forceinline__ void lz_step_body(const int bid, const int nblk,
    int m, int mode, int I, double lz_tol, const double* __restrict__ rbuf,
    const double* __restrict__ wbuf, double* __restrict__ Qg, LzState* __restrict__ state) {
  extern __shared__ double lz_sh[];
  double* q = lz_sh;              // m
  double* w = q + m;              // m
  double* r = w + m;              // m
  double* qp = r + m;             // m: previous Lanczos vector
  double* scratch = qp + m;       // 8 * kPowerWaves
  double* cf = scratch + 8 * kPowerWaves;  // kLzMax + 1
  double* part = cf + kLzMax + 1;          // sym_part_len(m)
  const int j = bid;
  LzState* stp = state + j;
  if (stp->done) return;
  const int it = stp->it;
  const double sigma = stp->sigma;
  double* Q = Qg + (long)j * (kLzMax + 1) * m;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double sv = 0.0, rv = 0.0;
  for (int e = threadIdx.x; e < m; e += kPowerThreads) {
    const double qe = Q[(long)it * m + e];
    q[e] = qe;
    qp[e] = it > 0 ? Q[(long)(it - 1) * m + e] : 0.0;
    r[e] = rbuf[(long)j * m + e];
    w[e] = wbuf[(long)j * m + e];
    sv += qe;
    rv += r[e] * qe;
  }
  sv = block_sum_f64(sv, scratch);
  rv = block_sum_f64(rv, scratch + kPowerWaves);
  // w = M q = S q - (r 1^T q + 1 r^T q)/m + sigma (1^T q)/m^2 1
  const double inv_m = 1.0 / (double)m;
  double a = 0.0, b0 = 0.0;
  for (int e = threadIdx.x; e < m; e += kPowerThreads) {
    const double we = w[e] - (r[e] * sv + rv) * inv_m + sigma * sv * inv_m * inv_m;
    w[e] = we;
    a += we * we;
    b0 += q[e] * we;
  }
  const double nw2 = block_sum_f64(a, scratch + 2 * kPowerWaves);
  const double qw = block_sum_f64(b0, scratch + 3 * kPowerWaves);
  const double bprev = it > 0 ? stp->beta[it] : 0.0;
  // three-term recurrence, then full reorthogonalization (classical
  // Gram-Schmidt against q_0..q_it, repeated when it removed more than
  // 1 - 1/sqrt(2) of the norm: Daniel-Gragg-Kaufman-Stewart)
  double w1 = 0.0;
  for (int e = threadIdx.x; e < m; e += kPowerThreads) {
    const double xv = w[e] - qw * q[e] - bprev * qp[e];
    w[e] = xv;
    w1 += xv * xv;
  }
  w1 = block_sum_f64(w1, scratch + 4 * kPowerWaves);
  double b = sqrt(w1);
  const int H = m <= kPowerThreads ? kPowerThreads / m : 1;
  for (int pass = 0; pass < 2; ++pass) {
    for (int l = wid; l <= it; l += kPowerWaves) {
      const double* ql = Q + (long)l * m;
      double a0 = 0.0, a1 = 0.0;
      int e = lane;
      for (; e + 64 < m; e += 128) {
        a0 += ql[e] * w[e];
        a1 += ql[e + 64] * w[e + 64];
      }
      if (e < m) a0 += ql[e] * w[e];
      const double acc = wave_sum_f64(a0 + a1);
      if (lane == 0) cf[l] = acc;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < m * H; t += kPowerThreads) {
      const int e = t % m, g = t / m;
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
      int l = g;
      for (; l + 3 * H <= it; l += 4 * H) {
        a0 += cf[l] * Q[(long)l * m + e];
        a1 += cf[l + H] * Q[(long)(l + H) * m + e];
        a2 += cf[l + 2 * H] * Q[(long)(l + 2 * H) * m + e];
        a3 += cf[l + 3 * H] * Q[(long)(l + 3 * H) * m + e];
      }
      for (; l <= it; l += H) a0 += cf[l] * Q[(long)l * m + e];
      part[(long)g * m + e] = (a0 + a1) + (a2 + a3);
    }
    __syncthreads();
    double w2 = 0.0;
    for (int e = threadIdx.x; e < m; e += kPowerThreads) {
      double s2 = 0.0;
      for (int g = 0; g < H; ++g) s2 += part[(long)g * m + e];
      const double xv = w[e] - s2;
      w[e] = xv;
      w2 += xv * xv;
    }
    w2 = block_sum_f64(w2, scratch + 5 * kPowerWaves);
    const bool again = w2 < 0.5 * w1;  // |w| fell below |w1| / sqrt(2)
    b = sqrt(w2);
    w1 = w2;
    if (!again) break;
  }
  if (threadIdx.x == 0) stp->alpha[it] = qw;
  const int len = it + 1;
  const int jmax = m < kLzMax ? m : kLzMax;
  const bool breakdown = !(b > 1e-12 * sqrt(nw2));
  int finished = breakdown || len == jmax;
  const bool check = mode == kLzPower ? lz_power_check(len) : (len >= 4 && (len & 3) == 0);
  if (breakdown || len == jmax || check) {
    __syncthreads();  // alpha[it] visible to wave 0
    if (wid == 0) {
      double* y = cf;  // kLzMax doubles of LDS
      int conv;
      if (mode == kLzPower) {
        tridiag_power_wave(stp->alpha, stp->beta, len, I, y);
        const int prev_len = stp->prev_len;
        double d = 0.0;
        for (int e = lane; e < len; e += 64) {
          const double pv = e < prev_len ? stp->y[e] : 0.0;  // previous y, zero padded
          const double dd = y[e] - pv;
          d += dd * dd;
        }
        d = wave_sum_f64(d);
        conv = breakdown || (prev_len > 0 && sqrt(d) <= lz_tol);
      } else {
        // top Ritz pair (theta, Q y); its residual |M Q y - theta Q y| is
        // b |y[len-1]| (b = norm of the next Lanczos residual)
        const double theta = tridiag_top_eig_wave(stp->alpha, stp->beta, len, part, y);
        wave_lds_sync();
        conv = breakdown || b * fabs(y[len - 1]) <= lz_tol * fabs(theta);
      }
      for (int e = lane; e < len; e += 64) stp->y[e] = y[e];
      if (lane == 0) {
        stp->prev_len = len;
        stp->conv = conv;
        scratch[0] = conv;
      }
    }
    __syncthreads();
    if (scratch[0] != 0.0) finished = 1;
  }
  if (finished) {
    if (threadIdx.x == 0) {
      stp->len = len;
      stp->done = 1;
    }
    return;
  }
  double* qn = Q + (long)(it + 1) * m;
  for (int e = threadIdx.x; e < m; e += kPowerThreads) qn[e] = w[e] / b;
  if (threadIdx.x == 0) {
    stp->beta[it + 1] = b;
    stp->it = it + 1;
    stp->len = len;
  }
}

__global__ __launch_bounds__(kPowerThreads) void k_lz_step(
    int m, int mode, int I, double lz_tol, const double* __restrict__ rbuf,
    const double* __restrict__ wbuf, double* __restrict__ Qg, LzState* __restrict__ state) {
  lz_step_body(blockIdx.x, gridDim.x, m, mode, I, lz_tol, rbuf, wbuf, Qg, state);
}

// v = Q y of every cluster (zeros when its Lanczos run was not accepted).
__device__ __forceinline__ void lz_finish_body(const int bid, const int nblk,
    int m, const int* __restrict__ counts, const double* __restrict__ Qg,
    const LzState* __restrict__ state, double* __restrict__ vout, int* __restrict__ steps_out,
    int* __restrict__ conv_out) {
  const int j = bid;
  const LzState& z = state[j];
  const int len = counts[j] == 0 ? 0 : z.len;
  const int conv = counts[j] != 0 && z.conv;
  const double* Q = Qg + (long)j * (kLzMax + 1) * m;
  for (int e = threadIdx.x; e < m; e += kPowerThreads) {
    double acc = 0.0;
    if (conv)
      for (int l = 0; l < len; ++l) acc += z.y[l] * Q[(long)l * m + e];
    vout[(long)j * m + e] = acc;
  }
  if (threadIdx.x == 0) {
    steps_out[j] = len;
    conv_out[j] = conv;
  }
}

__global__ __launch_bounds__(kPowerThreads) void k_lz_finish(
    int m, const int* __restrict__ counts, const double* __restrict__ Qg,
    const LzState* __restrict__ state, double* __restrict__ vout, int* __restrict__ steps_out,
    int* __restrict__ conv_out) {
  lz_finish_body(blockIdx.x, gridDim.x, m, counts, Qg, state, vout, steps_out, conv_out);
}

// ------------------------------------ Lanczos, several blocks per cluster
// One launch in which the G member blocks of a cluster share its Lanczos
// process: member b owns rows [b R, b R + R) of S_j (in LDS when they fit)
// and the same rows of the basis Q (LDS, one row of kLzCoopLQ columns per
// matrix row), so a step reads S_j once over G CUs instead of once on one
// CU.  Step it (q = q_it held whole by every member):
//   A  w = M q on the member's rows (M q = S q - (r 1^T q + 1 r^T q)/m +
//      sigma (1^T q)/m^2 1); partial dots Q^T w and |w|^2 published
//   B  c = Q^T w (partials summed in member order); w -= Q c (classical
//      Gram-Schmidt against the whole basis)
//   C  only when B left less than kLzCoopReorth of |w|^2 (|c|^2 against
//      |w|^2): partial dots of the new w published, d = Q^T w, w -= Q d
//   D  the member's rows of w published; beta = |w|, q_{it+1} = w / beta
//      assembled from every member's rows.
// alpha_it = c_it (+ d_it).  Every member sums the same values in the same
// order, so alpha, beta and the acceptance test of y = T^I e1 (run by every
// member) agree bit for bit and all members leave the loop at the same step.
//
// Exchange without flags or fences: a published FP64 value is two 64-bit
// words, each holding a 32-bit tag and one 32-bit half of the value, stored
// and loaded with device-scope (relaxed atomic) accesses; a reader polls the
// words until both carry the tag it expects.  tag = launch generation << 10
// | exchange epoch, so words left by an earlier epoch or launch never match
// and the buffer needs no reset between launches (zero at allocation; the
// generation starts at 1).  A slot is rewritten three epochs later at the
// earliest, and only after every member has read it (each member writes
// the next epoch's values only after reading every member's values of the
// epoch before).  Readers wait with a wall-clock bound; a member that times
// out marks the cluster failed (fail word = generation) and leaves, and the
// cluster then reports "not accepted" (its caller runs the plain power
// iteration).
constexpr int kLzCoopThreads = 256;
constexpr int kLzCoopWaves = kLzCoopThreads / kWave;
constexpr int kLzCoopLQ = kLzMax + 1;    // basis columns per row (odd: spreads LDS banks)
constexpr int kLzCoopPart = kLzMax + 2;  // partials per member: dots l <= it, then |w|^2
constexpr int kLzCoopEpochBits = 10;
// The second Gram-Schmidt pass runs when the first left less than this
// fraction of |w|^2: one pass leaves components of about eps |w| along the
// basis, i.e. eps |w| / |w - Q c| <= 100 eps relative to the new vector.
constexpr double kLzCoopReorth = 1e-4;

__device__ __forceinline__ double coop_sum(double v, double* red) {
  v = wave_sum_f64(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

// Two block sums with one pair of barriers.
__device__ __forceinline__ void coop_sum2(double& a, double& b, double* red) {
  a = wave_sum_f64(a);
  b = wave_sum_f64(b);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = a;
    red[4 + (threadIdx.x >> 6)] = b;
  }
  __syncthreads();
  a = (red[0] + red[1]) + (red[2] + red[3]);
  b = (red[4] + red[5]) + (red[6] + red[7]);
}

// Block max with one pair of barriers.
__device__ __forceinline__ double coop_max(double v, double* red4) {
  v = wave_max_f64(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red4[threadIdx.x >> 6] = v;
  __syncthreads();
  return fmax(fmax(red4[0], red4[1]), fmax(red4[2], red4[3]));
}

// Publishes v in the slot pair at p (tag in the upper halves).
__device__ __forceinline__ void put_tagged(uint64_t* p, double v, uint32_t tag) {
  const uint64_t t = (uint64_t)tag << 32;
  st_dev_u64(p, t | (uint32_t)__double2hiint(v));
  st_dev_u64(p + 1, t | (uint32_t)__double2loint(v));
}

// Reads the n <= N values published at the slot pairs p + 2 u stride
// (u < n) into v[u], polling until every word carries `tag`.  All loads of a
// round are in flight together (N slots, the ones past n repeat the last
// value, so N should be the smallest that covers n).  False on timeout or
// when the cluster was marked failed.
template <int N>
__device__ __forceinline__ bool poll_tagged(const uint64_t* p, long stride, int n, uint32_t tag,
                                            double* v, const uint32_t* fail, uint32_t gen,
                                            long long timeout) {
  if (n <= 0) return true;
  if (timeout < 0) return false;  // a negative limit: give up at once
  const long long t0 = wall_clock64();
  for (int round = 1;; ++round) {
    uint64_t h[N], l[N];
#pragma unroll
    for (int u = 0; u < N; ++u) {
      const long o = 2 * (long)(u < n ? u : n - 1) * stride;
      h[u] = ld_dev_u64(p + o);
      l[u] = ld_dev_u64(p + o + 1);
    }
    bool all = true;
#pragma unroll
    for (int u = 0; u < N; ++u)
      all = all && (uint32_t)(h[u] >> 32) == tag && (uint32_t)(l[u] >> 32) == tag;
    if (all) {
#pragma unroll
      for (int u = 0; u < N; ++u)
        if (u < n) v[u] = __hiloint2double((int)(uint32_t)h[u], (int)(uint32_t)l[u]);
      return true;
    }
    // the failure word and the clock every 8 rounds (a device-scope load
    // per round would double the round time)
    if ((round & 7) == 0 &&
        (__hip_atomic_load(const_cast<uint32_t*>(fail), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT) == gen ||
         wall_clock64() - t0 > timeout))
      return false;
    __builtin_amdgcn_s_sleep(1);
  }
}

// Sum in member order of the values published by the G members at
// p + 2 g stride (g < G), polled up to 16 at a time.
__device__ __forceinline__ bool poll_sum(const uint64_t* p, long stride, int G, uint32_t tag,
                                         const uint32_t* fail, uint32_t gen, long long timeout,
                                         double& sum) {
  double s = 0.0;
  for (int g0 = 0; g0 < G; g0 += 16) {
    double v[16];
    const int n = G - g0 < 16 ? G - g0 : 16;
    const uint64_t* q = p + 2 * (long)g0 * stride;
    const bool ok = n <= 4    ? poll_tagged<4>(q, stride, n, tag, v, fail, gen, timeout)
                    : n <= 8  ? poll_tagged<8>(q, stride, n, tag, v, fail, gen, timeout)
                              : poll_tagged<16>(q, stride, n, tag, v, fail, gen, timeout);
    if (!ok) return false;
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (u < n) s += v[u];
  }
  sum = s;
  return true;
}

// Reads this thread's rows tid + 256 u (u < n <= 16) of a published vector
// (slot pairs from p), with the smallest polling batch that covers n.
__device__ __forceinline__ bool poll_rows(const uint64_t* p, int n, uint32_t tag, double* v,
                                          const uint32_t* fail, uint32_t gen, long long timeout) {
  const uint64_t* q = p + 2L * threadIdx.x;
  if (n <= 1) return poll_tagged<1>(q, kLzCoopThreads, n, tag, v, fail, gen, timeout);
  if (n <= 2) return poll_tagged<2>(q, kLzCoopThreads, n, tag, v, fail, gen, timeout);
  if (n <= 4) return poll_tagged<4>(q, kLzCoopThreads, n, tag, v, fail, gen, timeout);
  if (n <= 8) return poll_tagged<8>(q, kLzCoopThreads, n, tag, v, fail, gen, timeout);
  return poll_tagged<16>(q, kLzCoopThreads, n, tag, v, fail, gen, timeout);
}

// Row stride of the member's S rows in LDS: odd, so that the lanes of a
// wave, which read one column of consecutive rows, hit distinct banks.
__host__ __device__ constexpr int lz_coop_sl_stride(int m) { return m | 1; }

// out[e] = sum_c S[e][c] x[c] for the member's nr <= 64 rows held in LDS
// (row stride st), FP64 accumulation.  Thread (s, e): row e, column range s
// of ns; the ns partials are summed in LDS (tmp) in part order.
__device__ __forceinline__ void coop_rows_matvec_lds(const float* Sl, int st, int m, int nr,
                                                     const double* x, double* tmp, double* out) {
  const int Rp = nr > 32 ? 64 : 32;
  const int ns = kLzCoopThreads / Rp;
  const int e = threadIdx.x % Rp, s = threadIdx.x / Rp;
  const int seg = (m + ns - 1) / ns;
  const int c0 = s * seg, c1 = c0 + seg < m ? c0 + seg : m;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  if (e < nr) {
    const float* row = Sl + (long)e * st;
    int c = c0;
    for (; c + 3 < c1; c += 4) {
      a0 += (double)row[c] * x[c];
      a1 += (double)row[c + 1] * x[c + 1];
      a2 += (double)row[c + 2] * x[c + 2];
      a3 += (double)row[c + 3] * x[c + 3];
    }
    for (; c < c1; ++c) a0 += (double)row[c] * x[c];
  }
  tmp[s * Rp + e] = (a0 + a1) + (a2 + a3);
  __syncthreads();
  for (int i = threadIdx.x; i < nr; i += kLzCoopThreads) {
    double t = 0.0;
    for (int ss = 0; ss < ns; ++ss) t += tmp[ss * Rp + i];
    out[i] = t;
  }
}

// out[e] = sum_c S[e][c] x[c] for the member's nr rows in global memory (row
// stride m), FP64 accumulation; a wave takes 4 rows at a time.
__device__ __forceinline__ void coop_rows_matvec(const float* __restrict__ Sr, int m, int nr,
                                                 const double* x, double* out) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int e0 = wid * 4; e0 < nr; e0 += kLzCoopWaves * 4) {
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    const float* row[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) row[u] = Sr + (long)(e0 + u < nr ? e0 + u : nr - 1) * m;
    for (int c = lane; c < m; c += kWave) {
      const double xc = x[c];
#pragma unroll
      for (int u = 0; u < 4; ++u) a[u] += (double)row[u][c] * xc;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double s = wave_sum_f64(a[u]);
      if (lane == 0 && e0 + u < nr) out[e0 + u] = s;
    }
  }
}

// Publishes dst[l] = sum_{e < nr} Qr[e][l] w[e] for l <= it and dst[it + 1]
// = |w|^2 over the member's rows.  Thread (g, l): column l, rows g, g + ng,
// ...; the ng row-group partials are then summed in LDS (tmp:
// kLzCoopThreads doubles) in group order.
__device__ __forceinline__ void coop_partials(const double* Qr, const double* w, int nr, int it,
                                              double* tmp, uint64_t* dst, uint32_t tag) {
  const int nc = it + 2;
  int ncp = 16;
  while (ncp < nc) ncp <<= 1;
  const int ng = kLzCoopThreads / ncp;
  const int l = threadIdx.x % ncp, g = threadIdx.x / ncp;
  double a0 = 0.0, a1 = 0.0;
  if (l <= it) {
    int e = g;
    for (; e + ng < nr; e += 2 * ng) {
      a0 += Qr[(long)e * kLzCoopLQ + l] * w[e];
      a1 += Qr[(long)(e + ng) * kLzCoopLQ + l] * w[e + ng];
    }
    if (e < nr) a0 += Qr[(long)e * kLzCoopLQ + l] * w[e];
  } else if (l == it + 1) {
    for (int e = g; e < nr; e += ng) a0 += w[e] * w[e];
  }
  tmp[g * ncp + l] = a0 + a1;
  __syncthreads();
  for (int c = threadIdx.x; c < nc; c += kLzCoopThreads) {
    double t = 0.0;
    for (int gg = 0; gg < ng; ++gg) t += tmp[gg * ncp + c];
    put_tagged(dst + 2 * c, t, tag);
  }
}

// w[e] -= sum_{l <= it} cf[l] Qr[e][l] on the member's rows (R <= 64).
// Thread (s, e): row e, columns s, s + ns, ...; the ns partials are summed
// in LDS (tmp) in part order.  Ends with a barrier.
__device__ __forceinline__ void coop_project(const double* Qr, const double* cf, int nr, int it,
                                             double* tmp, double* w) {
  const int Rp = nr > 32 ? 64 : 32;
  const int ns = kLzCoopThreads / Rp;
  const int e = threadIdx.x % Rp, s = threadIdx.x / Rp;
  double a0 = 0.0, a1 = 0.0;
  if (e < nr) {
    const double* qe = Qr + (long)e * kLzCoopLQ;
    int l = s;
    for (; l + ns <= it; l += 2 * ns) {
      a0 += cf[l] * qe[l];
      a1 += cf[l + ns] * qe[l + ns];
    }
    if (l <= it) a0 += cf[l] * qe[l];
  }
  tmp[s * Rp + e] = a0 + a1;
  __syncthreads();
  for (int i = threadIdx.x; i < nr; i += kLzCoopThreads) {
    double t = 0.0;
    for (int ss = 0; ss < ns; ++ss) t += tmp[ss * Rp + i];
    w[i] -= t;
  }
  __syncthreads();
}

// Exchange words per cluster: partials P1, P2 (G x kLzCoopPart pairs each),
// rows of r / w (m pairs), end marks (G words).
__host__ __device__ constexpr long lz_coop_xch_len(int m, int G) {
  return 2L * (2L * G * kLzCoopPart + m) + G;
}
__host__ __device__ constexpr long lz_coop_lds_doubles(int m, int R) {
  return 2L * m + (long)R * kLzCoopLQ + R + kLzCoopPart + 2L * kLzCoopLQ + 2L * kLzMax + 16 +
         kLzCoopThreads + kLzSqDoubles;
}

template <bool kSLds>
__global__ __launch_bounds__(kLzCoopThreads) void k_lz_coop(
    const float* __restrict__ S, int m, int k, int G, int R, const int* __restrict__ counts,
    uint64_t seed, int iter, int I, double lz_tol, uint64_t* __restrict__ xch,
    uint32_t* __restrict__ fails, uint32_t gen, long long timeout, double* __restrict__ vout,
    double* __restrict__ rout, int* __restrict__ steps_out, int* __restrict__ conv_out) {
  extern __shared__ double cz[];
  double* q = cz;                                // m: q_it, every row
  double* r = q + m;                             // m: r = S 1, every row
  double* Qr = r + m;                            // R x kLzCoopLQ: the member's basis rows
  double* w = Qr + (long)R * kLzCoopLQ;          // R
  double* cf = w + R;                            // kLzCoopPart
  double* alpha = cf + kLzCoopPart;              // kLzCoopLQ
  double* beta = alpha + kLzCoopLQ;              // kLzCoopLQ
  double* y = beta + kLzCoopLQ;                  // kLzMax
  double* yp = y + kLzMax;                       // kLzMax
  double* red = yp + kLzMax;                     // 8 (block sums)
  int* ok = reinterpret_cast<int*>(red + 8);     // 2 ints
  double* tmp = red + 16;                        // kLzCoopThreads: partial sums
  double* sq = tmp + kLzCoopThreads;             // kLzSqDoubles: acceptance test
  float* Sl = reinterpret_cast<float*>(sq + kLzSqDoubles);  // R x lz_coop_sl_stride(m)
  // member -> (cluster, index): consecutive blocks go to consecutive XCDs,
  // so numbering XCD by XCD puts the members of a cluster on one XCD
  const int per = gridDim.x / 8;
  const int p = (blockIdx.x % 8) * per + blockIdx.x / 8;
  const int j = p / G, b = p % G;
  if (j >= k || counts[j] == 0) return;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int e0 = b * R;
  const int nr = m - e0 < R ? m - e0 : R;
  const float* Sg = S + (long)j * m * m + (long)e0 * m;  // the member's rows
  uint64_t* X = xch + (long)j * lz_coop_xch_len(m, G);
  uint64_t* P1 = X;                               // [G][kLzCoopPart] pairs
  uint64_t* P2 = P1 + 2L * G * kLzCoopPart;
  uint64_t* Wx = P2 + 2L * G * kLzCoopPart;       // m pairs: rows of r, then of w
  uint64_t* Done = Wx + 2L * m;                   // G words
  uint32_t* fail = fails + j;
  const uint32_t tag0 = gen << kLzCoopEpochBits;
  const int sst = lz_coop_sl_stride(m);
  auto matvec = [&](const double* x, double* out) {
    if constexpr (kSLds)
      coop_rows_matvec_lds(Sl, sst, m, nr, x, tmp, out);
    else
      coop_rows_matvec(Sg, m, nr, x, out);
  };
  int len = 0, conv = 0;
  auto give_up = [&]() {
    if (tid == 0) __hip_atomic_store(fail, gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (b == 0 && tid == 0) {
      steps_out[j] = len;
      conv_out[j] = 0;
    }
  };
  // every thread's polls must have succeeded (the reduction is a barrier)
  auto all_ok = [&](bool good) { return __syncthreads_or(!good) == 0; };
  if constexpr (kSLds)
    for (long i = tid; i < (long)nr * m; i += kLzCoopThreads) {
      const long e = i / m, c = i - e * m;
      Sl[e * sst + c] = Sg[i];
    }
  // r = S 1 on the member's rows, exchanged; q_0 = power_v0 / |power_v0|
  for (int i = tid; i < m; i += kLzCoopThreads) q[i] = 1.0;
  __syncthreads();
  matvec(q, w);
  __syncthreads();
  uint32_t tag = tag0 | 1;
  for (int e = tid; e < nr; e += kLzCoopThreads) {
    put_tagged(Wx + 2L * (e0 + e), w[e], tag);
    rout[(long)j * m + e0 + e] = w[e];
  }
  double n0 = 0.0;
  for (int i = tid; i < m; i += kLzCoopThreads) {
    const double v = power_v0(seed, iter, j, i);
    n0 += v * v;
  }
  n0 = sqrt(coop_sum(n0, red));
  for (int i = tid; i < m; i += kLzCoopThreads) q[i] = power_v0(seed, iter, j, i) / n0;
  {
    double v[16];
    const int n = tid < m ? (m - 1 - tid) / kLzCoopThreads + 1 : 0;  // rows tid + 256 u
    const bool good = poll_rows(Wx, n, tag, v, fail, gen, timeout);
    if (!all_ok(good)) return give_up();
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (u < n) r[tid + u * kLzCoopThreads] = v[u];
  }
  double sg = 0.0;
  for (int i = tid; i < m; i += kLzCoopThreads) sg += r[i];
  const double sigma = coop_sum(sg, red);
  for (int e = tid; e < nr; e += kLzCoopThreads) Qr[(long)e * kLzCoopLQ] = q[e0 + e];
  if (tid == 0) beta[0] = 0.0;
  __syncthreads();
  const int jmax = m < kLzMax ? m : kLzMax;
  const double inv_m = 1.0 / (double)m;
  int prev_len = 0;
  for (int it = 0; it < jmax; ++it) {
    // A
    double sv = 0.0, rv = 0.0;
    for (int i = tid; i < m; i += kLzCoopThreads) {
      sv += q[i];
      rv += r[i] * q[i];
    }
    coop_sum2(sv, rv, red);
    matvec(q, w);
    __syncthreads();
    for (int e = tid; e < nr; e += kLzCoopThreads)
      w[e] = w[e] - (r[e0 + e] * sv + rv) * inv_m + sigma * sv * inv_m * inv_m;
    __syncthreads();
    tag = tag0 | (uint32_t)(2 + 3 * it);
    coop_partials(Qr, w, nr, it, tmp, P1 + 2L * b * kLzCoopPart, tag);
    // B
    bool good = true;
    for (int l = tid; l <= it + 1; l += kLzCoopThreads)
      good = good && poll_sum(P1 + 2L * l, kLzCoopPart, G, tag, fail, gen, timeout, cf[l]);
    if (!all_ok(good)) return give_up();
    const double nw2 = cf[it + 1];
    double a = cf[it];
    // |w - Q c|^2 = |w|^2 - |c|^2 over the orthonormal basis; the same sum
    // in the same order in every thread and member, so all take the same
    // branch
    double cc = 0.0;
    for (int l = 0; l <= it; ++l) cc += cf[l] * cf[l];
    const bool again = !(nw2 - cc >= kLzCoopReorth * nw2);
    coop_project(Qr, cf, nr, it, tmp, w);
    if (again) {
      // C
      tag = tag0 | (uint32_t)(3 + 3 * it);
      coop_partials(Qr, w, nr, it, tmp, P2 + 2L * b * kLzCoopPart, tag);
      for (int l = tid; l <= it + 1; l += kLzCoopThreads)
        good = good && poll_sum(P2 + 2L * l, kLzCoopPart, G, tag, fail, gen, timeout, cf[l]);
      if (!all_ok(good)) return give_up();
      a += cf[it];
      coop_project(Qr, cf, nr, it, tmp, w);
    }
    tag = tag0 | (uint32_t)(4 + 3 * it);
    for (int e = tid; e < nr; e += kLzCoopThreads) put_tagged(Wx + 2L * (e0 + e), w[e], tag);
    // D: every member's rows of w; |w|^2 from them (the same block sum in
    // every member)
    double wv[16];
    const int nw = tid < m ? (m - 1 - tid) / kLzCoopThreads + 1 : 0;
    good = poll_rows(Wx, nw, tag, wv, fail, gen, timeout);
    if (!all_ok(good)) return give_up();
    double bb = 0.0;
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (u < nw) bb += wv[u] * wv[u];
    const double bnorm = sqrt(coop_sum(bb, red));
    if (tid == 0) alpha[it] = a;
    len = it + 1;
    const bool breakdown = !(bnorm > 1e-12 * sqrt(nw2));
    if (breakdown || len == jmax || lz_power_check(len)) {
      __syncthreads();  // alpha[it] visible
      if (len <= kLzSqMax) {
        tridiag_power_block<kLzCoopThreads>(alpha, beta, len, I, sq, y);
      } else {
        if (wid == 0) tridiag_power_wave(alpha, beta, len, I, y);
        __syncthreads();
      }
      if (wid == 0) {
        double d = 0.0;
        for (int e = lane; e < len; e += 64) {
          const double pv = e < prev_len ? yp[e] : 0.0;  // previous y, zero padded
          const double dd = y[e] - pv;
          d += dd * dd;
        }
        d = wave_sum_f64(d);
        if (lane == 0) ok[1] = breakdown || (prev_len > 0 && sqrt(d) <= lz_tol);
        for (int e = lane; e < len; e += 64) yp[e] = y[e];
      }
      prev_len = len;
      __syncthreads();
      conv = ok[1];
      if (conv || breakdown) break;
    }
    if (len == jmax) break;
    if (tid == 0) beta[it + 1] = bnorm;
    const double inv_b = 1.0 / bnorm;
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (u < nw) q[tid + u * kLzCoopThreads] = wv[u] * inv_b;
    __syncthreads();
    for (int e = tid; e < nr; e += kLzCoopThreads)
      Qr[(long)e * kLzCoopLQ + it + 1] = q[e0 + e];
    __syncthreads();
  }
  // v = Q y on the member's rows
  for (int e = tid; e < nr; e += kLzCoopThreads) {
    double acc = 0.0;
    if (conv)
      for (int l = 0; l < len; ++l) acc += y[l] * Qr[(long)e * kLzCoopLQ + l];
    vout[(long)j * m + e0 + e] = acc;
  }
  // member 0 reports the cluster once every member has written its rows
  tag = tag0 | (uint32_t)(5 + 3 * (len - 1));
  __syncthreads();
  if (tid == 0) st_dev_u64(Done + b, (uint64_t)tag << 32);
  if (b == 0) {
    bool good = true;
    if (tid < kWave && timeout < 0) good = false;  // a negative limit: give up at once
    if (tid < kWave && timeout >= 0) {
      const long long t0 = wall_clock64();
      for (;;) {
        bool mine = true;
        for (int g = tid; g < G; g += kWave) mine = mine && (uint32_t)(ld_dev_u64(Done + g) >> 32) == tag;
        if (__all(mine)) break;
        if (__hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen ||
            wall_clock64() - t0 > timeout) {
          good = false;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    if (tid == 0) {
      steps_out[j] = len;
      conv_out[j] = good ? conv : 0;
    }
  }
}

// ---------------------------------------------------------------- launchers
// --------------------------------------------------------- grouped forms
// The kernels above for many jobs in one launch (GroupTable): each block
// finds its job and runs the body as that job's block bid of nblk.
#define TCX_GROUP_JOB(T)                                                       \
  const int gj_ = group_job(first, njobs, blockIdx.x);                         \
  const T a = jobs[gj_];                                                       \
  const int bid = blockIdx.x - first[gj_], nblk = first[gj_ + 1] - first[gj_];

// Fills of many segments: blockIdx.y is the segment.
__global__ void k_fill_segs(const FillSeg* __restrict__ segs) {
  const FillSeg sg = segs[blockIdx.y];
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < sg.count;
       i += (long)gridDim.x * blockDim.x)
    sg.p[i] = sg.value;
}

__global__ __launch_bounds__(256) void k_group_count_group(const GroupCountJob* __restrict__ jobs,
                                                           const int* __restrict__ first,
                                                           int njobs) {
  TCX_GROUP_JOB(GroupCountJob)
  group_count_body(bid, nblk, a.label, a.n, a.k, a.dist, a.B, a.blk_count, a.minrow, a.lds_min,
                   a.bad);
}

__global__ __launch_bounds__(256) void k_group_scan_group(const GroupScanJob* __restrict__ jobs,
                                                          const int* __restrict__ first,
                                                          int njobs) {
  TCX_GROUP_JOB(GroupScanJob)
  group_scan_body(bid, nblk, a.nrowblk, a.K, a.blk_count, a.blk_base, a.counts);
}

__global__ __launch_bounds__(1024) void k_group_offsets_group(
    const GroupOffsetsJob* __restrict__ jobs, const int* __restrict__ first, int njobs) {
  TCX_GROUP_JOB(GroupOffsetsJob)
  (void)bid;
  (void)nblk;
  group_offsets_body(a.k, a.B, a.kcounts, a.counts, a.offsets, a.koff);
}

__global__ __launch_bounds__(kGroupChunk) void k_group_place_group(
    const GroupPlaceJob* __restrict__ jobs, const int* __restrict__ first, int njobs) {
  TCX_GROUP_JOB(GroupPlaceJob)
  group_place_par_body(bid, nblk, a.label, a.n, a.k, a.dist, a.B, a.blk_base, a.koff, a.perm);
}

__global__ __launch_bounds__(256) void k_row_stats_group(const RowStatsJob* __restrict__ jobs,
                                                         const int* __restrict__ first,
                                                         int njobs) {
  TCX_GROUP_JOB(RowStatsJob)
  row_stats_body(bid, nblk, a.x, a.rows, a.m, a.stats);
}

__global__ __launch_bounds__(256) void k_align_pack_group(const AlignPackJob* __restrict__ jobs,
                                                          const int* __restrict__ first,
                                                          int njobs) {
  TCX_GROUP_JOB(AlignPackJob)
  align_pack_body(bid, nblk, a.x, a.m, a.perm, a.n_pad, a.offsets, a.minrow, a.k, a.shift,
                  a.czero, a.stats, a.yhi, a.ylo, a.a0);
}

template <int R>
__global__ __launch_bounds__(256) void k_align_pack_lds_group(
    const AlignPackJob* __restrict__ jobs, const int* __restrict__ first, int njobs) {
  TCX_GROUP_JOB(AlignPackJob)
  align_pack_lds_body<R>(bid, nblk, a.x, a.m, a.perm, a.offsets, a.minrow, a.k, a.shift,
                         a.czero, a.yhi, a.ylo, a.a0);
}

// One item (a kSyrkSub block of one job's cluster over the cluster's
// columns) per wave: items[4 * block + wave], the waves past nitems idle.
template <bool kSplit>
__global__ __launch_bounds__(256) void k_syrk_group(const SyrkJob* __restrict__ jobs,
                                                    const GroupSyrkItem* __restrict__ items,
                                                    int nitems) {
  const int i = __builtin_amdgcn_readfirstlane((int)blockIdx.x * 4 + (int)(threadIdx.x >> 6));
  if (i >= nitems) return;
  const GroupSyrkItem gi = items[i];
  const SyrkJob jb = jobs[gi.job];
  syrk_wave<kSplit>(gi.it.ti * 4, gi.it.tj * 4, gi.it.p0 >> 4, gi.it.p1 >> 4, panel_blocks(jb.m),
                    jb.yhi, jb.ylo, jb.part + (long)gi.it.part * kSyrkSub * kSyrkSub, kSyrkSub);
}

__global__ __launch_bounds__(256) void k_syrk_reduce_group(const SyrkJob* __restrict__ jobs,
                                                           const GroupSyrkTile* __restrict__ tiles) {
  constexpr int kSub = kSyrkSub / 64;
  const GroupSyrkTile* gt = tiles + blockIdx.x / (kSub * kSub);
  const SyrkJob jb = jobs[gt->job];
  syrk_reduce_body(blockIdx.x % (kSub * kSub), gt->t, jb.part, jb.m, jb.S);
}

__global__ __launch_bounds__(kPowerThreads) void k_power_lanczos_group(
    const PowerLanczosJob* __restrict__ jobs, const int* __restrict__ first, int njobs,
    double lz_tol) {
  TCX_GROUP_JOB(PowerLanczosJob)
  power_lanczos_body(bid, nblk, a.S, a.m, a.counts, a.seed, a.iter, a.I, lz_tol, a.Q, a.vout,
                     a.rout, a.steps_out, a.conv_out);
}

__global__ __launch_bounds__(kMvThreads) void k_cluster_matvec_group(
    const ClusterMatvecJob* __restrict__ jobs, const int* __restrict__ first, int njobs) {
  TCX_GROUP_JOB(ClusterMatvecJob)
  cluster_matvec_body(bid, nblk, a.S, a.m, a.k, a.x, a.xstride, a.y, a.state);
}

__global__ __launch_bounds__(kPowerThreads) void k_lz_setup_group(const LzJob* __restrict__ jobs,
                                                                  const int* __restrict__ first,
                                                                  int njobs) {
  TCX_GROUP_JOB(LzJob)
  lz_setup_body(bid, nblk, a.m, a.counts, a.mask, a.seed, a.iter, a.rbuf, a.Q, a.state);
  if (threadIdx.x == 0 && a.done_out) a.done_out[bid] = a.state[bid].done;  // thread 0 wrote it
}

__global__ __launch_bounds__(kPowerThreads) void k_lz_step_group(const LzJob* __restrict__ jobs,
                                                                 const int* __restrict__ first,
                                                                 int njobs, int mode,
                                                                 double lz_tol) {
  TCX_GROUP_JOB(LzJob)
  lz_step_body(bid, nblk, a.m, mode, a.I, lz_tol, a.rbuf, a.wbuf, a.Q, a.state);
  if (threadIdx.x == 0 && a.done_out) a.done_out[bid] = a.state[bid].done;  // thread 0 wrote it
}

__global__ __launch_bounds__(kPowerThreads) void k_lz_finish_group(const LzJob* __restrict__ jobs,
                                                                   const int* __restrict__ first,
                                                                   int njobs) {
  TCX_GROUP_JOB(LzJob)
  lz_finish_body(bid, nblk, a.m, a.counts, a.Q, a.state, a.vout, a.steps_out, a.conv_out);
}

__global__ __launch_bounds__(kPowerThreads) void k_power_group(const PowerJob* __restrict__ jobs,
                                                               const int* __restrict__ first,
                                                               int njobs) {
  TCX_GROUP_JOB(PowerJob)
  power_body(bid, nblk, a.S, a.m, a.counts, a.a0, a.seed, a.iter, a.max_steps, a.tol, a.vin,
             a.svin, a.rsum, a.steps_in, a.conv_in, a.eig_mode, a.mask, a.defer_fallback, a.c_out,
             a.info);
}

hipError_t launch_fill_int(int* p, long count, int value, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  long blocks = (count + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_int, dim3((unsigned)blocks), dim3(256), 0, st, p, count, value);
  return hipGetLastError();
}

int group_bins(long n, int k, bool have_dist) {
  if (!have_dist || k > kPlaceParMaxK) return 1;
  const long nblk = (n + kGroupChunk - 1) / kGroupChunk;
  int B = kGroupBins;
  while (B > 1 && ((long)k * B > kPlaceParMaxK || nblk * k * B > kGroupMaxBlkKeys)) B >>= 1;
  return B;
}

hipError_t launch_group_count(const int* label, long n, int k, const float* dist, int B,
                              int* blk_count, int* minrow, int* bad, hipStream_t st) {
  long nblk = (n + kGroupChunk - 1) / kGroupChunk;
  // per-block cluster minima in LDS while counts and minima fit 64 KB
  const size_t lds = sizeof(int) * ((size_t)k * B + k);
  const int lds_min = lds <= 65536;
  hipLaunchKernelGGL(k_group_count, dim3((unsigned)nblk), dim3(256),
                     lds_min ? lds : sizeof(int) * (size_t)k * B, st, label, n, k, dist, B,
                     blk_count, minrow, lds_min, bad);
  return hipGetLastError();
}

hipError_t launch_group_scan(int nblk, int k, int B, const int* blk_count, int* blk_base,
                             int* kcounts, int* counts, long* offsets, long* koff,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_group_scan, dim3(k * B), dim3(256), 0, st, nblk, k * B, blk_count,
                     blk_base, kcounts);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_group_offsets, dim3(1), dim3(1024), 0, st, k, B, kcounts, counts, offsets,
                     koff);
  return hipGetLastError();
}

hipError_t launch_group_place(const int* label, long n, int k, const float* dist, int B,
                              const int* blk_base, const long* koff, int* perm, hipStream_t st) {
  long nblk = (n + kGroupChunk - 1) / kGroupChunk;
  if ((long)k * B <= kPlaceParMaxK)
    hipLaunchKernelGGL(k_group_place_par, dim3((unsigned)nblk), dim3(kGroupChunk),
                       sizeof(int) * kPlaceWaves * k * B, st, label, n, k, dist, B, blk_base, koff,
                       perm);
  else if (B == 1)
    hipLaunchKernelGGL(k_group_place, dim3((unsigned)nblk), dim3(64), sizeof(int) * k, st,
                       label, n, k, blk_base, koff, perm);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_row_stats(const float* x, long rows, int m, double2* stats, hipStream_t st) {
  if (rows <= 0) return hipSuccess;
  const long blocks = (rows + 3) / 4;
  hipLaunchKernelGGL(k_row_stats, dim3((unsigned)blocks), dim3(256), 0, st, x, rows, m, stats);
  return hipGetLastError();
}

bool align_pack_uses_row_stats(int m) { return m > kPackLdsMaxM; }

hipError_t launch_align_pack(const float* x, int m, const int* perm, long n_pad,
                             const long* offsets, const int* minrow, int k, const int* shift,
                             const int* czero, const double2* stats, uint16_t* yhi,
                             uint16_t* ylo, float* a0, hipStream_t st) {
  if (n_pad <= 0) return hipSuccess;
  if (m <= kPackLdsMaxM) {
    const int R = align_pack_form(m);
    const size_t lds = sizeof(float) * (size_t)R * (m + 4);
    const dim3 grid((unsigned)(n_pad / R));
    if (R == 64)
      hipLaunchKernelGGL(k_align_pack_lds<64>, grid, dim3(256), lds, st, x, m, perm, offsets,
                         minrow, k, shift, czero, yhi, ylo, a0);
    else if (R == 32)
      hipLaunchKernelGGL(k_align_pack_lds<32>, grid, dim3(256), lds, st, x, m, perm, offsets,
                         minrow, k, shift, czero, yhi, ylo, a0);
    else
      hipLaunchKernelGGL(k_align_pack_lds<16>, grid, dim3(256), lds, st, x, m, perm, offsets,
                         minrow, k, shift, czero, yhi, ylo, a0);
    return hipGetLastError();
  }
  long blocks = n_pad / 64;
  hipLaunchKernelGGL(k_align_pack, dim3((unsigned)blocks), dim3(256), 0, st, x, m, perm, n_pad,
                     offsets, minrow, k, shift, czero, stats, yhi, ylo, a0);
  return hipGetLastError();
}

long syrk_blocks(int m, int ngroups) {
  const long T = (m + kSyrkSub - 1) / kSyrkSub, U = T * (T + 1) / 2;
  return 8 * (((ngroups + 7) / 8 * U + 3) / 4);
}

hipError_t launch_syrk(const SyrkGroup* groups, int ngroups, const uint16_t* yhi,
                       const uint16_t* ylo, int m, int split, float* part, hipStream_t st) {
  if (ngroups <= 0) return hipSuccess;
  const dim3 grid((unsigned)syrk_blocks(m, ngroups));
  if (split)
    hipLaunchKernelGGL(k_syrk<true>, grid, dim3(256), 0, st, groups, ngroups, yhi, ylo, m, part);
  else
    hipLaunchKernelGGL(k_syrk<false>, grid, dim3(256), 0, st, groups, ngroups, yhi, ylo, m, part);
  return hipGetLastError();
}

hipError_t launch_syrk_reduce(const SyrkCluster* clusters, int nclusters, const float* part, int m,
                              float* S, hipStream_t st) {
  if (nclusters <= 0) return hipSuccess;
  constexpr int kSub = kSyrkSub / 64;
  const long T = (m + kSyrkSub - 1) / kSyrkSub, U = T * (T + 1) / 2;
  hipLaunchKernelGGL(k_syrk_reduce, dim3((unsigned)(nclusters * U * kSub * kSub)), dim3(256), 0,
                     st, clusters, part, m, S);
  return hipGetLastError();
}

static size_t power_smem(int m) {
  return sizeof(double) * (3 * (size_t)m + 8 * kPowerWaves + sym_part_len(m));
}

hipError_t launch_power(const float* S, int m, int k, const int* counts, const float* a0,
                        uint64_t seed, int iter, int max_steps, double tol, const double* vin,
                        const double* svin, const double* rsum, const int* steps_in,
                        const int* conv_in, int eig_mode, const int* mask, int defer_fallback,
                        float* c_out, PowerInfo* info, hipStream_t st) {
  hipLaunchKernelGGL(k_power, dim3(k), dim3(kPowerThreads), power_smem(m), st, S, m, counts,
                     a0, seed, iter, max_steps, tol, vin, svin, rsum, steps_in, conv_in, eig_mode,
                     mask, defer_fallback, c_out, info);
  return hipGetLastError();
}

static size_t lanczos_smem(int m) {
  return sizeof(double) * (4 * (size_t)m + 8 * kPowerWaves + 5 * (kLzMax + 1) +
                           std::max((size_t)sym_part_len(m), (size_t)kLzSqDoubles));
}

hipError_t launch_power_lanczos(const float* S, int m, int k, const int* counts, uint64_t seed,
                                int iter, int I, double lz_tol, double* Q, double* vout,
                                double* rout, int* steps_out, int* conv_out, hipStream_t st) {
  hipLaunchKernelGGL(k_power_lanczos, dim3(k), dim3(kPowerThreads), lanczos_smem(m), st, S, m,
                     counts, seed, iter, I, lz_tol, Q, vout, rout, steps_out, conv_out);
  return hipGetLastError();
}

bool plan_lz_coop(int m, int k, int cus, LzCoopPlan* p) {
  if (m < kLzCoopMinM || k <= 0 || cus < 8) return false;
  int G = 1;
  while (G < 16 && 2L * G * k <= cus) G *= 2;
  int R = (m + G - 1) / G;
  if (R > kLzCoopMaxR) R = kLzCoopMaxR;
  G = (m + R - 1) / R;
  R = (m + G - 1) / G;  // rows balanced over the members
  if (G > kLzCoopMaxG || (long)(G - 1) * R >= m) return false;  // every member owns rows
  const size_t base = sizeof(double) * (size_t)lz_coop_lds_doubles(m, R);
  const size_t sbytes = sizeof(float) * (size_t)R * lz_coop_sl_stride(m);
  const bool s_lds = sbytes <= kLzCoopSLdsBytes && base + sbytes <= kLzCoopLdsMax;
  const size_t lds = base + (s_lds ? sbytes : 0);
  if (lds > kLzCoopLdsMax) return false;
  const long nb = ((long)k * G + 7) / 8 * 8;
  // every member must be resident at once
  int per_cu = 0;
  const void* fn = s_lds ? (const void*)k_lz_coop<true> : (const void*)k_lz_coop<false>;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kLzCoopThreads, lds) !=
      hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  if ((long)per_cu * cus < nb) return false;
  p->G = G;
  p->R = R;
  p->nb = (int)nb;
  p->s_lds = s_lds;
  p->lds = lds;
  p->xch_len = lz_coop_xch_len(m, G);
  return true;
}

hipError_t launch_lz_coop(const LzCoopPlan& p, const float* S, int m, int k, const int* counts,
                          uint64_t seed, int iter, int I, double lz_tol, uint64_t* xch,
                          uint32_t* fails, uint32_t gen, long long timeout, double* vout,
                          double* rout, int* steps_out, int* conv_out, hipStream_t st) {
  if (gen == 0 || gen >= (1u << (32 - kLzCoopEpochBits))) return hipErrorInvalidValue;
  if (p.s_lds)
    hipLaunchKernelGGL(k_lz_coop<true>, dim3(p.nb), dim3(kLzCoopThreads), p.lds, st, S, m, k, p.G,
                       p.R, counts, seed, iter, I, lz_tol, xch, fails, gen, timeout, vout, rout,
                       steps_out, conv_out);
  else
    hipLaunchKernelGGL(k_lz_coop<false>, dim3(p.nb), dim3(kLzCoopThreads), p.lds, st, S, m, k,
                       p.G, p.R, counts, seed, iter, I, lz_tol, xch, fails, gen, timeout, vout,
                       rout, steps_out, conv_out);
  return hipGetLastError();
}

static size_t lz_step_smem(int m) {
  return sizeof(double) * (4 * (size_t)m + 8 * kPowerWaves + (kLzMax + 1) + sym_part_len(m));
}

hipError_t launch_cluster_matvec(const float* S, int m, int k, const double* x, long xstride,
                                 double* y, const LzState* state, hipStream_t st) {
  if (k <= 0) return hipSuccess;
  const long panels = (m + kMvRows - 1) / kMvRows;
  hipLaunchKernelGGL(k_cluster_matvec, dim3((unsigned)(panels * k)), dim3(kMvThreads),
                     sizeof(double) * m, st, S, m, k, x, xstride, y, state);
  return hipGetLastError();
}

hipError_t launch_lz_setup(int m, int k, const int* counts, const int* mask, uint64_t seed,
                           int iter, const double* rbuf, double* Q, LzState* state,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_lz_setup, dim3(k), dim3(kPowerThreads), 0, st, m, counts, mask, seed, iter,
                     rbuf, Q, state);
  return hipGetLastError();
}

hipError_t launch_lz_step(int m, int k, int mode, int I, double lz_tol, const double* rbuf,
                          const double* wbuf, double* Q, LzState* state, hipStream_t st) {
  hipLaunchKernelGGL(k_lz_step, dim3(k), dim3(kPowerThreads), lz_step_smem(m), st, m, mode, I,
                     lz_tol, rbuf, wbuf, Q, state);
  return hipGetLastError();
}

hipError_t launch_lz_finish(int m, int k, const int* counts, const double* Q,
                            const LzState* state, double* vout, int* steps_out, int* conv_out,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_lz_finish, dim3(k), dim3(kPowerThreads), 0, st, m, counts, Q, state, vout,
                     steps_out, conv_out);
  return hipGetLastError();
}

// ------------------------------------------------ grouped-form launchers
hipError_t launch_fill_segs(const FillSeg* segs, int nseg, long max_count, hipStream_t st) {
  if (nseg <= 0 || max_count <= 0) return hipSuccess;
  long blocks = (max_count + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  for (int s0 = 0; s0 < nseg; s0 += 65535) {  // grid y limit
    const int ns = nseg - s0 < 65535 ? nseg - s0 : 65535;
    hipLaunchKernelGGL(k_fill_segs, dim3((unsigned)blocks, (unsigned)ns), dim3(256), 0, st,
                       segs + s0);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

size_t group_count_lds(int k, int B, int* lds_min) {
  const size_t lds = sizeof(int) * ((size_t)k * B + k);
  *lds_min = lds <= 65536;
  return *lds_min ? lds : sizeof(int) * (size_t)k * B;
}


// ------------------------------------------------ Gram form of thin clusters
// For a cluster of c members whose padded segment is shorter than m (and at
// most kGramMaxND), M = P Y^T Y P = Yc^T Yc (Yc = Y P, rows centred) is evaluated
// through the c x c Gram matrix G = Yc Yc^T = Y Y^T - s s^T / m (s = Y 1):
// M^I v0 = Yc^T G^(I-1) (Yc v0), and M v = Yc^T (G g) for v = Yc^T g.  The
// same algebra as the S form, on a matrix (m / c)^2 times smaller, so every
// Lanczos step of a thin cluster moves c^2 instead of m^2 values.
//   k_gram_tiles  G (order n = c rounded up to 4, the padding members zero)
//                 in 64 x 64 tiles of the segment from the packed panel
//                 (Y = hi + lo, exact in FP32), FP32 MFMA 32x32x2 with FP32
//                 accumulation,
//                 the s s^T / m term in the epilogue; entries with a <= b
//                 stored with their mirrors, so G is exactly symmetric
//   (start)       u0 = Y (v0 - mean(v0)) (FP64), v0 = power_v0(seed, iter,
//                 cluster, .), into the cluster's r row, formed by the
//                 diagonal tiles from the rows they load
//   (Lanczos)     power_lanczos_cluster in its Gram form: g = G^(I-1) u0
//                 and h = G g, into the cluster's w and sv rows
//   k_gram_map    v = P Y^T g, S v -> P Y^T h (= M v) and r = 0, the inputs
//                 the power kernel takes for the S form (r = 0: no further
//                 centring)
typedef float f32x16g __attribute__((ext_vector_type(16)));

__device__ __forceinline__ void panel_load8(const uint16_t* __restrict__ yhi,
                                            const uint16_t* __restrict__ ylo, long off,
                                            float* y) {
  const uint4 h = *reinterpret_cast<const uint4*>(yhi + off);
  const uint4 l = *reinterpret_cast<const uint4*>(ylo + off);
  const uint32_t hw[4] = {h.x, h.y, h.z, h.w}, lw[4] = {l.x, l.y, l.z, l.w};
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    y[2 * u] = __uint_as_float(hw[u] << 16) + __uint_as_float(lw[u] << 16);
    y[2 * u + 1] = __uint_as_float(hw[u] & 0xffff0000u) + __uint_as_float(lw[u] & 0xffff0000u);
  }
}

__global__ __launch_bounds__(256) void k_gram_tiles(const GramItem* __restrict__ items) {
  const GramItem it = items[blockIdx.x];
  constexpr int kLd = 64 + 4;
  __shared__ float As[32 * kLd], Bs[32 * kLd];
  __shared__ float ssA[64], ssB[64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool diag = it.ti == it.tj;
  const int tm = panel_blocks(it.m);
  // loader: chunk ch (16 members) of each side, lane ln: time ln % 32,
  // members 8 * (ln / 32) .. + 7 of the chunk
  const int ch = wave, tt = lane & 31, qo = ch * 16 + 8 * (lane >> 5);
  const bool va = it.ti * 64 + ch * 16 < it.nd, vb = it.tj * 64 + ch * 16 < it.nd;
  const long ca = (it.p0 >> 4) + it.ti * 4 + ch, cb = (it.p0 >> 4) + it.tj * 4 + ch;
  float sa[8], sb[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) sa[e] = sb[e] = 0.f;
  const int wi = wave >> 1, wj = wave & 1;
  const bool active = !(diag && wi > wj);
  f32x16g acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  // the start vector of the tile's members (diagonal tiles with u0): the
  // mean of v0 over the m times, then u0[a] = sum_t y_a[t] (v0[t] - mean)
  __shared__ double vred[4];
  double vmean = 0.0;
  double su[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) su[e] = 0.0;
  const bool start = it.u0 != nullptr;  // uniform: the whole block
  if (start) {
    double t0 = 0.0;
    for (int t = tid; t < it.m; t += 256) t0 += power_v0(it.seed, it.iter, it.cluster, t);
    t0 = wave_sum_f64(t0);
    if (lane == 0) vred[wave] = t0;
    __syncthreads();
    vmean = ((vred[0] + vred[1]) + (vred[2] + vred[3])) / (double)it.m;
  }
  for (int r32 = 0; r32 < tm; ++r32) {
    float ya[8], yb[8];
    if (va) {
      panel_load8(it.yhi, it.ylo, (ca * tm + r32) * 512 + lane * 8, ya);
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) ya[e] = 0.f;
    }
    if (!diag && vb) {
      panel_load8(it.yhi, it.ylo, (cb * tm + r32) * 512 + lane * 8, yb);
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) yb[e] = diag ? ya[e] : 0.f;
    }
    __syncthreads();  // the previous step's operands are consumed
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      As[tt * kLd + qo + e] = ya[e];
      Bs[tt * kLd + qo + e] = yb[e];
      sa[e] += ya[e];
      sb[e] += yb[e];
    }
    if (start) {
      const int t = r32 * 32 + tt;
      if (t < it.m) {  // rows t >= m of the panel are zero
        const double vc = power_v0(it.seed, it.iter, it.cluster, t) - vmean;
#pragma unroll
        for (int e = 0; e < 8; ++e) su[e] += (double)ya[e] * vc;
      }
    }
    __syncthreads();
    if (active) {
      const float* ap = As + (lane >> 5) * kLd + wi * 32 + (lane & 31);
      const float* bp = Bs + (lane >> 5) * kLd + wj * 32 + (lane & 31);
#pragma unroll
      for (int kk = 0; kk < 16; ++kk)
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ap[2 * kk * kLd], bp[2 * kk * kLd], acc, 0, 0,
                                                   0);
    }
  }
  // row sums: lanes of one half-wave hold the same 8 members at 32 times
#pragma unroll
  for (int e = 0; e < 8; ++e) {
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
      sa[e] += __shfl_xor(sa[e], o);
      sb[e] += __shfl_xor(sb[e], o);
    }
  }
  if (tt == 0) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      ssA[qo + e] = sa[e];
      ssB[qo + e] = sb[e];
    }
  }
  if (start) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
#pragma unroll
      for (int o = 1; o < 32; o <<= 1) su[e] += __shfl_xor(su[e], o);
      const int a = it.ti * 64 + qo + e;
      if (tt == 0 && a < it.n) it.u0[a] = su[e];
    }
  }
  __syncthreads();
  if (!active) return;
  const float inv_m = 1.0f / (float)it.m;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = wi * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
    const int jj = wj * 32 + (lane & 31);
    const int a = it.ti * 64 + i, b = it.tj * 64 + jj;
    if (a < it.n && b < it.n && a <= b) {
      const float v = acc[r] - ssA[i] * ssB[jj] * inv_m;
      it.G[(long)a * it.n + b] = v;
      it.G[(long)b * it.n + a] = v;
    }
  }
}

// Sum over a block of 256 threads (4 waves); every thread gets the sum.
__device__ __forceinline__ double block256_sum_f64(double v, double* red) {
  v = wave_sum_f64(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(kPowerThreads) void k_gram_lanczos(const GramLzJob* __restrict__ jobs,
                                                                double lz_tol) {
  const GramLzJob a = jobs[blockIdx.x];
  power_lanczos_cluster(a.G, a.nd, a.u0, 0, 0, 0, a.I, lz_tol, a.Q, a.g, nullptr, a.h, a.steps,
                        a.conv);
}

__global__ __launch_bounds__(256) void k_gram_map(const GramVecJob* __restrict__ jobs) {
  const GramVecJob a = jobs[blockIdx.x];
  extern __shared__ double gsh[];
  double* g = gsh;          // nd
  double* h = g + a.nd;     // nd
  double* vt = h + a.nd;    // m
  double* st = vt + a.m;    // m
  __shared__ double red[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < a.nd; i += 256) {  // members n .. nd - 1 are padding
    g[i] = i < a.n ? a.g[i] : 0.0;
    h[i] = i < a.n ? a.h[i] : 0.0;
  }
  __syncthreads();
  const int tm = panel_blocks(a.m), nch = a.nd / 16;
  for (int r32 = wave; r32 < tm; r32 += 4) {
    const int t = r32 * 32 + (lane & 31);
    if (r32 * 32 >= a.m) break;
    double av = 0.0, as = 0.0;
    for (int cc = 0; cc < nch; ++cc) {
      float y[8];
      panel_load8(a.yhi, a.ylo, (((a.p0 >> 4) + cc) * (long)tm + r32) * 512 + lane * 8, y);
      const int q = cc * 16 + 8 * (lane >> 5);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        av += (double)y[e] * g[q + e];
        as += (double)y[e] * h[q + e];
      }
    }
    av += __shfl_xor(av, 32);
    as += __shfl_xor(as, 32);
    if (lane < 32 && t < a.m) {
      vt[t] = av;
      st[t] = as;
    }
  }
  __syncthreads();
  double sv = 0.0, ss = 0.0;
  for (int t = threadIdx.x; t < a.m; t += 256) {
    sv += vt[t];
    ss += st[t];
  }
  const double mv = block256_sum_f64(sv, red) / (double)a.m;
  const double ms = block256_sum_f64(ss, red) / (double)a.m;
  for (int t = threadIdx.x; t < a.m; t += 256) {
    a.v[t] = vt[t] - mv;
    a.sv[t] = st[t] - ms;
    a.r[t] = 0.0;
  }
}


// ------------------------------------------- one wave per small cluster
// The Lanczos evaluation of power_lanczos_cluster for dimension d <=
// kLzWaveMaxD by one wave, with no block barriers: vectors are held in
// registers (element e = lane + 64 u), the matvec reads column c of the
// symmetric matrix as one coalesced row segment with x[c] broadcast from
// the wave's LDS, sums are DPP wave sums, and T^I e1 is the one-wave recurrence
// (tridiag_power_wave).  Four clusters per block of 256 threads.  Same
// steps and acceptance rule as the block form (which takes the recurrence
// by binary powering for short T: the two agree to rounding).
constexpr int kLzWaveU = kLzWaveMaxD / 64;
constexpr int kLzWaveLds = 3 * (kLzMax + 1) + 2 * kLzMax + kLzWaveMaxD;  // alpha, beta, cf, y, yp, x

// y = S x for symmetric S (FP32, d x d) through the wave's LDS slot xb (d
// doubles); entries e >= d of y come back zero.  Loads past the last row
// are clamped to it.
// kW > 1: the kW waves of the block share one cluster, wave `part` takes
// the columns [part * d/kW, (part+1) * d/kW) and the partial products are
// summed through `red` (kW x kLzWaveMaxD doubles) in wave order, so every
// wave ends with the same y.
template <int kW>
__device__ __forceinline__ void wave_sym_matvec(const float* __restrict__ S, int d, int U,
                                                const double (&x)[kLzWaveU],
                                                double (&y)[kLzWaveU], double* xb, int part,
                                                double* red) {
  constexpr int kCols = 8;  // columns whose loads are in flight together
  const int lane = threadIdx.x & 63;
  int ec[kLzWaveU];
#pragma unroll
  for (int u = 0; u < kLzWaveU; ++u) {
    y[u] = 0.0;
    const int e = lane + 64 * u;
    ec[u] = e < d ? e : d - 1;
    if (e < d) xb[e] = x[u];
  }
  wave_lds_sync();
  const int share = (d + kW - 1) / kW;
  const int cb = kW > 1 ? part * share : 0, ce = kW > 1 ? (cb + share < d ? cb + share : d) : d;
  for (int c0 = cb; c0 < ce; c0 += kCols) {
    float sv[kCols][kLzWaveU];
#pragma unroll
    for (int t = 0; t < kCols; ++t) {
      const int c = c0 + t < ce ? c0 + t : ce - 1;
      const float* col = S + (long)c * d;
#pragma unroll
      for (int u = 0; u < kLzWaveU; ++u) sv[t][u] = u < U ? col[ec[u]] : 0.f;
    }
#pragma unroll
    for (int t = 0; t < kCols; ++t) {
      const double xc = c0 + t < ce ? xb[c0 + t] : 0.0;
#pragma unroll
      for (int u = 0; u < kLzWaveU; ++u) y[u] += (double)sv[t][u] * xc;
    }
  }
  if constexpr (kW > 1) {
#pragma unroll
    for (int u = 0; u < kLzWaveU; ++u) {
      const int e = lane + 64 * u;
      if (e < d) red[part * kLzWaveMaxD + e] = y[u];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kLzWaveU; ++u) {
      const int e = lane + 64 * u;
      double t = 0.0;
      if (e < d)
#pragma unroll
        for (int p = 0; p < kW; ++p) t += red[p * kLzWaveMaxD + e];
      y[u] = t;
    }
    __syncthreads();  // red is rewritten by the next product
  }
#pragma unroll
  for (int u = 0; u < kLzWaveU; ++u)
    if (lane + 64 * u >= d) y[u] = 0.0;
  wave_lds_sync();  // xb is rewritten by the next product
}

// Rows l0 .. l0 + kLzWaveRows - 1 of the basis (rows past `last`, and
// entries e >= d, read as zero), loaded together.
constexpr int kLzWaveRows = 8;
__device__ __forceinline__ void wave_basis_rows(const double* __restrict__ Q, int d, int U, int l0,
                                                int last,
                                                double (&qv)[kLzWaveRows][kLzWaveU]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int t = 0; t < kLzWaveRows; ++t) {
    const double* ql = Q + (long)(l0 + t) * d;
#pragma unroll
    for (int u = 0; u < kLzWaveU; ++u) {
      const int e = lane + 64 * u;
      qv[t][u] = (u < U && l0 + t <= last && e < d) ? ql[e] : 0.0;
    }
  }
}

// kW = 1: four clusters per block, one wave each.  kW = 4: one cluster per
// block, the four waves sharing each matvec by columns (the product's
// latency chain is a quarter as long) and computing everything else alike;
// wave 0 writes the results, every wave the (identical) basis rows it reads
// back itself.
template <int kW>
__global__ __launch_bounds__(256, 3) void k_lanczos_wave(const LzWaveJob* __restrict__ jobs,
                                                      int njobs, double lz_tol) {
  __shared__ double lds[4][kLzWaveLds];
  __shared__ double red[kW > 1 ? kW * kLzWaveMaxD : 1];
  // the wave index is uniform in the wave: readfirstlane lets the job's
  // fields live in scalar registers
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int jid = kW > 1 ? (int)blockIdx.x : (int)blockIdx.x * 4 + wv;
  if (jid >= njobs) return;  // whole wave (kW = 1) or whole block
  const int part = kW > 1 ? wv : 0;
  const bool writer = part == 0;
  const LzWaveJob a = jobs[jid];
  double* alpha = lds[wv];
  double* beta = alpha + kLzMax + 1;
  double* cf = beta + kLzMax + 1;
  double* y = cf + kLzMax + 1;
  double* yp = y + kLzMax;
  double* xb = yp + kLzMax;
  const int d = a.d, U = (d + 63) / 64;
  const float* S = a.S;
  double q[kLzWaveU], w[kLzWaveU], qp[kLzWaveU], r[kLzWaveU];
  double sigma = 0.0, n0 = 0.0;
  const double inv_d = 1.0 / (double)d;
  if (a.gram) {
#pragma unroll
    for (int u = 0; u < kLzWaveU; ++u) {
      const int e = lane + 64 * u;
      q[u] = e < d ? a.vstart[e] : 0.0;
      r[u] = qp[u] = 0.0;
      n0 += q[u] * q[u];
    }
    n0 = sqrt(wave_sum_f64(n0));
    if (!(n0 > 0.0)) {  // wave-uniform
#pragma unroll
      for (int u = 0; u < kLzWaveU; ++u) {
        const int e = lane + 64 * u;
        if (e < d && writer) {
          a.v[e] = 0.0;
          a.h[e] = 0.0;
        }
      }
      if (lane == 0 && writer) {
        *a.steps = 0;
        *a.conv = 1;
      }
      return;
    }
  } else {
    double one[kLzWaveU];
#pragma unroll
    for (int u = 0; u < kLzWaveU; ++u) one[u] = lane + 64 * u < d ? 1.0 : 0.0;
    wave_sym_matvec<kW>(S, d, U, one, r, xb, part, red);  // r = S 1
#pragma unroll
    for (int u = 0; u < kLzWaveU; ++u) {
      const int e = lane + 64 * u;
      sigma += r[u];
      q[u] = e < d ? power_v0(a.seed, a.iter, a.cluster, e) : 0.0;
      qp[u] = 0.0;
      n0 += q[u] * q[u];
      if (e < d && writer) a.r[e] = r[u];
    }
    sigma = wave_sum_f64(sigma);
    n0 = sqrt(wave_sum_f64(n0));
  }
#pragma unroll
  for (int u = 0; u < kLzWaveU; ++u) {
    const int e = lane + 64 * u;
    q[u] /= n0;
    if (e < d) a.Q[e] = q[u];
  }
  const int jmax = d < kLzMax ? d : kLzMax;
  int len = 0, conv = 0, prev_len = 0;
  double bprev = 0.0;
  for (int it = 0; it < jmax; ++it) {
    wave_sym_matvec<kW>(S, d, U, q, w, xb, part, red);  // w = S q
    if (!a.gram) {                   // w = M q = P S P q
      double sv = 0.0, rv = 0.0;
#pragma unroll
      for (int u = 0; u < kLzWaveU; ++u) {
        sv += q[u];
        rv += r[u] * q[u];
      }
      sv = wave_sum_f64(sv);
      rv = wave_sum_f64(rv);
#pragma unroll
      for (int u = 0; u < kLzWaveU; ++u)
        if (lane + 64 * u < d)
          w[u] = w[u] - (r[u] * sv + rv) * inv_d + sigma * sv * inv_d * inv_d;
    }
    double nw2 = 0.0, qw = 0.0;
#pragma unroll
    for (int u = 0; u < kLzWaveU; ++u) {
      nw2 += w[u] * w[u];
      qw += q[u] * w[u];
    }
    nw2 = wave_sum_f64(nw2);
    qw = wave_sum_f64(qw);
    double w1 = 0.0;
#pragma unroll
    for (int u = 0; u < kLzWaveU; ++u) {
      w[u] = w[u] - qw * q[u] - bprev * qp[u];
      w1 += w[u] * w[u];
    }
    w1 = wave_sum_f64(w1);
    double b = sqrt(w1);
    for (int pass = 0; pass < 2; ++pass) {
      // classical Gram-Schmidt against q_0..q_it, repeated when it removed
      // more than 1 - 1/sqrt(2) of the norm; basis rows in groups of
      // kLzWaveRows so that their loads are in flight together
      for (int l0 = 0; l0 <= it; l0 += kLzWaveRows) {
        double qv[kLzWaveRows][kLzWaveU];
        wave_basis_rows(a.Q, d, U, l0, it, qv);
#pragma unroll
        for (int t = 0; t < kLzWaveRows; ++t) {
          double dt = 0.0;
#pragma unroll
          for (int u = 0; u < kLzWaveU; ++u) dt += qv[t][u] * w[u];
          dt = wave_sum_f64(dt);
          if (lane == 0 && l0 + t <= it) cf[l0 + t] = dt;
        }
      }
      wave_lds_sync();
      for (int l0 = 0; l0 <= it; l0 += kLzWaveRows) {
        double qv[kLzWaveRows][kLzWaveU];
        wave_basis_rows(a.Q, d, U, l0, it, qv);
#pragma unroll
        for (int t = 0; t < kLzWaveRows; ++t) {
          const double c = l0 + t <= it ? cf[l0 + t] : 0.0;
#pragma unroll
          for (int u = 0; u < kLzWaveU; ++u) w[u] -= c * qv[t][u];
        }
      }
      double w2 = 0.0;
#pragma unroll
      for (int u = 0; u < kLzWaveU; ++u) w2 += w[u] * w[u];
      w2 = wave_sum_f64(w2);
      const bool again = w2 < 0.5 * w1;
      b = sqrt(w2);
      w1 = w2;
      wave_lds_sync();  // cf is rewritten by the next pass
      if (!again) break;
    }
    if (lane == 0) alpha[it] = qw;
    len = it + 1;
    const bool breakdown = !(b > 1e-12 * sqrt(nw2));
    if (breakdown || len == jmax || lz_power_check(len)) {
      wave_lds_sync();
      tridiag_power_wave(alpha, beta, len, a.I, y);
      wave_lds_sync();
      double dd = 0.0;
      for (int e = lane; e < len; e += 64) {
        const double pv = e < prev_len ? yp[e] : 0.0;
        dd += (y[e] - pv) * (y[e] - pv);
      }
      dd = wave_sum_f64(dd);
      conv = breakdown || (prev_len > 0 && sqrt(dd) <= lz_tol);
      for (int e = lane; e < len; e += 64) yp[e] = y[e];
      wave_lds_sync();
      prev_len = len;
      if (conv || breakdown) break;
    }
    if (len == jmax) break;
    if (lane == 0) beta[it + 1] = b;
    const double inv_b = 1.0 / b;
    double* qn = a.Q + (long)(it + 1) * d;
#pragma unroll
    for (int u = 0; u < kLzWaveU; ++u) {
      const int e = lane + 64 * u;
      qp[u] = q[u];
      q[u] = w[u] * inv_b;
      if (e < d) qn[e] = q[u];
    }
    bprev = b;
  }
  // v = Q y (zero when the check never passed, in the S form)
  double v[kLzWaveU];
  int steps = len;
  if (conv) {
#pragma unroll
    for (int u = 0; u < kLzWaveU; ++u) v[u] = 0.0;
    for (int l0 = 0; l0 < len; l0 += kLzWaveRows) {
      double qv[kLzWaveRows][kLzWaveU];
      wave_basis_rows(a.Q, d, U, l0, len - 1, qv);
#pragma unroll
      for (int t = 0; t < kLzWaveRows; ++t) {
        const double c = l0 + t < len ? y[l0 + t] : 0.0;
#pragma unroll
        for (int u = 0; u < kLzWaveU; ++u) v[u] += c * qv[t][u];
      }
    }
  } else if (!a.gram) {
#pragma unroll
    for (int u = 0; u < kLzWaveU; ++u) v[u] = 0.0;
  } else {
    // Gram form: plain power steps from vstart (at most I, until the
    // Rayleigh quotient moves by <= 1e-12 relative); S v = 0 gives v = 0
#pragma unroll
    for (int u = 0; u < kLzWaveU; ++u) {
      const int e = lane + 64 * u;
      v[u] = e < d ? a.vstart[e] / n0 : 0.0;
    }
    double lam_prev = 0.0;
    for (int t = 0; t < a.I; ++t) {
      wave_sym_matvec<kW>(S, d, U, v, w, xb, part, red);
      double nw2 = 0.0, vw = 0.0;
#pragma unroll
      for (int u = 0; u < kLzWaveU; ++u) {
        nw2 += w[u] * w[u];
        vw += v[u] * w[u];
      }
      nw2 = wave_sum_f64(nw2);
      vw = wave_sum_f64(vw);
      if (!(nw2 > 0.0)) {
#pragma unroll
        for (int u = 0; u < kLzWaveU; ++u) v[u] = 0.0;
        break;
      }
      const double inv = 1.0 / sqrt(nw2);
#pragma unroll
      for (int u = 0; u < kLzWaveU; ++u) v[u] = w[u] * inv;
      ++steps;
      if (t > 0 && fabs(vw - lam_prev) <= 1e-12 * fabs(vw)) break;
      lam_prev = vw;
    }
    conv = 1;
  }
  wave_sym_matvec<kW>(S, d, U, v, w, xb, part, red);  // h = S v
#pragma unroll
  for (int u = 0; u < kLzWaveU; ++u) {
    const int e = lane + 64 * u;
    if (e < d && writer) {
      a.v[e] = v[u];
      a.h[e] = w[u];
    }
  }
  if (lane == 0 && writer) {
    *a.steps = steps;
    *a.conv = conv;
  }
}

#define TCX_GROUP_ARGS(T) static_cast<const T*>(g.jobs), g.first, g.njobs

hipError_t launch_group_count_group(const GroupTable& g, hipStream_t st) {
  if (g.blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_group_count_group, dim3(g.blocks), dim3(256), g.lds, st,
                     TCX_GROUP_ARGS(GroupCountJob));
  return hipGetLastError();
}

hipError_t launch_group_scan_group(const GroupTable& g, hipStream_t st) {
  if (g.blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_group_scan_group, dim3(g.blocks), dim3(256), 0, st,
                     TCX_GROUP_ARGS(GroupScanJob));
  return hipGetLastError();
}

hipError_t launch_group_offsets_group(const GroupTable& g, hipStream_t st) {
  if (g.blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_group_offsets_group, dim3(g.blocks), dim3(1024), 0, st,
                     TCX_GROUP_ARGS(GroupOffsetsJob));
  return hipGetLastError();
}

size_t group_place_lds(int k, int B) { return sizeof(int) * kPlaceWaves * (size_t)k * B; }

hipError_t launch_group_place_group(const GroupTable& g, hipStream_t st) {
  if (g.blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_group_place_group, dim3(g.blocks), dim3(kGroupChunk), g.lds, st,
                     TCX_GROUP_ARGS(GroupPlaceJob));
  return hipGetLastError();
}

hipError_t launch_row_stats_group(const GroupTable& g, hipStream_t st) {
  if (g.blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_row_stats_group, dim3(g.blocks), dim3(256), 0, st,
                     TCX_GROUP_ARGS(RowStatsJob));
  return hipGetLastError();
}

static_assert(kTunePackMaxM == kPackLdsMaxM && kTunePackMaxBytes == kPackLdsMaxBytes,
              "the tuner's align + pack limits are the kernel's");

int align_pack_form(int m) {
  if (m > kPackLdsMaxM) return 0;
  const int tuned = tuned_pack_rows(m);  // every column is packed alike in any block
  return tuned > 0 ? tuned : pack_lds_rows(m);
}
long align_pack_cols(int form) { return form ? form : 64; }
size_t align_pack_lds(int form, int m) { return form ? sizeof(float) * (size_t)form * (m + 4) : 0; }

hipError_t launch_align_pack_group(int form, const GroupTable& g, hipStream_t st) {
  if (g.blocks <= 0) return hipSuccess;
  switch (form) {
    case 0:
      hipLaunchKernelGGL(k_align_pack_group, dim3(g.blocks), dim3(256), 0, st,
                         TCX_GROUP_ARGS(AlignPackJob));
      break;
    case 16:
      hipLaunchKernelGGL(k_align_pack_lds_group<16>, dim3(g.blocks), dim3(256), g.lds, st,
                         TCX_GROUP_ARGS(AlignPackJob));
      break;
    case 32:
      hipLaunchKernelGGL(k_align_pack_lds_group<32>, dim3(g.blocks), dim3(256), g.lds, st,
                         TCX_GROUP_ARGS(AlignPackJob));
      break;
    case 64:
      hipLaunchKernelGGL(k_align_pack_lds_group<64>, dim3(g.blocks), dim3(256), g.lds, st,
                         TCX_GROUP_ARGS(AlignPackJob));
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_syrk_group(const SyrkJob* jobs, const GroupSyrkItem* items, int nitems,
                             int split, hipStream_t st) {
  if (nitems <= 0) return hipSuccess;
  const dim3 grid((unsigned)((nitems + 3) / 4));
  if (split)
    hipLaunchKernelGGL(k_syrk_group<true>, grid, dim3(256), 0, st, jobs, items, nitems);
  else
    hipLaunchKernelGGL(k_syrk_group<false>, grid, dim3(256), 0, st, jobs, items, nitems);
  return hipGetLastError();
}

hipError_t launch_syrk_reduce_group(const SyrkJob* jobs, const GroupSyrkTile* tiles, int ntiles,
                                    hipStream_t st) {
  if (ntiles <= 0) return hipSuccess;
  constexpr int kSub = kSyrkSub / 64;
  hipLaunchKernelGGL(k_syrk_reduce_group, dim3((unsigned)ntiles * kSub * kSub), dim3(256), 0, st,
                     jobs, tiles);
  return hipGetLastError();
}

hipError_t launch_gram_tiles(const GramItem* items, int nitems, hipStream_t st) {
  if (nitems <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gram_tiles, dim3((unsigned)nitems), dim3(256), 0, st, items);
  return hipGetLastError();
}

size_t gram_map_lds(int m, int nd) { return sizeof(double) * 2 * ((size_t)m + nd); }
size_t gram_lanczos_lds(int nd) { return lanczos_smem(nd); }

hipError_t launch_gram_lanczos(const GramLzJob* jobs, int njobs, size_t lds, double lz_tol,
                               hipStream_t st) {
  if (njobs <= 0) return hipSuccess;
  if (lds > lanczos_smem(kGramMaxND)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_gram_lanczos, dim3((unsigned)njobs), dim3(kPowerThreads), lds, st, jobs,
                     lz_tol);
  return hipGetLastError();
}

hipError_t launch_gram_map(const GramVecJob* jobs, int njobs, size_t lds, hipStream_t st) {
  if (njobs <= 0) return hipSuccess;
  if (lds > gram_map_lds(kGramMaxM, kGramMaxND)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_gram_map, dim3((unsigned)njobs), dim3(256), lds, st, jobs);
  return hipGetLastError();
}

hipError_t launch_lanczos_wave(const LzWaveJob* jobs, int njobs, int waves_per_cluster,
                               double lz_tol, hipStream_t st) {
  if (njobs <= 0) return hipSuccess;
  if (waves_per_cluster == 4)
    hipLaunchKernelGGL(k_lanczos_wave<4>, dim3((unsigned)njobs), dim3(256), 0, st, jobs, njobs,
                       lz_tol);
  else if (waves_per_cluster == 1)
    hipLaunchKernelGGL(k_lanczos_wave<1>, dim3((unsigned)((njobs + 3) / 4)), dim3(256), 0, st,
                       jobs, njobs, lz_tol);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

size_t power_lanczos_lds(int m) { return lanczos_smem(m); }

hipError_t launch_power_lanczos_group(const GroupTable& g, double lz_tol, hipStream_t st) {
  if (g.blocks <= 0) return hipSuccess;
  if (g.lds > lanczos_smem(kFusedLanczosMaxM)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_power_lanczos_group, dim3(g.blocks), dim3(kPowerThreads), g.lds, st,
                     TCX_GROUP_ARGS(PowerLanczosJob), lz_tol);
  return hipGetLastError();
}

long cluster_matvec_blocks(int m, int k) { return (long)((m + kMvRows - 1) / kMvRows) * k; }

hipError_t launch_cluster_matvec_group(const GroupTable& g, hipStream_t st) {
  if (g.blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cluster_matvec_group, dim3(g.blocks), dim3(kMvThreads), g.lds, st,
                     TCX_GROUP_ARGS(ClusterMatvecJob));
  return hipGetLastError();
}

hipError_t launch_lz_setup_group(const GroupTable& g, hipStream_t st) {
  if (g.blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lz_setup_group, dim3(g.blocks), dim3(kPowerThreads), 0, st,
                     TCX_GROUP_ARGS(LzJob));
  return hipGetLastError();
}

hipError_t launch_lz_step_group(const GroupTable& g, int mode, double lz_tol, hipStream_t st) {
  if (g.blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lz_step_group, dim3(g.blocks), dim3(kPowerThreads), g.lds, st,
                     TCX_GROUP_ARGS(LzJob), mode, lz_tol);
  return hipGetLastError();
}

hipError_t launch_lz_finish_group(const GroupTable& g, hipStream_t st) {
  if (g.blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lz_finish_group, dim3(g.blocks), dim3(kPowerThreads), 0, st,
                     TCX_GROUP_ARGS(LzJob));
  return hipGetLastError();
}

size_t power_lds(int m) { return power_smem(m); }
size_t lz_step_lds(int m) { return lz_step_smem(m); }

hipError_t launch_power_group(const GroupTable& g, hipStream_t st) {
  if (g.blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_power_group, dim3(g.blocks), dim3(kPowerThreads), g.lds, st,
                     TCX_GROUP_ARGS(PowerJob));
  return hipGetLastError();
}

hipError_t configure_extract_kernels() {
  hipError_t e = hipFuncSetAttribute((const void*)k_power,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)power_smem(4096));
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_power_lanczos,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lanczos_smem(kFusedLanczosMaxM));
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_lz_step, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lz_step_smem(4096));
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_lz_coop<true>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLzCoopLdsMax);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_lz_coop<false>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLzCoopLdsMax);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_cluster_matvec,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(double) * 4096));
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_align_pack_lds<64>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, kPackLdsMaxBytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_align_pack_lds<32>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, kPackLdsMaxBytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_align_pack_lds<16>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, kPackLdsMaxBytes);
  // grouped forms
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_power_group, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)power_smem(4096));
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_power_lanczos_group,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lanczos_smem(kFusedLanczosMaxM));
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_gram_lanczos, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lanczos_smem(kGramMaxND));
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_gram_map, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)gram_map_lds(kGramMaxM, kGramMaxND));
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_lz_step_group,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lz_step_smem(4096));
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_cluster_matvec_group,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(double) * 4096));
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_align_pack_lds_group<64>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, kPackLdsMaxBytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_align_pack_lds_group<32>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, kPackLdsMaxBytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_align_pack_lds_group<16>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, kPackLdsMaxBytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_group_count_group,
                            hipFuncAttributeMaxDynamicSharedMemorySize, 65536);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_group_place_group,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(int) * kPlaceWaves * kPlaceParMaxK));
  return e;
}

}  // namespace tcx

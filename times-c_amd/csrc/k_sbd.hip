// SBD kernels: z-normalization, centroid spectra, the fused
// FFT -> conj-multiply -> inverse FFT -> max-over-lags -> argmin assignment,
// the materialized NCC and the zero-fill shift.
//
// Layout: series are rows of a row-major FP32 [n][m] matrix.  A spectrum is
// F complex values stored lane-strided ([s][lane], s < P, lane < L) so that
// the lane group that produced it reads it back with coalesced loads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "tcx_device.hpp"
#include "tcx_internal.hpp"
#include "tcx_kernels.hpp"

namespace tcx {

// ---------------------------------------------------------------- z-norm
// One wave per row, FP64 accumulation of mean and of the centred sum of
// squares; (x - mean) / sqrt(ss / (m - ddof)), NaN -> 0 (nan_to_num).
// Semantics of zscore, Code/test/test_timesC.py:42-52.
__device__ __forceinline__ void znorm_body(const int bid, const int nblk,
    float* __restrict__ x, long rows,
                                               int m, int ddof) {
  const int lane = threadIdx.x & (kWave - 1);
  const long row = (long)bid * 4 + (threadIdx.x >> 6);
  if (row >= rows) return;
  float* r = x + row * (long)m;
  double s = 0.0;
  for (int t = lane; t < m; t += kWave) s += (double)r[t];
  s = wave_sum_f64(s);
  const double mean = s / (double)m;
  double ss = 0.0;
  for (int t = lane; t < m; t += kWave) {
    double d = (double)r[t] - mean;
    ss += d * d;
  }
  ss = wave_sum_f64(ss);
  const double sd = sqrt(ss / (double)(m - ddof));
  for (int t = lane; t < m; t += kWave) {
    double v = ((double)r[t] - mean) / sd;
    float f;
    if (isnan(v))
      f = 0.f;
    else if (isinf(v))
      f = v > 0 ? 3.402823466e38f : -3.402823466e38f;
    else
      f = (float)v;
    r[t] = f;
  }
}

__global__ __launch_bounds__(256) void k_znorm(float* __restrict__ x, long rows,
                                               int m, int ddof) {
  znorm_body(blockIdx.x, gridDim.x, x, rows, m, ddof);
}

// The same for m % 4 == 0, m <= 256 * NV: the row is read once, as float4
// vectors held in registers (lane l holds vectors l, l + 64, ...), and
// written once; the sums are the same two-pass FP64 sums (in another
// order).
template <int NV>
__device__ __forceinline__ void znorm_v_body(const int bid, const int nblk,
    float* __restrict__ x, long rows, int m,
                                                 int ddof) {
  const int lane = threadIdx.x & (kWave - 1);
  const long row = (long)bid * 4 + (threadIdx.x >> 6);
  if (row >= rows) return;
  float4* r = reinterpret_cast<float4*>(x + row * (long)m);
  const int nv = m >> 2;
  float4 v[NV];
#pragma unroll
  for (int u = 0; u < NV; ++u) {
    const int i = lane + kWave * u;
    v[u] = i < nv ? r[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  double s = 0.0;
#pragma unroll
  for (int u = 0; u < NV; ++u)
    s += ((double)v[u].x + (double)v[u].y) + ((double)v[u].z + (double)v[u].w);
  s = wave_sum_f64(s);
  const double mean = s / (double)m;
  double ss = 0.0;
#pragma unroll
  for (int u = 0; u < NV; ++u) {
    if (lane + kWave * u < nv) {
      const double a = (double)v[u].x - mean, b = (double)v[u].y - mean;
      const double c = (double)v[u].z - mean, d = (double)v[u].w - mean;
      ss += (a * a + b * b) + (c * c + d * d);
    }
  }
  ss = wave_sum_f64(ss);
  const double sd = sqrt(ss / (double)(m - ddof));
  auto z = [&](float e) {
    const double w = ((double)e - mean) / sd;
    return isnan(w) ? 0.f : (isinf(w) ? (w > 0 ? 3.402823466e38f : -3.402823466e38f) : (float)w);
  };
#pragma unroll
  for (int u = 0; u < NV; ++u) {
    const int i = lane + kWave * u;
    if (i < nv) r[i] = make_float4(z(v[u].x), z(v[u].y), z(v[u].z), z(v[u].w));
  }
}

template <int NV>
__global__ __launch_bounds__(256) void k_znorm_v(float* __restrict__ x, long rows, int m,
                                                 int ddof) {
  znorm_v_body<NV>(blockIdx.x, gridDim.x, x, rows, m, ddof);
}

// ------------------------------------------------------------- spectra
// Forward FFT of each row (zero padded to F, imaginary part 0), stored
// lane-strided, plus the row's L2 norm, an all-zero flag and (mag != NULL)
// the spectrum's magnitudes lane-major: |X[lane + L*s]| at lane*P + s, so a
// lane reads its values as 16-byte vectors.
template <int F>
__device__ __forceinline__ void spectra_body(const int bid, const int nblk,
    const float* __restrict__ y, int rows, int m, const float2* __restrict__ twg,
    float2* __restrict__ spec, float* __restrict__ norms, int* __restrict__ zero,
    float* __restrict__ mag, uint16_t* __restrict__ bmag) {
  using G = FftGeom<F>;
  constexpr int L = G::L, P = G::P;
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float* red = reinterpret_cast<float*>(smem + F + G::kGroups * fft_buf_len<F>());
  for (int i = threadIdx.x; i < F; i += blockDim.x) tw[i] = twg[i];
  __syncthreads();
  const int g = threadIdx.x / L, lane = threadIdx.x % L;
  float2* buf = smem + F + g * fft_buf_len<F>();
  for (int row = bid * G::kGroups + g; row < rows; row += nblk * G::kGroups) {
    const float* yr = y + (long)row * m;
    float2 v[P];
    float ss = 0.f;
    int nz = 0;
#pragma unroll
    for (int s = 0; s < P; ++s) {
      int t = lane + L * s;
      float a = t < m ? yr[t] : 0.f;
      v[s] = make_float2(a, 0.f);
      ss += a * a;
      nz |= (a != 0.f);
    }
    ss = group_sum<L>(ss, red);
    float fnz = group_sum<L>((float)nz, red);
    fft<F, L, false>(v, buf, tw, lane);
    float2* o = spec + (long)row * F;
#pragma unroll
    for (int s = 0; s < P; ++s) o[s * L + lane] = v[s];
    if (mag) {
      float* mo = mag + (long)row * F + lane * P;
#pragma unroll
      for (int s = 0; s < P; ++s) mo[s] = sqrtf(v[s].x * v[s].x + v[s].y * v[s].y);
    }
    if (bmag) {  // bound operand row: |C[f]| / |c| for f <= F/2, zero padded
      constexpr int KP = bound_kpad(F);
      const float rn = ss > 0.f ? 1.f / sqrtf(ss) : 0.f;
#pragma unroll
      for (int s = 0; s < P; ++s) {
        const int f = lane + L * s;
        if (f <= F / 2)
          bmag[bound_frag(row, f, KP)] = f16_bits_up(sqrtf(v[s].x * v[s].x + v[s].y * v[s].y) * rn);
      }
      for (int f = F / 2 + 1 + lane; f < KP; f += L) bmag[bound_frag(row, f, KP)] = 0;
    }
    if (lane == 0) {
      norms[row] = sqrtf(ss);
      if (zero) zero[row] = fnz == 0.f;
    }
  }
}

template <int F>
__global__ __launch_bounds__(FftGeom<F>::kBlock) void k_spectra(
    const float* __restrict__ y, int rows, int m, const float2* __restrict__ twg,
    float2* __restrict__ spec, float* __restrict__ norms, int* __restrict__ zero,
    float* __restrict__ mag, uint16_t* __restrict__ bmag) {
  spectra_body<F>(blockIdx.x, gridDim.x, y, rows, m, twg, spec, norms, zero, mag, bmag);
}

// ------------------------------------------------ fused SBD / assignment
// Two series share one complex transform: z = x_a + i x_b.  For centroid j
// with spectrum C_j, ifft(FFT(z) conj(C_j)) = cc_a + i cc_b, the two
// circular cross-correlations of length F (F >= 2m-1, so no aliasing).
// ASSIGN: per series keep the centroid with the smallest
//   dist = 1 - (max_lag cc / F) / (|x| |c_j|)       (first argmin)
// and the lag key of its maximum (first argmax), test_timesC.py:258-260.
// NCC: write cc / F / (|x| |y|) for all 2m-1 lags (ncc.cuh:467-517 layout
// out[x][y][lag]).
template <int F, int MODE>
__global__ __launch_bounds__(FftGeom<F>::kBlock) void k_sbd(
    const float* __restrict__ x, long n, int m, int k,
    const float2* __restrict__ cspec, const float* __restrict__ cnorm,
    const float2* __restrict__ twg, int rev, int* __restrict__ label,
    float* __restrict__ dist, int* __restrict__ shift, float* __restrict__ dmat,
    int* __restrict__ changed, float* __restrict__ ncc_out,
    const int* __restrict__ fixed_label) {
  using G = FftGeom<F>;
  constexpr int L = G::L, P = G::P;
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float* red = reinterpret_cast<float*>(smem + F + G::kGroups * fft_buf_len<F>());
  for (int i = threadIdx.x; i < F; i += blockDim.x) tw[i] = twg[i];
  __syncthreads();
  const int g = threadIdx.x / L, lane = threadIdx.x % L;
  float2* buf = smem + F + g * fft_buf_len<F>();
  const long npairs = (n + 1) / 2;
  const int nlag = 2 * m - 1;
  int s_lo, s_hi;
  lag_rows(lane, L, m, F, s_lo, s_hi);
  for (long pr = (long)blockIdx.x * G::kGroups + g; pr < npairs;
       pr += (long)gridDim.x * G::kGroups) {
    const long ia = 2 * pr, ib = 2 * pr + 1;
    const bool hasb = ib < n;
    const float* xa = x + ia * (long)m;
    const float* xb = x + (hasb ? ib : ia) * (long)m;
    float2 zx[P];
    float sa = 0.f, sb = 0.f;
#pragma unroll
    for (int s = 0; s < P; ++s) {
      int t = lane + L * s;
      float a = t < m ? xa[t] : 0.f;
      float b = (hasb && t < m) ? xb[t] : 0.f;
      zx[s] = make_float2(a, b);
      sa += a * a;
      sb += b * b;
    }
    sa = group_sum<L>(sa, red);
    sb = group_sum<L>(sb, red);
    const float na = sqrtf(sa), nb = sqrtf(sb);
    fft<F, L, false>(zx, buf, tw, lane);

    float bestA = INFINITY, bestB = INFINITY;
    int labA = 0, labB = 0, keyA = 0, keyB = 0;
    // fixed_label: report the lag for the given centroid instead of the argmin
    const int fixA = fixed_label ? fixed_label[ia] : -1;
    const int fixB = (fixed_label && hasb) ? fixed_label[ib] : -1;
    for (int j = 0; j < k; ++j) {
      const float2* cj = cspec + (long)j * F;
      float2 z[P];
#pragma unroll
      for (int s = 0; s < P; ++s) z[s] = cmulc(zx[s], cj[s * L + lane]);
      fft<F, L, true>(z, buf, tw, lane);
      const float cn = cnorm[j];
      const float dena = na * cn, denb = nb * cn;
      if constexpr (MODE == kModeAssign) {
        float va, vb;
        int ka, kb;
        lag_argmax2<L, P>(z, s_lo, s_hi, rev, lane, m, F, va, ka, vb, kb);
        group_argmax<L>(va, ka, red);
        group_argmax<L>(vb, kb, red);
        // den == 0: NumPy divides by inf, every lag is 0, argmax is index 0
        float da = 1.f, db = 1.f;
        if (dena != 0.f) da = 1.f - (va / (float)F) / dena; else ka = 0;
        if (denb != 0.f) db = 1.f - (vb / (float)F) / denb; else kb = 0;
        if (fixed_label ? (j == fixA) : (da < bestA)) { bestA = da; labA = j; keyA = ka; }
        if (fixed_label ? (j == fixB) : (db < bestB)) { bestB = db; labB = j; keyB = kb; }
        if (dmat && lane == 0) {
          dmat[ia * k + j] = da;
          if (hasb) dmat[ib * k + j] = db;
        }
      } else {
        const float sca = dena != 0.f ? 1.f : 0.f, scb = denb != 0.f ? 1.f : 0.f;
        const float dA = dena != 0.f ? dena : 1.f, dB = denb != 0.f ? denb : 1.f;
        float* oa = ncc_out + (ia * (long)k + j) * nlag;
        float* ob = ncc_out + (ib * (long)k + j) * nlag;
#pragma unroll
        for (int s = 0; s < P; ++s) {
          int key = lag_key(lane + L * s, m, F, 0);
          if (key >= 0) {
            oa[key] = sca * ((z[s].x / (float)F) / dA);
            if (hasb) ob[key] = scb * ((z[s].y / (float)F) / dB);
          }
        }
      }
    }
    if constexpr (MODE == kModeAssign) {
      if (lane == 0) {
        if (changed && label[ia] != labA) *changed = 1;
        label[ia] = labA;
        if (dist) dist[ia] = bestA;
        if (shift) shift[ia] = keyA + 1 - m;
        if (hasb) {
          if (changed && label[ib] != labB) *changed = 1;
          label[ib] = labB;
          if (dist) dist[ib] = bestB;
          if (shift) shift[ib] = keyB + 1 - m;
        }
      }
    }
  }
}

// The first N (<= P) magnitudes of a lane from its lane-major row segment
// (16-byte aligned, P a multiple of 4): N/4 vector loads and a scalar tail.
template <int P, int N>
__device__ __forceinline__ void load_mag(const float* __restrict__ src, float (&out)[N]) {
  static_assert(P % 4 == 0 && N <= P, "lane-major magnitudes");
#pragma unroll
  for (int i = 0; i + 4 <= N; i += 4) {
    const float4 q = *reinterpret_cast<const float4*>(src + i);
    out[i] = q.x;
    out[i + 1] = q.y;
    out[i + 2] = q.z;
    out[i + 3] = q.w;
  }
#pragma unroll
  for (int i = N / 4 * 4; i < N; ++i) out[i] = src[i];
}

// ------------------------------------------- pruned assignment (exact)
// For every centroid j the lag maximum of the correlation is bounded by the
// spectra:  max_l cc[l] = max_l (1/F) |sum_f X_f conj(C_f) e^{2 pi i f l/F}|
// <= (1/F) sum_f |X_f| |C_f|,  so  dist_j >= lb_j = 1 - ub_j  with
// ub_j = sum_f |X_f||C_f| / (F |x| |c_j|).  Per pair of series the
// centroids are evaluated (inverse FFT, exact distance) in decreasing
// order of their bound; a centroid whose lower bound (with a rounding
// margin) exceeds the best exact distance found cannot be the argmin and
// is skipped.  Ties keep NumPy's first-index rule.  The result equals the
// dense kernel's; the pair's two series share each evaluated transform.
// |X_a|, |X_b| come from the packed transform Z = FFT(x_a + i x_b):
// X_a = (Z_f + conj(Z_{F-f}))/2, X_b = (Z_f - conj(Z_{F-f}))/(2i).
// order (may be NULL): row indices to pair up, -1 = hole; rows with the
// same previous label are adjacent there, so both series of a pair tend to
// share their candidates.
constexpr float kUbRel = 1e-4f;   // rounding margin of the bound
constexpr float kUbAbs = 1e-5f;

// a' + b' after v_permlane16_swap (O = 16: odd rows of a with even rows of
// b) or v_permlane32_swap (O = 32: upper half of a with lower half of b).
template <int O>
__device__ __forceinline__ float swap_sum(float a, float b) {
  static_assert(O == 16 || O == 32, "row or half swap");
  if constexpr (O == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false,
                                                    false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
  } else {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false,
                                                    false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
  }
}

// Reduce-scatter of the CNT values v[0..CNT) of every lane over a group of
// L lanes (one wave or less): at lane-bit O a lane keeps the lower or upper
// half of its values (by that bit) plus the partner's matching half; once a
// single value is left the remaining levels are plain butterfly sums.
// Afterwards a lane holds the group sums of values base .. base + CNT/L - 1
// (one value when CNT <= L).
template <int CNT, int O, int L, int N>
__device__ __forceinline__ void reduce_scatter(float (&v)[N], int lane, int& base) {
  if constexpr (O >= 1) {
    const bool up = (lane & O) != 0;
    if constexpr (O >= 16) {
      // rows of 16 (O = 16) or halves of 32 (O = 32) swapped between two
      // registers a, b: afterwards a lane without bit O holds its own a and
      // its partner's a, a lane with bit O its partner's b and its own b, so
      // the sum of the two is each lane's kept half plus the partner's
      if constexpr (CNT > 1) {
        constexpr int HF = CNT / 2;
#pragma unroll
        for (int i = 0; i < HF; ++i) v[i] = swap_sum<O>(v[i], v[i + HF]);
        if (up) base += HF;
        reduce_scatter<HF, O / 2, L>(v, lane, base);
      } else {
        v[0] = swap_sum<O>(v[0], v[0]);
        reduce_scatter<1, O / 2, L>(v, lane, base);
      }
    } else {
      // partner by DPP: lane ^ 1, lane ^ 2, the 8-lane mirror (bit 2 flips),
      // the 16-lane mirror (bit 3 flips); each pairs the two halves of the
      // current subgroup one to one
      constexpr int CTRL = O == 1 ? 0xB1 : O == 2 ? 0x4E : O == 4 ? 0x141 : 0x140;
      if constexpr (CNT > 1) {
        constexpr int HF = CNT / 2;
#pragma unroll
        for (int i = 0; i < HF; ++i) {
          const float send = up ? v[i] : v[i + HF];
          const float keep = up ? v[i + HF] : v[i];
          v[i] = keep + dpp_f32<CTRL>(send);
        }
        if (up) base += HF;
        reduce_scatter<HF, O / 2, L>(v, lane, base);
      } else {
        v[0] += dpp_f32<CTRL>(v[0]);
        reduce_scatter<1, O / 2, L>(v, lane, base);
      }
    }
  }
}

// Partner values of the packed spectrum Z = FFT(x_a + i x_b) held
// lane-strided (Z[lane + L*s] in zx[s]): for s <= P/2, (px[s], py[s]) =
// Z[F - f] with f = lane + L*s, from lane' = (L - lane) % L, s' = P-1-s
// (lane 0: s' = (P - s) % P).  Then X_a[f] = (Z[f] + conj(Z[F-f])) / 2 and
// X_b[f] = (Z[f] - conj(Z[F-f])) / (2i).  Groups wider than a wave exchange
// through `buf` (LDS) between barriers.
template <int F, int L, int P>
__device__ __forceinline__ void partner_half(const float2 (&zx)[P], float2* buf, int lane,
                                             float (&px)[P / 2 + 1], float (&py)[P / 2 + 1]) {
  constexpr int H = P / 2;
  const int pl = (L - lane) & (L - 1);
  if constexpr (L > kWave) {
#pragma unroll
    for (int s = 0; s < P; ++s) buf[lds_pad(lane + L * s)] = zx[s];
    __syncthreads();
  }
#pragma unroll
  for (int s = 0; s <= H; ++s) {
    if constexpr (L > kWave) {
      const float2 q = buf[lds_pad(pl + L * (P - 1 - s))];
      px[s] = q.x;
      py[s] = q.y;
    } else {
      px[s] = __shfl(zx[P - 1 - s].x, pl, L);
      py[s] = __shfl(zx[P - 1 - s].y, pl, L);
    }
    if (lane == 0) {
      px[s] = zx[(P - s) & (P - 1)].x;
      py[s] = zx[(P - s) & (P - 1)].y;
    }
  }
  if constexpr (L > kWave) __syncthreads();  // reads done before buf is reused
}

// Half spectra of every series, once per fit (the series do not change
// between iterations): row r holds X_r[f] for f = 0 .. F/2 (F/2 + 1 values,
// X_r[F-f] = conj(X_r[f])), and xn[r] = |x_r|.  Rows are transformed in
// natural pairs packed as x_a + i x_b and split with partner_half.
template <int F>
__device__ __forceinline__ void series_spec_body(const int bid, const int nblk,
    const float* __restrict__ x, long n, int m, const float2* __restrict__ twg,
    float2* __restrict__ xs, float* __restrict__ xn, uint16_t* __restrict__ xm) {
  using G = FftGeom<F>;
  constexpr int L = G::L, P = G::P, H = P / 2;
  constexpr long HS = F / 2 + 1;
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float* red = reinterpret_cast<float*>(smem + F + G::kGroups * fft_buf_len<F>());
  for (int i = threadIdx.x; i < F; i += blockDim.x) tw[i] = twg[i];
  __syncthreads();
  const int g = threadIdx.x / L, lane = threadIdx.x % L;
  float2* buf = smem + F + g * fft_buf_len<F>();
  const long npairs = (n + 1) / 2;
  for (long pr = (long)bid * G::kGroups + g; pr < npairs;
       pr += (long)nblk * G::kGroups) {
    const long ia = 2 * pr, ib = 2 * pr + 1;
    const bool hasb = ib < n;
    const float* xa = x + ia * (long)m;
    const float* xb = x + (hasb ? ib : ia) * (long)m;
    float2 zx[P];
    float sa = 0.f, sb = 0.f;
#pragma unroll
    for (int s = 0; s < P; ++s) {
      const int t = lane + L * s;
      const float a = t < m ? xa[t] : 0.f;
      const float b = (hasb && t < m) ? xb[t] : 0.f;
      zx[s] = make_float2(a, b);
      sa += a * a;
      sb += b * b;
    }
    sa = group_sum<L>(sa, red);
    sb = group_sum<L>(sb, red);
    if (lane == 0) {
      xn[ia] = sqrtf(sa);
      if (hasb) xn[ib] = sqrtf(sb);
    }
    fft<F, L, false>(zx, buf, tw, lane);
    float px[H + 1], py[H + 1];
    partner_half<F, L, P>(zx, buf, lane, px, py);
    float2* oa = xs + ia * HS;
    float2* ob = xs + ib * HS;
    // bound rows: weight * |X_f| / |x|, weight 2 inside, 1 at f = 0, F/2
    // (the 1/F is applied to the product)
    constexpr int KP = bound_kpad(F);
    const float ra = sa > 0.f ? 1.f / sqrtf(sa) : 0.f;
    const float rb = sb > 0.f ? 1.f / sqrtf(sb) : 0.f;
#pragma unroll
    for (int s = 0; s <= H; ++s) {
      const int f = lane + L * s;
      if (s < H || lane == 0) {  // f < F/2, or f = F/2
        const int fo = s < H ? f : F / 2;
        const float2 A = make_float2(0.5f * (zx[s].x + px[s]), 0.5f * (zx[s].y - py[s]));
        const float2 B = make_float2(0.5f * (zx[s].y + py[s]), 0.5f * (px[s] - zx[s].x));
        oa[fo] = A;
        if (hasb) ob[fo] = B;
        if (xm) {
          const float w = (fo == 0 || fo == F / 2) ? 1.f : 2.f;
          xm[bound_frag(ia, fo, KP)] = f16_bits_up(w * sqrtf(A.x * A.x + A.y * A.y) * ra);
          if (hasb) xm[bound_frag(ib, fo, KP)] = f16_bits_up(w * sqrtf(B.x * B.x + B.y * B.y) * rb);
        }
      }
    }
    if (xm)
      for (int f = F / 2 + 1 + lane; f < KP; f += L) {  // padding of the K dimension
        xm[bound_frag(ia, f, KP)] = 0;
        if (hasb) xm[bound_frag(ib, f, KP)] = 0;
      }
  }
}

template <int F>
__global__ __launch_bounds__(FftGeom<F>::kBlock) void k_series_spec(
    const float* __restrict__ x, long n, int m, const float2* __restrict__ twg,
    float2* __restrict__ xs, float* __restrict__ xn, uint16_t* __restrict__ xm) {
  series_spec_body<F>(blockIdx.x, gridDim.x, x, n, m, twg, xs, xn, xm);
}

// ------------------------------------------ bounds of the pruned assignment
// ub[i][j] = sum_f xm[i][f] cm[j][f] / F for every series i and centroid j:
// the bounds k_assign_prune otherwise forms per pair, as one product of the
// per-fit magnitude rows xm (k_series_spec) and the centroids' rows cm =
// |C_j[f]| / |c_j| (k_spectra), both binary16 rounded up in fragment order
// (bound_frag), on v_mfma_f32_16x16x16_f16 with FP32 accumulation (the
// products of two binary16 values are exact in FP32).  A wave computes 16
// series x 16 centroids per tile and NT tiles per pass over its rows of xm;
// for the 16 columns of a step lane l reads the l-th 8-byte vector of the
// fragment block of each operand (one contiguous 512 B per wave).  Result
// element (i, j) is in lane j + 16 (i / 4), register i % 4.  Rows past n (or
// centroids past k) in the last block are computed from whatever the
// padding holds and never stored: a result element depends on its own row
// and column only.
typedef _Float16 bound_h4 __attribute__((ext_vector_type(4)));
typedef float bound_f4 __attribute__((ext_vector_type(4)));
template <int NT, int MT, int NS>
__device__ __forceinline__ void bounds_body(const int bid, const int nblk,
    const uint16_t* __restrict__ xm, long n, int kpad,
                                                const uint16_t* __restrict__ cm, int k, int ld,
                                                float scale, float* __restrict__ ub) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const long blk0 = ((long)bid * 4 + wid) * MT;  // first 16-row block of the wave
  if (blk0 * 16 >= n) return;
  const long nrb = (n + 15) >> 4;  // row blocks of 16
  const int r = lane & 15, kq = lane >> 4;
  const int nsteps = kpad >> 4;
  const bound_h4* xa[MT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const long b = blk0 + mt < nrb ? blk0 + mt : blk0;  // blocks past n: computed, not stored
    xa[mt] = reinterpret_cast<const bound_h4*>(xm) + b * nsteps * 64 + lane;
  }
  for (int t0 = 0; t0 * 16 < k; t0 += NT) {
    bound_f4 acc[MT][NT];
    const bound_h4* cb[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int jt = (t0 + t) * 16 < k ? t0 + t : t0;  // tiles past k: computed, not stored
      cb[t] = reinterpret_cast<const bound_h4*>(cm) + (long)jt * nsteps * 64 + lane;
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) acc[mt][t] = bound_f4{0.f, 0.f, 0.f, 0.f};
    }
    for (int s0 = 0; s0 < nsteps; s0 += NS) {
      bound_h4 a[MT][NS], bv[NT][NS];
#pragma unroll
      for (int u = 0; u < NS; ++u) {
        const int s = s0 + u < nsteps ? s0 + u : nsteps - 1;  // clamped, dropped below
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) a[mt][u] = xa[mt][(long)s * 64];
#pragma unroll
        for (int t = 0; t < NT; ++t) bv[t][u] = cb[t][(long)s * 64];
      }
#pragma unroll
      for (int u = 0; u < NS; ++u) {
        if (s0 + u < nsteps) {  // uniform
#pragma unroll
          for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int t = 0; t < NT; ++t)
              acc[mt][t] =
                  __builtin_amdgcn_mfma_f32_16x16x16f16(a[mt][u], bv[t][u], acc[mt][t], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int j = (t0 + t) * 16 + r;
        if (j >= k) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const long i = (blk0 + mt) * 16 + 4 * kq + q;
          if (i < n) ub[i * ld + j] = acc[mt][t][q] * scale;
        }
      }
  }
}

template <int NT, int MT, int NS>
__global__ __launch_bounds__(256) void k_bounds(const uint16_t* __restrict__ xm, long n, int kpad,
                                                const uint16_t* __restrict__ cm, int k, int ld,
                                                float scale, float* __restrict__ ub) {
  bounds_body<NT, MT, NS>(blockIdx.x, gridDim.x, xm, n, kpad, cm, k, ld, scale, ub);
}

// Bounds of every centroid for one series pair (group of L lanes) from the
// weighted half-spectrum magnitudes ma, mb: ub = sum_f |X_f| |C_f| /
// (F |x| |c|), written to ubs[0, k) (series a) and ubs[k, 2k) (series b).
template <int F, int L, int P>
__device__ __forceinline__ void prune_bounds(const float* ma, const float* mb, float na, float nb,
                                             const float* __restrict__ cmag,
                                             const float* __restrict__ cnorm, int k, int lane,
                                             float* red, float* ubs) {
  constexpr int H = P / 2;
  if constexpr (L <= kWave) {
    // KC centroids at a time; the 2*KC per-lane partial sums are reduced
    // by a reduce-scatter over the L lanes (each level halves the values
    // a lane keeps)
    constexpr int KC = 4;
    for (int j0 = 0; j0 < k; j0 += KC) {
      float v[2 * KC];
#pragma unroll
      for (int c = 0; c < KC; ++c) {
        float aa = 0.f, ab = 0.f;
        if (j0 + c < k) {
          float cv[H + 1];
          load_mag<P, H + 1>(cmag + (long)(j0 + c) * F + lane * P, cv);
#pragma unroll
          for (int s = 0; s <= H; ++s) {
            aa += ma[s] * cv[s];
            ab += mb[s] * cv[s];
          }
        }
        v[c] = aa;
        v[KC + c] = ab;
      }
      int base = 0;
      reduce_scatter<2 * KC, L / 2, L>(v, lane, base);
      constexpr int kHeld = 2 * KC >= L ? 2 * KC / L : 1;
#pragma unroll
      for (int i = 0; i < kHeld; ++i) {
        const int idx = base + i, ser = idx / KC, jj = j0 + idx % KC;
        if (jj < k) {
          const float den = (ser ? nb : na) * cnorm[jj];
          ubs[ser * k + jj] = den != 0.f ? (v[i] / (float)F) / den : 0.f;
        }
      }
    }
  } else {
    for (int j = 0; j < k; ++j) {
      float cv[H + 1];
      load_mag<P, H + 1>(cmag + (long)j * F + lane * P, cv);
      float aa = 0.f, ab = 0.f;
#pragma unroll
      for (int s = 0; s <= H; ++s) {
        aa += ma[s] * cv[s];
        ab += mb[s] * cv[s];
      }
      aa = group_sum<L>(aa, red);
      ab = group_sum<L>(ab, red);
      if (lane == 0) {
        const float cn = cnorm[j];
        const float dena = na * cn, denb = nb * cn;
        ubs[j] = dena != 0.f ? (aa / (float)F) / dena : 0.f;
        ubs[k + j] = denb != 0.f ? (ab / (float)F) / denb : 0.f;
      }
    }
  }
}

// Launch geometry of the pruned assignment (that of the dense kernel).
template <int F>
struct PruneGeom {
  static constexpr int L = FftGeom<F>::L;
  static constexpr int kBlock = FftGeom<F>::kBlock;
  static constexpr int kGroups = kBlock / L;
  // waves per SIMD the register allocation must allow: three 256-thread
  // blocks fit a CU's LDS up to F = 1024 (168 VGPRs per lane)
  static constexpr int kMinWaves = F <= 1024 ? 3 : 1;
};

template <int F, bool kPre>
__device__ __forceinline__ void assign_prune_body(const int bid, const int nblk,
    const float* __restrict__ x, const float2* __restrict__ xs, const float* __restrict__ xn,
    const float* __restrict__ ubg, int ub_ld, long n, long npairs, int m, int k,
    const float2* __restrict__ cspec, const float* __restrict__ cmag,
    const float* __restrict__ cnorm, const float2* __restrict__ twg, int rev,
    const int* __restrict__ order, int* __restrict__ label, float* __restrict__ dist,
    int* __restrict__ shift, int* __restrict__ changed, unsigned long long* __restrict__ evals) {
  using G = PruneGeom<F>;
  constexpr int L = G::L, P = FftGeom<F>::P;
  extern __shared__ float2 smem[];
  float2* tw = smem;
  float* red = reinterpret_cast<float*>(smem + F + G::kGroups * fft_buf_len<F>());
  float* ubs_all = red + 64;
  for (int i = threadIdx.x; i < F; i += blockDim.x) tw[i] = twg[i];
  __syncthreads();
  const int g = threadIdx.x / L, lane = threadIdx.x % L;
  float2* buf = smem + F + g * fft_buf_len<F>();
  float* ubs = ubs_all + (long)g * 2 * k;  // [0,k): series a, [k,2k): series b
  unsigned long long my_evals = 0;
  int s_lo, s_hi;
  lag_rows(lane, L, m, F, s_lo, s_hi);
  for (long pr = (long)bid * G::kGroups + g; pr < npairs;
       pr += (long)nblk * G::kGroups) {
    long ia = order ? order[2 * pr] : 2 * pr;
    long ib = order ? (2 * pr + 1 < 2 * npairs ? order[2 * pr + 1] : -1) : 2 * pr + 1;
    if (ia >= n) ia = -1;
    if (ib >= n) ib = -1;
    if (ia < 0) {
      ia = ib;
      ib = -1;
    }
    if (ia < 0) continue;  // hole pair (uniform for the group)
    const bool hasb = ib >= 0;
    // packed spectrum zx[s] = Z[lane + L*s], Z = X_a + i X_b; without the
    // per-fit data also the weighted half-spectrum magnitudes ma, mb
    // (f = lane + L*s for s < P/2, f = F/2 at lane 0, s = P/2; weight 2
    // inside, 1 at f = 0 and f = F/2) for the bounds
    constexpr int H = P / 2;
    float2 zx[P];
    float na, nb;
    if constexpr (kPre) {
      // from the per-fit half spectra: X[f] for f <= F/2, conj(X[F-f]) above
      constexpr long HS = F / 2 + 1;
      // lane bases: f = lane + L*s below F/2 and F - f above, so that every
      // load is a base plus an immediate offset
      const float2* pa = xs + ia * HS + lane;
      const float2* pb = xs + (hasb ? ib : ia) * HS + lane;
      const float2* qa = xs + ia * HS + (F - L * H) - lane;
      const float2* qb = xs + (hasb ? ib : ia) * HS + (F - L * H) - lane;
      na = xn[ia];
      nb = hasb ? xn[ib] : 0.f;
      // all loads of the pair in flight at once (one latency per pair)
      auto assemble = [&](int s) {
        float2 A, B;
        if (s < H) {
          A = pa[L * s];
          B = hasb ? pb[L * s] : make_float2(0.f, 0.f);
        } else {
          const float2 a2 = qa[-L * (s - H)];
          const float2 b2 = hasb ? qb[-L * (s - H)] : make_float2(0.f, 0.f);
          A = make_float2(a2.x, -a2.y);
          B = make_float2(b2.x, -b2.y);
        }
        zx[s] = make_float2(A.x - B.y, A.y + B.x);
      };
#pragma unroll
      for (int s = 0; s < P; ++s) assemble(s);
      // the bounds, computed for every series by k_bounds
      for (int j = lane; j < k; j += L) {
        ubs[j] = ubg[ia * ub_ld + j];
        ubs[k + j] = hasb ? ubg[ib * ub_ld + j] : 0.f;
      }
    } else {
      const float* xa = x + ia * (long)m;
      const float* xb = x + (hasb ? ib : ia) * (long)m;
      float sa = 0.f, sb = 0.f;
#pragma unroll
      for (int s = 0; s < P; ++s) {
        const int t = lane + L * s;
        const float a = t < m ? xa[t] : 0.f;
        const float b = (hasb && t < m) ? xb[t] : 0.f;
        zx[s] = make_float2(a, b);
        sa += a * a;
        sb += b * b;
      }
      sa = group_sum<L>(sa, red);
      sb = group_sum<L>(sb, red);
      na = sqrtf(sa);
      nb = sqrtf(sb);
      fft<F, L, false>(zx, buf, tw, lane);
      // |X_a[f]| = |Z[f] + conj(Z[F-f])| / 2, |X_b[f]| = |Z[f] - conj(Z[F-f])| / 2
      float px[H + 1], py[H + 1];
      float ma[H + 1], mb[H + 1];
      partner_half<F, L, P>(zx, buf, lane, px, py);
#pragma unroll
      for (int s = 0; s <= H; ++s) {
        const float ar = zx[s].x + px[s], ai = zx[s].y - py[s];
        const float br = zx[s].x - px[s], bi = zx[s].y + py[s];
        const float wgt = s == H ? (lane == 0 ? 0.5f : 0.f) : ((s == 0 && lane == 0) ? 0.5f : 1.f);
        ma[s] = wgt * sqrtf(ar * ar + ai * ai);
        mb[s] = wgt * sqrtf(br * br + bi * bi);
      }
      prune_bounds<F, L, P>(ma, mb, na, nb, cmag, cnorm, k, lane, red, ubs);
    }
    group_sync<L>();
    float bestA = INFINITY, bestB = hasb ? INFINITY : -INFINITY;
    int labA = 0, labB = 0, keyA = 0, keyB = 0;
    // one-wave group with k <= 64: centroid j's two bounds stay in lane j's
    // registers (-1: evaluated), the next candidate is a wave max and a ballot
    const bool inreg = L == kWave && k <= kWave;  // uniform
    float myA = -1.f, myB = -1.f;
    if (inreg && lane < k) {
      myA = ubs[lane];
      myB = ubs[k + lane];
    }
    for (int round = 0; round < k; ++round) {
      // candidate with the largest bound among centroids that can still win
      float sc = -INFINITY;
      int cj = 0x7fffffff;
      if (inreg) {
        const bool oka = (myA >= 0.f) & (1.f - (myA * (1.f + kUbRel) + kUbAbs) <= bestA);
        const bool okb = (myB >= 0.f) & (1.f - (myB * (1.f + kUbRel) + kUbAbs) <= bestB);
        float v = oka ? myA : -INFINITY;
        v = okb ? fmaxf(v, myB) : v;
        sc = wave_max(v);
        if (sc > -INFINITY) cj = __builtin_ctzll(__ballot(v == sc));  // first j on ties
      } else {
        for (int j = lane; j < k; j += L) {  // j increases: strict '>' keeps the first
          const float ua = ubs[j], ub = ubs[k + j];
          const bool oka = (ua >= 0.f) & (1.f - (ua * (1.f + kUbRel) + kUbAbs) <= bestA);
          const bool okb = (ub >= 0.f) & (1.f - (ub * (1.f + kUbRel) + kUbAbs) <= bestB);
          float v = oka ? ua : -INFINITY;
          v = okb ? fmaxf(v, ub) : v;
          const bool take = v > sc;
          sc = take ? v : sc;
          cj = take ? j : cj;
        }
        group_argmax<L>(sc, cj, red);
      }
      if (!(sc > -INFINITY)) break;
      const int j = cj;
      group_sync<L>();
      if (inreg) {
        if (lane == j) {
          myA = -1.f;  // evaluated
          myB = -1.f;
        }
      } else if (lane == 0) {
        ubs[j] = -1.f;  // evaluated
        ubs[k + j] = -1.f;
      }
      const float cn = cnorm[j];
      const float dena = na * cn, denb = nb * cn;
      float da = 1.f, db = 1.f;
      int ka = 0, kb = 0;
      if (cn != 0.f) {
        ++my_evals;
        const float2* cjs = cspec + (long)j * F;
        float2 z[P];
#pragma unroll
        for (int s = 0; s < P; ++s) z[s] = cmulc(zx[s], cjs[s * L + lane]);
        fft<F, L, true>(z, buf, tw, lane);
        float va, vb;
        if constexpr (L == kWave) {
          // m through an empty asm: the row tests are then a few scalar
          // operations here instead of loop invariants held (and spilled)
          // across the whole loop
          int mo = m;
          asm volatile("" : "+s"(mo));
          lag_max_wave<P>(z, rev, lane, mo, F, va, ka, vb, kb);
        } else if constexpr (L > kWave) {
          int mo = m;
          asm volatile("" : "+s"(mo));
          lag_max_group<L, P>(z, rev, lane, mo, F, red, va, ka, vb, kb);
        } else {
          lag_argmax2<L, P>(z, s_lo, s_hi, rev, lane, m, F, va, ka, vb, kb);
          group_argmax<L>(va, ka, red);
          group_argmax<L>(vb, kb, red);
        }
        if (dena != 0.f) da = 1.f - (va / (float)F) / dena; else ka = 0;
        if (denb != 0.f) db = 1.f - (vb / (float)F) / denb; else kb = 0;
      }
      if (da < bestA || (da == bestA && j < labA)) { bestA = da; labA = j; keyA = ka; }
      if (hasb && (db < bestB || (db == bestB && j < labB))) { bestB = db; labB = j; keyB = kb; }
      group_sync<L>();
    }
    if (lane == 0) {
      if (changed && label[ia] != labA) *changed = 1;
      label[ia] = labA;
      if (dist) dist[ia] = bestA;
      if (shift) shift[ia] = keyA + 1 - m;
      if (hasb) {
        if (changed && label[ib] != labB) *changed = 1;
        label[ib] = labB;
        if (dist) dist[ib] = bestB;
        if (shift) shift[ib] = keyB + 1 - m;
      }
    }
  }
  if (evals && lane == 0 && my_evals) atomicAdd(evals, my_evals);
}

template <int F, bool kPre>
__global__ __launch_bounds__(PruneGeom<F>::kBlock, PruneGeom<F>::kMinWaves) void k_assign_prune(
    const float* __restrict__ x, const float2* __restrict__ xs, const float* __restrict__ xn,
    const float* __restrict__ ubg, int ub_ld, long n, long npairs, int m, int k,
    const float2* __restrict__ cspec, const float* __restrict__ cmag,
    const float* __restrict__ cnorm, const float2* __restrict__ twg, int rev,
    const int* __restrict__ order, int* __restrict__ label, float* __restrict__ dist,
    int* __restrict__ shift, int* __restrict__ changed, unsigned long long* __restrict__ evals) {
  assign_prune_body<F, kPre>(blockIdx.x, gridDim.x, x, xs, xn, ubg, ub_ld, n, npairs, m, k, cspec, cmag, cnorm, twg, rev, order, label, dist, shift, changed, evals);
}

// ---------------------------------------------------------- zero-fill shift
// out[t] = in[t - s] (0 outside), s = shift[row]; skipped when *skip != 0.
// roll_zeropad, test_timesC.py:55-77.  In place: one block per row, the row
// is staged in LDS.
__global__ __launch_bounds__(256) void k_apply_shift(float* __restrict__ x, long rows, int m,
                                                    const int* __restrict__ shift,
                                                    const int* __restrict__ skip) {
  extern __shared__ float row_s[];
  if (skip && *skip) return;
  for (long row = blockIdx.x; row < rows; row += gridDim.x) {
    float* r = x + row * (long)m;
    for (int t = threadIdx.x; t < m; t += blockDim.x) row_s[t] = r[t];
    __syncthreads();
    const int s = shift[row];
    for (int t = threadIdx.x; t < m; t += blockDim.x) {
      long src = (long)t - s;
      r[t] = (src >= 0 && src < m) ? row_s[src] : 0.f;
    }
    __syncthreads();
  }
}

// --------------------------------------------------------- grouped forms
// The kernels above for many jobs in one launch (GroupTable): each block
// finds its job and runs the body as that job's block bid of nblk.
#define TCX_GROUP_JOB(T)                                                       \
  const int gj_ = group_job(first, njobs, blockIdx.x);                         \
  const T a = jobs[gj_];                                                       \
  const int bid = blockIdx.x - first[gj_], nblk = first[gj_ + 1] - first[gj_];

__global__ __launch_bounds__(256) void k_znorm_group(const ZnormJob* __restrict__ jobs,
                                                     const int* __restrict__ first, int njobs) {
  TCX_GROUP_JOB(ZnormJob)
  znorm_body(bid, nblk, a.x, a.rows, a.m, a.ddof);
}

template <int NV>
__global__ __launch_bounds__(256) void k_znorm_v_group(const ZnormJob* __restrict__ jobs,
                                                       const int* __restrict__ first, int njobs) {
  TCX_GROUP_JOB(ZnormJob)
  znorm_v_body<NV>(bid, nblk, a.x, a.rows, a.m, a.ddof);
}

template <int F>
__global__ __launch_bounds__(FftGeom<F>::kBlock) void k_spectra_group(
    const SpectraJob* __restrict__ jobs, const int* __restrict__ first, int njobs,
    const float2* __restrict__ twg) {
  TCX_GROUP_JOB(SpectraJob)
  spectra_body<F>(bid, nblk, a.y, a.rows, a.m, twg, a.spec, a.norms, a.zero, a.mag, a.bmag);
}

template <int F>
__global__ __launch_bounds__(FftGeom<F>::kBlock) void k_series_spec_group(
    const SeriesSpecJob* __restrict__ jobs, const int* __restrict__ first, int njobs,
    const float2* __restrict__ twg) {
  TCX_GROUP_JOB(SeriesSpecJob)
  series_spec_body<F>(bid, nblk, a.x, a.n, a.m, twg, a.xs, a.xn, a.xm);
}

template <int NT, int MT, int NS>
__global__ __launch_bounds__(256) void k_bounds_group(const BoundsJob* __restrict__ jobs,
                                                      const int* __restrict__ first, int njobs,
                                                      int kpad, float scale) {
  TCX_GROUP_JOB(BoundsJob)
  bounds_body<NT, MT, NS>(bid, nblk, a.xm, a.n, kpad, a.cm, a.k, a.ld, scale, a.ub);
}

template <int F>
__global__ __launch_bounds__(PruneGeom<F>::kBlock, PruneGeom<F>::kMinWaves) void k_assign_prune_group(
    const PruneJob* __restrict__ jobs, const int* __restrict__ first, int njobs,
    const float2* __restrict__ twg) {
  TCX_GROUP_JOB(PruneJob)
  const long npairs = a.npad ? *a.npad / 2 : a.npairs;  // the grid covers the largest count
  assign_prune_body<F, true>(bid, nblk, a.x, a.xs, a.xn, a.ub, a.ub_ld, a.n, npairs, a.m, a.k,
                             a.cspec, a.cmag, a.cnorm, twg, a.rev, a.order, a.label, a.dist,
                             a.shift, a.changed, a.evals);
}

// ---------------------------------------------------------------- launchers
template <int F>
static size_t sbd_smem() {
  using G = FftGeom<F>;
  return sizeof(float2) * (F + G::kGroups * fft_buf_len<F>()) + sizeof(float) * 64;
}

template <int F>
static hipError_t launch_spectra_F(const float* y, int rows, int m, const float2* tw,
                                   float2* spec, float* norms, int* zero, float* mag,
                                   uint16_t* bmag, hipStream_t st) {
  using G = FftGeom<F>;
  int blocks = (rows + G::kGroups - 1) / G::kGroups;
  if (blocks < 1) blocks = 1;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL((k_spectra<F>), dim3(blocks), dim3(G::kBlock), sbd_smem<F>(), st, y,
                     rows, m, tw, spec, norms, zero, mag, bmag);
  return hipGetLastError();
}

template <int F, int MODE>
static hipError_t launch_sbd_F(const SbdArgs& a, hipStream_t st) {
  using G = FftGeom<F>;
  long npairs = (a.n + 1) / 2;
  long blocks = (npairs + G::kGroups - 1) / G::kGroups;
  if (blocks < 1) blocks = 1;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL((k_sbd<F, MODE>), dim3((unsigned)blocks), dim3(G::kBlock),
                     sbd_smem<F>(), st, a.x, a.n, a.m, a.k, a.cspec, a.cnorm, a.tw, a.rev,
                     a.label, a.dist, a.shift, a.dmat, a.changed, a.ncc_out, a.fixed_label);
  return hipGetLastError();
}

template <int F>
static size_t prune_smem(int k) {
  using G = PruneGeom<F>;
  return sizeof(float2) * (F + G::kGroups * fft_buf_len<F>()) + sizeof(float) * 64 +
         sizeof(float) * 2 * (size_t)k * G::kGroups;
}

// Largest k whose per-group bound lists fit the LDS of one workgroup.
template <int F>
static int prune_max_k_F() {
  const size_t fixed = prune_smem<F>(0);
  const size_t per_k = sizeof(float) * 2 * PruneGeom<F>::kGroups;
  const size_t fit = fixed < (size_t)kLdsBytes ? ((size_t)kLdsBytes - fixed) / per_k : 0;
  return fit < (size_t)kPruneMaxK ? (int)fit : kPruneMaxK;
}

template <int F>
static hipError_t launch_prune_F(const PruneArgs& a, hipStream_t st) {
  using G = PruneGeom<F>;
  const int tuned = tuned_prune_cap(F, a.k, a.n);  // the grid changes no result
  const long cap = tuned > 0 ? tuned : 8192;
  long blocks = (a.npairs + G::kGroups - 1) / G::kGroups;
  if (blocks < 1) blocks = 1;
  if (blocks > cap) blocks = cap;
  if (a.xs && a.ub)
    hipLaunchKernelGGL((k_assign_prune<F, true>), dim3((unsigned)blocks), dim3(G::kBlock),
                       prune_smem<F>(a.k), st, a.x, a.xs, a.xn, a.ub, a.ub_ld, a.n, a.npairs, a.m,
                       a.k, a.cspec, a.cmag, a.cnorm, a.tw, a.rev, a.order, a.label, a.dist,
                       a.shift, a.changed, a.evals);
  else
    hipLaunchKernelGGL((k_assign_prune<F, false>), dim3((unsigned)blocks), dim3(G::kBlock),
                       prune_smem<F>(a.k), st, a.x, nullptr, nullptr, nullptr, 0, a.n, a.npairs,
                       a.m, a.k, a.cspec, a.cmag, a.cnorm, a.tw, a.rev, a.order, a.label, a.dist,
                       a.shift, a.changed, a.evals);
  return hipGetLastError();
}

template <int F>
static hipError_t launch_series_spec_F(const float* x, long n, int m, const float2* tw,
                                       float2* xs, float* xn, uint16_t* xm, hipStream_t st) {
  using G = FftGeom<F>;
  long blocks = ((n + 1) / 2 + G::kGroups - 1) / G::kGroups;
  if (blocks < 1) blocks = 1;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL((k_series_spec<F>), dim3((unsigned)blocks), dim3(G::kBlock), sbd_smem<F>(),
                     st, x, n, m, tw, xs, xn, xm);
  return hipGetLastError();
}

// Host: the twiddle table of the length-F transform in the layout the
// Stockham passes read (tw_offset).
void fft_twiddle_table(int F, float2* out) {
  const int P = F / fft_lanes(F);
  const double two_pi = 6.283185307179586476925286766559;
  for (int ns = 1; ns < F;) {
    const int R = F / ns < P ? F / ns : P;
    if (ns > 1) {
      const int off = tw_offset(F, P, ns);
      for (int r = 1; r < R; ++r)
        for (int kk = 0; kk < ns; ++kk) {
          const double a = two_pi * (double)r * (double)kk / ((double)ns * (double)R);
          out[off + (r - 1) * ns + kk] = make_float2((float)std::cos(a), (float)-std::sin(a));
        }
    }
    ns *= R;
  }
}

#define TCX_FOR_EACH_F(X) \
  X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) X(8192)

hipError_t launch_spectra(int F, const float* y, int rows, int m, const float2* tw,
                          float2* spec, float* norms, int* zero, float* mag, uint16_t* bmag,
                          hipStream_t st) {
  switch (F) {
#define X(FF) \
  case FF:    \
    return launch_spectra_F<FF>(y, rows, m, tw, spec, norms, zero, mag, bmag, st);
    TCX_FOR_EACH_F(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

int prune_max_k(int F) {
  switch (F) {
#define X(FF) \
  case FF:    \
    return prune_max_k_F<FF>();
    TCX_FOR_EACH_F(X)
#undef X
    default:
      return 0;
  }
}

hipError_t launch_series_spec(int F, const float* x, long n, int m, const float2* tw, float2* xs,
                              float* xn, uint16_t* xm, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  switch (F) {
#define X(FF) \
  case FF:    \
    return launch_series_spec_F<FF>(x, n, m, tw, xs, xn, xm, st);
    TCX_FOR_EACH_F(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_bounds(int F, const uint16_t* xm, long n, const uint16_t* cm, int k, float* ub,
                         int ld, hipStream_t st) {
  if (n <= 0 || k <= 0) return hipSuccess;
  if (ld < k || F < 2) return hipErrorInvalidValue;
  const int kpad = bound_kpad(F);
  const float scale = 1.f / (float)F;  // a power of two: exact
  // MT row blocks of 16 per wave share each centroid operand load, NS
  // steps of 16 columns in flight per batch (for one centroid tile the
  // whole row up to 33 steps: F <= 1024 in one batch); 4 waves per block
  auto grid = [&](int mt) { return dim3((unsigned)((n + 64L * mt - 1) / (64L * mt))); };
  const int nsteps = kpad >> 4;
  if (k <= 16 && nsteps <= 9)
    hipLaunchKernelGGL((k_bounds<1, 1, 9>), grid(1), dim3(256), 0, st, xm, n, kpad, cm, k, ld,
                       scale, ub);
  else if (k <= 16 && nsteps <= 17)
    hipLaunchKernelGGL((k_bounds<1, 1, 17>), grid(1), dim3(256), 0, st, xm, n, kpad, cm, k, ld,
                       scale, ub);
  else if (k <= 16)
    hipLaunchKernelGGL((k_bounds<1, 1, 33>), grid(1), dim3(256), 0, st, xm, n, kpad, cm, k, ld,
                       scale, ub);
  else if (k <= 32)
    hipLaunchKernelGGL((k_bounds<2, 2, 11>), grid(2), dim3(256), 0, st, xm, n, kpad, cm, k, ld,
                       scale, ub);
  else
    hipLaunchKernelGGL((k_bounds<4, 1, 11>), grid(1), dim3(256), 0, st, xm, n, kpad, cm, k, ld,
                       scale, ub);
  return hipGetLastError();
}

hipError_t launch_assign_prune(int F, const PruneArgs& a, hipStream_t st) {
  if (a.k > prune_max_k(F)) return hipErrorInvalidValue;
  switch (F) {
#define X(FF) \
  case FF:    \
    return launch_prune_F<FF>(a, st);
    TCX_FOR_EACH_F(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_sbd(int F, int mode, const SbdArgs& a, hipStream_t st) {
  switch (F) {
#define X(FF)                                                     \
  case FF:                                                        \
    return mode == kModeAssign ? launch_sbd_F<FF, kModeAssign>(a, st) \
                               : launch_sbd_F<FF, kModeNcc>(a, st);
    TCX_FOR_EACH_F(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
}

// Kernels whose dynamic LDS exceeds the default per-launch limit.
hipError_t configure_sbd_kernels() {
  hipError_t e = hipSuccess;
#define X(FF)                                                                                 \
  if (e == hipSuccess)                                                                        \
    e = hipFuncSetAttribute((const void*)k_sbd<FF, kModeAssign>,                              \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)sbd_smem<FF>()); \
  if (e == hipSuccess)                                                                        \
    e = hipFuncSetAttribute((const void*)k_sbd<FF, kModeNcc>,                                 \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)sbd_smem<FF>()); \
  if (e == hipSuccess)                                                                        \
    e = hipFuncSetAttribute((const void*)k_spectra<FF>,                                       \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)sbd_smem<FF>()); \
  if (e == hipSuccess)                                                                        \
    e = hipFuncSetAttribute((const void*)k_assign_prune<FF, false>,                           \
                            hipFuncAttributeMaxDynamicSharedMemorySize,                       \
                            (int)prune_smem<FF>(prune_max_k_F<FF>()));                        \
  if (e == hipSuccess)                                                                        \
    e = hipFuncSetAttribute((const void*)k_assign_prune<FF, true>,                            \
                            hipFuncAttributeMaxDynamicSharedMemorySize,                       \
                            (int)prune_smem<FF>(prune_max_k_F<FF>()));                        \
  if (e == hipSuccess)                                                                        \
    e = hipFuncSetAttribute((const void*)k_series_spec<FF>,                                   \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)sbd_smem<FF>()); \
  if (e == hipSuccess)                                                                        \
    e = hipFuncSetAttribute((const void*)k_series_spec_group<FF>,                             \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)sbd_smem<FF>()); \
  if (e == hipSuccess)                                                                        \
    e = hipFuncSetAttribute((const void*)k_spectra_group<FF>,                                 \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)sbd_smem<FF>()); \
  if (e == hipSuccess)                                                                        \
    e = hipFuncSetAttribute((const void*)k_assign_prune_group<FF>,                            \
                            hipFuncAttributeMaxDynamicSharedMemorySize,                       \
                            (int)prune_smem<FF>(prune_max_k_F<FF>()));
  TCX_FOR_EACH_F(X)
#undef X
  return e;
}

hipError_t launch_znorm(float* x, long rows, int m, int ddof, hipStream_t st) {
  if (rows <= 0) return hipSuccess;
  long blocks = (rows + 3) / 4;
  if ((m & 3) == 0 && m <= 256 * 16 && (reinterpret_cast<uintptr_t>(x) & 15) == 0) {
    if (m <= 256 * 2)
      hipLaunchKernelGGL(k_znorm_v<2>, dim3((unsigned)blocks), dim3(256), 0, st, x, rows, m, ddof);
    else if (m <= 256 * 4)
      hipLaunchKernelGGL(k_znorm_v<4>, dim3((unsigned)blocks), dim3(256), 0, st, x, rows, m, ddof);
    else if (m <= 256 * 8)
      hipLaunchKernelGGL(k_znorm_v<8>, dim3((unsigned)blocks), dim3(256), 0, st, x, rows, m, ddof);
    else
      hipLaunchKernelGGL(k_znorm_v<16>, dim3((unsigned)blocks), dim3(256), 0, st, x, rows, m,
                         ddof);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_znorm, dim3((unsigned)blocks), dim3(256), 0, st, x, rows, m, ddof);
  return hipGetLastError();
}

// ------------------------------------------------ grouped-form launchers
long znorm_blocks(long rows) { return (rows + 3) / 4; }

int znorm_vectors(const float* x, int m) {
  if ((m & 3) != 0 || m > 256 * 16 || (reinterpret_cast<uintptr_t>(x) & 15) != 0) return 0;
  return m <= 256 * 2 ? 2 : (m <= 256 * 4 ? 4 : (m <= 256 * 8 ? 8 : 16));
}

hipError_t launch_znorm_group(const GroupTable& g, int nv, hipStream_t st) {
  if (g.blocks <= 0) return hipSuccess;
  const ZnormJob* j = static_cast<const ZnormJob*>(g.jobs);
  switch (nv) {
    case 0:
      hipLaunchKernelGGL(k_znorm_group, dim3(g.blocks), dim3(256), 0, st, j, g.first, g.njobs);
      break;
    case 2:
      hipLaunchKernelGGL(k_znorm_v_group<2>, dim3(g.blocks), dim3(256), 0, st, j, g.first,
                         g.njobs);
      break;
    case 4:
      hipLaunchKernelGGL(k_znorm_v_group<4>, dim3(g.blocks), dim3(256), 0, st, j, g.first,
                         g.njobs);
      break;
    case 8:
      hipLaunchKernelGGL(k_znorm_v_group<8>, dim3(g.blocks), dim3(256), 0, st, j, g.first,
                         g.njobs);
      break;
    case 16:
      hipLaunchKernelGGL(k_znorm_v_group<16>, dim3(g.blocks), dim3(256), 0, st, j, g.first,
                         g.njobs);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int F>
static long groups_blocks(long units) {  // launch_spectra_F / launch_*_F block counts
  long b = (units + FftGeom<F>::kGroups - 1) / FftGeom<F>::kGroups;
  return b < 1 ? 1 : b;
}

long series_spec_blocks(int F, long n) {
  switch (F) {
#define X(FF) \
  case FF:    \
    return std::min(groups_blocks<FF>((n + 1) / 2), 8192L);
    TCX_FOR_EACH_F(X)
#undef X
    default:
      return 0;
  }
}

long spectra_blocks(int F, long rows) {
  switch (F) {
#define X(FF) \
  case FF:    \
    return std::min(groups_blocks<FF>(rows), 4096L);
    TCX_FOR_EACH_F(X)
#undef X
    default:
      return 0;
  }
}

long prune_blocks(int F, long npairs, int k, long n) {
  const int tuned = tuned_prune_cap(F, k, n);  // the grid changes no result
  const long cap = tuned > 0 ? tuned : 8192;
  switch (F) {
#define X(FF)                                                                   \
  case FF: {                                                                    \
    long b = (npairs + PruneGeom<FF>::kGroups - 1) / PruneGeom<FF>::kGroups;     \
    return b < 1 ? 1 : (b > cap ? cap : b);                                     \
  }
    TCX_FOR_EACH_F(X)
#undef X
    default:
      return 0;
  }
}

size_t prune_lds(int F, int k) {
  switch (F) {
#define X(FF) \
  case FF:    \
    return prune_smem<FF>(k);
    TCX_FOR_EACH_F(X)
#undef X
    default:
      return 0;
  }
}

hipError_t launch_series_spec_group(int F, const GroupTable& g, const float2* tw, hipStream_t st) {
  if (g.blocks <= 0) return hipSuccess;
  const SeriesSpecJob* j = static_cast<const SeriesSpecJob*>(g.jobs);
  switch (F) {
#define X(FF)                                                                                    \
  case FF:                                                                                       \
    hipLaunchKernelGGL((k_series_spec_group<FF>), dim3(g.blocks), dim3(FftGeom<FF>::kBlock),      \
                       sbd_smem<FF>(), st, j, g.first, g.njobs, tw);                            \
    break;
    TCX_FOR_EACH_F(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_spectra_group(int F, const GroupTable& g, const float2* tw, hipStream_t st) {
  if (g.blocks <= 0) return hipSuccess;
  const SpectraJob* j = static_cast<const SpectraJob*>(g.jobs);
  switch (F) {
#define X(FF)                                                                                    \
  case FF:                                                                                       \
    hipLaunchKernelGGL((k_spectra_group<FF>), dim3(g.blocks), dim3(FftGeom<FF>::kBlock),          \
                       sbd_smem<FF>(), st, j, g.first, g.njobs, tw);                            \
    break;
    TCX_FOR_EACH_F(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// The forms of launch_bounds: 0 <1,1,9>, 1 <1,1,17>, 2 <1,1,33> (k <= 16),
// 3 <2,2,11> (k <= 32), 4 <4,1,11>.
int bounds_form(int F, int k) {
  const int nsteps = bound_kpad(F) >> 4;
  if (k <= 16) return nsteps <= 9 ? 0 : (nsteps <= 17 ? 1 : 2);
  return k <= 32 ? 3 : 4;
}

long bounds_blocks(int form, long n) {
  const long mt = form == 3 ? 2 : 1;
  return (n + 64 * mt - 1) / (64 * mt);
}

hipError_t launch_bounds_group(int F, int form, const GroupTable& g, hipStream_t st) {
  if (g.blocks <= 0) return hipSuccess;
  const BoundsJob* j = static_cast<const BoundsJob*>(g.jobs);
  const int kpad = bound_kpad(F);
  const float scale = 1.f / (float)F;
  switch (form) {
    case 0:
      hipLaunchKernelGGL((k_bounds_group<1, 1, 9>), dim3(g.blocks), dim3(256), 0, st, j, g.first,
                         g.njobs, kpad, scale);
      break;
    case 1:
      hipLaunchKernelGGL((k_bounds_group<1, 1, 17>), dim3(g.blocks), dim3(256), 0, st, j, g.first,
                         g.njobs, kpad, scale);
      break;
    case 2:
      hipLaunchKernelGGL((k_bounds_group<1, 1, 33>), dim3(g.blocks), dim3(256), 0, st, j, g.first,
                         g.njobs, kpad, scale);
      break;
    case 3:
      hipLaunchKernelGGL((k_bounds_group<2, 2, 11>), dim3(g.blocks), dim3(256), 0, st, j, g.first,
                         g.njobs, kpad, scale);
      break;
    case 4:
      hipLaunchKernelGGL((k_bounds_group<4, 1, 11>), dim3(g.blocks), dim3(256), 0, st, j, g.first,
                         g.njobs, kpad, scale);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_assign_prune_group(int F, const GroupTable& g, const float2* tw,
                                     hipStream_t st) {
  if (g.blocks <= 0) return hipSuccess;
  const PruneJob* j = static_cast<const PruneJob*>(g.jobs);
  switch (F) {
#define X(FF)                                                                                    \
  case FF:                                                                                       \
    if (g.lds > prune_smem<FF>(prune_max_k_F<FF>())) return hipErrorInvalidValue;                \
    hipLaunchKernelGGL((k_assign_prune_group<FF>), dim3(g.blocks), dim3(PruneGeom<FF>::kBlock),   \
                       g.lds, st, j, g.first, g.njobs, tw);                                     \
    break;
    TCX_FOR_EACH_F(X)
#undef X
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_apply_shift(float* x, long rows, int m, const int* shift, const int* skip,
                              hipStream_t st) {
  if (rows <= 0) return hipSuccess;
  long blocks = rows < 4096 ? rows : 4096;
  hipLaunchKernelGGL(k_apply_shift, dim3((unsigned)blocks), dim3(256), sizeof(float) * m, st,
                     x, rows, m, shift, skip);
  return hipGetLastError();
}

}  // namespace tcx

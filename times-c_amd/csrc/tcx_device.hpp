// Device-side building blocks for gfx950 (CDNA4, wave64): complex helpers,
// lane-group reductions and the register/LDS Stockham FFT used by every SBD
// kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tcx {

constexpr int kWave = 64;

// ------------------------------------------------------------- complex ops
__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// ------------------------------------------------------ grouped launches
// Job of block b in a grouped launch: the largest j with first[j] <= b
// (first: prefix of the jobs' block counts, jobs without blocks skipped).
__device__ __forceinline__ int group_job(const int* __restrict__ first, int njobs, int b) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// ------------------------------------------------------- synchronisation
// LDS exchange inside one wavefront: a wave's LDS operations complete in
// issue order, so only the compiler must be kept from reordering.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Device-scope 64-bit stores and loads (relaxed atomics at agent scope:
// coherent across the XCDs' L2 caches without a cache-wide fence), used to
// exchange values between the blocks of one launch.
__device__ __forceinline__ void st_dev_u64(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_dev_u64(const uint64_t* p) {
  return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev_f64(uint64_t* p, double v) {
  st_dev_u64(p, (uint64_t)__double_as_longlong(v));
}
__device__ __forceinline__ double ld_dev_f64(const uint64_t* p) {
  return __longlong_as_double((long long)ld_dev_u64(p));
}
// Waits until this wave's vector-memory operations have completed; a
// device-scope store that completed is visible to every block.
__device__ __forceinline__ void wait_vmem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// IEEE binary16 bits of v >= 0 rounded toward +infinity: the nearest
// value (hardware conversion), stepped to the next binary16 when it fell
// below v (for v >= 0 the next larger bit pattern is the next value up).
// NaN stays NaN.
__device__ __forceinline__ uint16_t f16_bits_up(float v) {
  const _Float16 h = (_Float16)v;
  const uint16_t b = __builtin_bit_cast(uint16_t, h);
  return (float)h < v ? (uint16_t)(b + 1) : b;
}

// Sync of an FFT lane group of L lanes: wave-local when the group lives in
// one wave, a block barrier otherwise (then the group is the whole block).
template <int L>
__device__ __forceinline__ void group_sync() {
  if constexpr (L <= kWave)
    wave_lds_sync();
  else
    __syncthreads();
}

// ------------------------------------------------------ group reductions
// Reductions over L lanes.  For L <= 64 the group is a contiguous lane
// range of one wave; for L > 64 the group is the whole block and `scratch`
// (at least 2*L/64 words of LDS) carries the per-wave partials.
// Inside a wave the partners come from DPP (lane ^ 1, lane ^ 2, the mirror
// in 8 lanes, the mirror in a row of 16), then the four row results are
// read out with readlane: VALU/SALU only, no LDS.  Every lane of a group
// ends with the same value.
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float readlane_f32(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

template <int L>
__device__ __forceinline__ float group_sum(float v, float* scratch) {
  constexpr int W = L < kWave ? L : kWave;
  if constexpr (W >= 2) v += dpp_f32<0xB1>(v);   // lane ^ 1
  if constexpr (W >= 4) v += dpp_f32<0x4E>(v);   // lane ^ 2
  if constexpr (W >= 8) v += dpp_f32<0x141>(v);  // mirror in 8 lanes
  if constexpr (W >= 16) v += dpp_f32<0x140>(v); // mirror in 16 lanes
  if constexpr (W == 32) {
    const float a = readlane_f32(v, 0) + readlane_f32(v, 16);
    const float b = readlane_f32(v, 32) + readlane_f32(v, 48);
    v = (threadIdx.x & 32) ? b : a;
  } else if constexpr (W == 64) {
    v = (readlane_f32(v, 0) + readlane_f32(v, 16)) + (readlane_f32(v, 32) + readlane_f32(v, 48));
  }
  if constexpr (L > kWave) {
    constexpr int NW = L / kWave;
    int wid = threadIdx.x / kWave;
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) scratch[wid] = v;
    __syncthreads();
    float s = 0.f;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += scratch[w];
    v = s;
  }
  return v;
}

// Max with the smallest key on ties (NumPy argmax: first index wins).  The
// pair travels as one 64-bit word: the value mapped to an unsigned integer
// of the same order (-0 folded into +0, so the two compare equal as floats
// do) above the complemented key, so the largest word is the largest value
// with the smallest key, and each reduction step is one 64-bit compare.  The
// value is never NaN here (the lag scans skip NaN); a lane without a lag
// holds (-inf, INT_MAX), which loses to every lag.
__device__ __forceinline__ uint64_t argmax_pack(float v, int key) {
  const uint32_t b = __float_as_uint(v + 0.0f);  // -0 + 0 = +0
  const uint32_t o = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  return ((uint64_t)o << 32) | (uint32_t)~key;
}
__device__ __forceinline__ void argmax_unpack(uint64_t p, float& v, int& key) {
  const uint32_t o = (uint32_t)(p >> 32);
  v = __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
  key = (int)~(uint32_t)p;
}
template <int CTRL>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t p) {
  const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(p >> 32), CTRL, 0xF, 0xF,
                                                         false);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)p, CTRL, 0xF, 0xF, false);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t p, int lane) {
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(p >> 32), lane);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)p, lane);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t max_u64(uint64_t a, uint64_t b) { return b > a ? b : a; }

template <int L>
__device__ __forceinline__ void group_argmax(float& v, int& key, float* scratch) {
  constexpr int W = L < kWave ? L : kWave;
  uint64_t p = argmax_pack(v, key);
  if constexpr (W >= 2) p = max_u64(p, dpp_u64<0xB1>(p));   // lane ^ 1
  if constexpr (W >= 4) p = max_u64(p, dpp_u64<0x4E>(p));   // lane ^ 2
  if constexpr (W >= 8) p = max_u64(p, dpp_u64<0x141>(p));  // mirror in 8 lanes
  if constexpr (W >= 16) p = max_u64(p, dpp_u64<0x140>(p)); // mirror in 16 lanes
  if constexpr (W == 32) {
    const uint64_t a = max_u64(readlane_u64(p, 0), readlane_u64(p, 16));
    const uint64_t b = max_u64(readlane_u64(p, 32), readlane_u64(p, 48));
    p = (threadIdx.x & 32) ? b : a;
  } else if constexpr (W == 64) {
    p = max_u64(max_u64(readlane_u64(p, 0), readlane_u64(p, 16)),
                max_u64(readlane_u64(p, 32), readlane_u64(p, 48)));
  }
  if constexpr (L > kWave) {
    constexpr int NW = L / kWave;
    const int wid = threadIdx.x / kWave;
    uint32_t* hs = reinterpret_cast<uint32_t*>(scratch);
    uint32_t* ls = hs + NW;
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) {
      hs[wid] = (uint32_t)(p >> 32);
      ls[wid] = (uint32_t)p;
    }
    __syncthreads();
    p = ((uint64_t)hs[0] << 32) | ls[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) p = max_u64(p, ((uint64_t)hs[w] << 32) | ls[w]);
  }
  argmax_unpack(p, v, key);
}

// group_argmax for a 64-lane group whose values all sit in lanes 0..15 (the
// other lanes hold (-inf, INT_MAX)): the reduction stays inside the first
// row of 16 lanes and lane 0's result is broadcast.
__device__ __forceinline__ void group_argmax_row0(float& v, int& key) {
  uint64_t p = argmax_pack(v, key);
  p = max_u64(p, dpp_u64<0xB1>(p));
  p = max_u64(p, dpp_u64<0x4E>(p));
  p = max_u64(p, dpp_u64<0x141>(p));
  p = max_u64(p, dpp_u64<0x140>(p));
  argmax_unpack(readlane_u64(p, 0), v, key);
}

// Cross-lane moves of a double through DPP (VALU data-parallel primitives,
// no LDS round trip): CTRL is the DPP control word.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, false),
                          __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// Whole-wave shifts by one lane (DPP wave_shr:1 / wave_shl:1): lane l gets
// v of lane l-1 (resp. l+1); lane 0 (resp. 63) gets fill.
__device__ __forceinline__ double wave_shr1_f64(double v, double fill) {
  return __hiloint2double(
      __builtin_amdgcn_update_dpp(__double2hiint(fill), __double2hiint(v), 0x138, 0xF, 0xF, false),
      __builtin_amdgcn_update_dpp(__double2loint(fill), __double2loint(v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ double wave_shl1_f64(double v, double fill) {
  return __hiloint2double(
      __builtin_amdgcn_update_dpp(__double2hiint(fill), __double2hiint(v), 0x130, 0xF, 0xF, false),
      __builtin_amdgcn_update_dpp(__double2loint(fill), __double2loint(v), 0x130, 0xF, 0xF, false));
}

// Maximum over the 64 lanes of a wave (all lanes active), same pattern as
// wave_sum_f64.
__device__ __forceinline__ double wave_max_f64(double v) {
  v = fmax(v, dpp_f64<0xB1>(v));
  v = fmax(v, dpp_f64<0x4E>(v));
  v = fmax(v, dpp_f64<0x124>(v));
  v = fmax(v, dpp_f64<0x128>(v));
  return fmax(fmax(readlane_f64(v, 0), readlane_f64(v, 16)),
              fmax(readlane_f64(v, 32), readlane_f64(v, 48)));
}

// Sum over the 64 lanes of a wave (all lanes active); every lane gets the
// same value (fixed order: quads, rows of 16, then the four rows).
__device__ __forceinline__ double wave_sum_f64(double v) {
  v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]: lane ^ 1
  v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]: lane ^ 2
  v += dpp_f64<0x124>(v);  // row_ror:4
  v += dpp_f64<0x128>(v);  // row_ror:8 -> each lane holds its row's sum
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// ---------------------------------------------------------- in-register DFT
// exp(-2 pi i q / 32): table of cos / sin used for compile-time butterflies.
constexpr float kCos32[32] = {
    1.0f, 0.98078528040323043f, 0.92387953251128674f, 0.83146961230254524f,
    0.70710678118654757f, 0.55557023301960229f, 0.38268343236508984f, 0.19509032201612833f,
    0.0f, -0.19509032201612819f, -0.38268343236508973f, -0.55557023301960196f,
    -0.70710678118654746f, -0.83146961230254535f, -0.92387953251128674f, -0.98078528040323043f,
    -1.0f, -0.98078528040323043f, -0.92387953251128685f, -0.83146961230254546f,
    -0.70710678118654768f, -0.55557023301960218f, -0.38268343236509034f, -0.19509032201612866f,
    0.0f, 0.1950903220161283f, 0.38268343236509f, 0.55557023301960184f,
    0.70710678118654735f, 0.83146961230254524f, 0.92387953251128652f, 0.98078528040323032f,
};
constexpr float kSin32[32] = {
    0.0f, 0.19509032201612825f, 0.38268343236508978f, 0.55557023301960218f,
    0.70710678118654746f, 0.83146961230254524f, 0.92387953251128674f, 0.98078528040323043f,
    1.0f, 0.98078528040323043f, 0.92387953251128674f, 0.83146961230254546f,
    0.70710678118654757f, 0.55557023301960218f, 0.38268343236508989f, 0.19509032201612861f,
    0.0f, -0.19509032201612836f, -0.38268343236508967f, -0.55557023301960196f,
    -0.70710678118654746f, -0.83146961230254524f, -0.92387953251128652f, -0.98078528040323032f,
    -1.0f, -0.98078528040323043f, -0.92387953251128663f, -0.83146961230254546f,
    -0.70710678118654768f, -0.55557023301960218f, -0.38268343236509039f, -0.19509032201612872f,
};

// v * W_R^k, W_R = exp(-2 pi i / R) (forward) or its conjugate (inverse).
template <int R, int K, bool INV>
__device__ __forceinline__ float2 twmul(float2 v) {
  constexpr int q = (K * (32 / R)) & 31;
  if constexpr (q == 0) {
    return v;
  } else if constexpr (q == 8) {  // forward: -i
    return INV ? make_float2(-v.y, v.x) : make_float2(v.y, -v.x);
  } else if constexpr (q == 16) {
    return make_float2(-v.x, -v.y);
  } else if constexpr (q == 24) {  // forward: +i
    return INV ? make_float2(v.y, -v.x) : make_float2(-v.y, v.x);
  } else {
    constexpr float c = kCos32[q];
    constexpr float s = INV ? kSin32[q] : -kSin32[q];
    return make_float2(v.x * c - v.y * s, v.x * s + v.y * c);
  }
}

template <int R, bool INV, int K>
__device__ __forceinline__ void dft_combine(float2* v, const float2* e, const float2* o) {
  if constexpr (K < R / 2) {
    float2 t = twmul<R, K, INV>(o[K]);
    v[K] = cadd(e[K], t);
    v[K + R / 2] = csub(e[K], t);
    dft_combine<R, INV, K + 1>(v, e, o);
  }
}

// Length-R DFT (R <= 32, power of two) of v[0..R-1], natural order in and
// out, radix-2 decimation in time with compile-time twiddles.
template <int R, bool INV>
__device__ __forceinline__ void dft(float2* v) {
  static_assert(R >= 1 && R <= 32 && (R & (R - 1)) == 0, "radix");
  if constexpr (R == 2) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R > 2) {
    float2 e[R / 2], o[R / 2];
#pragma unroll
    for (int i = 0; i < R / 2; ++i) {
      e[i] = v[2 * i];
      o[i] = v[2 * i + 1];
    }
    dft<R / 2, INV>(e);
    dft<R / 2, INV>(o);
    dft_combine<R, INV, 0>(v, e, o);
  }
}

// --------------------------------------------------------- Stockham FFT
// A length-F complex FFT held by a group of L lanes, P = F/L values per
// lane, "lane-strided natural order" on input and output:
//     v[s] of lane l  <->  element l + L*s.
// Passes are radix P (the first reads registers directly, the last leaves
// its output in registers); between passes the data goes through the
// group's LDS buffer `buf`.  `tw` holds exp(-2 pi i q / F), q < F, in LDS.
// The forward transform's output order equals the inverse transform's input
// order, so spectra are multiplied in place without any reordering.
__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 5); }
template <int F>
constexpr int fft_buf_len() {
  return F + F / 32 + 1;
}

// Lanes of an FFT group of length F (see FftGeom).
__host__ __device__ constexpr int fft_lanes(int F) {
  return F <= 4 ? 1 : F <= 64 ? F / 4 : F <= 256 ? 16 : F / 16;
}

// Twiddle table layout: the passes of a length-F transform with P values
// per lane have strides Ns = 1, R_1, R_1 R_2, ... (R = min(F / Ns, P)); the
// pass with stride Ns > 1 reads entries off(Ns) + (r - 1) Ns + kk =
// exp(-2 pi i r kk / (Ns R)), r = 1..R-1, kk = 0..Ns-1, so the lanes of a
// wave read consecutive entries (no LDS bank conflicts).  The table has
// F - R_1 < F entries.
__host__ __device__ constexpr int tw_offset(int F, int P, int Ns) {
  int off = 0;
  for (int ns = 1; ns < Ns;) {
    const int R = F / ns < P ? F / ns : P;
    if (ns > 1) off += (R - 1) * ns;
    ns *= R;
  }
  return off;
}

template <int F, int L, bool INV, int Ns>
__device__ __forceinline__ void stockham_pass(float2 (&v)[F / L], float2* buf,
                                              const float2* tw, int lane) {
  constexpr int P = F / L;
  constexpr int R = (F / Ns < P) ? F / Ns : P;
  constexpr int NB = P / R;
  constexpr bool kFirst = (Ns == 1);
  constexpr bool kLast = (Ns * R == F);
  static_assert(!kFirst || NB == 1, "first pass is one radix-P butterfly");
  float2 u[P];
  if constexpr (kFirst) {
#pragma unroll
    for (int r = 0; r < P; ++r) u[r] = v[r];
  } else {
    // element lane + L*t + r*(F/R): with F/R a multiple of 32 the offset
    // r*(F/R) keeps the low five bits of lane + L*t
    if constexpr ((F / R) % 32 == 0) {
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        const int pb = lds_pad(lane + L * t);
        const float2* src = buf + pb;
        constexpr int kStep = (F / R) + (F / R) / 32;
#pragma unroll
        for (int r = 0; r < R; ++r) u[t * R + r] = src[r * kStep];
      }
    } else {
#pragma unroll
      for (int t = 0; t < NB; ++t)
#pragma unroll
        for (int r = 0; r < R; ++r)
          u[t * R + r] = buf[lds_pad(lane + L * t + r * (F / R))];
    }
    if constexpr (L > kWave) __syncthreads();  // reads done before writes
  }
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    const int j = lane + L * t;
    const int kk = j & (Ns - 1);
    float2* b = u + t * R;
    if constexpr (!kFirst) {
      constexpr int kOff = tw_offset(F, P, Ns);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const float2 w = tw[kOff + (r - 1) * Ns + kk];
        b[r] = INV ? cmulc(b[r], w) : cmul(b[r], w);
      }
    }
    dft<R, INV>(b);
    if constexpr (!kLast) {
      // base + r*Ns: (j-kk)*R is a multiple of Ns*R and kk < Ns, so
      // (base & 31) + ((r*Ns) & 31) < 32 for every r < R (Ns, R powers of
      // two): one padded base plus immediate offsets
      const int base = (j - kk) * R + kk;
      float2* dst = buf + lds_pad(base);
      constexpr int kStep = Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * kStep + ((r * kStep) >> 5)] = b[r];
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) v[t + r * NB] = b[r];
    }
  }
  if constexpr (!kLast) {
    group_sync<L>();
    stockham_pass<F, L, INV, Ns * R>(v, buf, tw, lane);
  }
}

// Unnormalized transform: forward uses exp(-2 pi i f t / F), inverse
// exp(+2 pi i f t / F) (no 1/F).
template <int F, int L, bool INV>
__device__ __forceinline__ void fft(float2 (&v)[F / L], float2* buf,
                                    const float2* tw, int lane) {
  stockham_pass<F, L, INV, 1>(v, buf, tw, lane);
}

// Lane-group geometry per FFT length: L lanes, P = F/L values per lane.
template <int F>
struct FftGeom {
  static constexpr int L = fft_lanes(F);
  static constexpr int P = F / L;
  static constexpr int kBlock = L <= kWave ? 256 : L;  // threads per block
  static constexpr int kGroups = kBlock / L;           // lane groups per block
};

// Lag key of circular correlation position tau (0 <= tau < F): the index in
// NumPy's concatenation cc[-(m-1):] ++ cc[:m] (test_timesC.py:113), or -1
// if tau is not one of the 2m-1 lags.  rev=1 gives the index in the
// reversed array, i.e. the order of cc(centroid, member).
__device__ __forceinline__ int lag_key(int tau, int m, int F, int rev) {
  int idx = tau <= m - 1 ? tau + m - 1 : (tau >= F - (m - 1) ? tau - F + m - 1 : -1);
  if (idx >= 0 && rev) idx = 2 * m - 2 - idx;
  return idx;
}

// Per-lane argmax over the lags of two inverse transforms packed as
// z[s].x, z[s].y at tau = lane + L*s (P values per lane, L*P = F).  tau is
// a lag when tau <= m-1 (s <= s_lo) or tau >= F-m+1 (s >= s_hi); lag_rows
// gives this lane's s_lo and s_hi.  tau < F/2 holds the lags 0..m-1 (keys
// m-1+tau) and tau >= F/2 the negative lags (keys tau-F+m-1, all smaller),
// so visiting s = P/2..P-1 and then 0..P/2-1 meets the keys in increasing
// order (the reverse order when rev, whose keys are mirrored); a strict '>'
// then keeps the smallest key among equal values.  No lag taken (all -inf
// or NaN): key INT_MAX.
__device__ __forceinline__ void lag_rows(int lane, int L, int m, int F, int& s_lo, int& s_hi) {
  const int a = m - 1 - lane;                 // tau <= m-1  <=>  L*s <= a
  s_lo = a >= 0 ? a / L : -1;
  const int b = F - (m - 1) - lane;           // tau >= F-m+1  <=>  L*s >= b
  s_hi = b <= 0 ? 0 : (b + L - 1) / L;
}

template <int L, int P>
__device__ __forceinline__ void lag_argmax2(const float2* z, int s_lo, int s_hi, int rev, int lane,
                                            int m, int F, float& va, int& ka, float& vb,
                                            int& kb) {
  float bx = -INFINITY, by = -INFINITY;
  int sx = 0, sy = 0;
  // s_lo < P/2 <= s_hi, so each half needs one of the two tests
  auto visit = [&](int s) {
    const bool ok = s < P / 2 ? s <= s_lo : s >= s_hi;
    const float zx = ok ? z[s].x : -INFINITY;
    const float zy = ok ? z[s].y : -INFINITY;
    const bool tx = zx > bx, ty = zy > by;
    bx = tx ? zx : bx;
    sx = tx ? s : sx;
    by = ty ? zy : by;
    sy = ty ? s : sy;
  };
  if (!rev) {
#pragma unroll
    for (int s = P / 2; s < P; ++s) visit(s);
#pragma unroll
    for (int s = 0; s < P / 2; ++s) visit(s);
  } else {
#pragma unroll
    for (int s = P / 2 - 1; s >= 0; --s) visit(s);
#pragma unroll
    for (int s = P - 1; s >= P / 2; --s) visit(s);
  }
  va = bx;
  vb = by;
  ka = bx > -INFINITY ? lag_key(lane + L * sx, m, F, rev) : 0x7fffffff;
  kb = by > -INFINITY ? lag_key(lane + L * sy, m, F, rev) : 0x7fffffff;
}

// Order-preserving map of a float's bits to a signed integer (a < b as
// floats => ord(a) < ord(b); -0 maps below +0); its own inverse.
__device__ __forceinline__ int f32_ord(float v) {
  const int b = __float_as_int(v);
  return b ^ ((b >> 31) & 0x7fffffff);
}
__device__ __forceinline__ float ord_f32(int k) { return __int_as_float(k ^ ((k >> 31) & 0x7fffffff)); }
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}

// Wave maximum of v, which must not be NaN (every lane gets it), on the
// integer images: no NaN canonicalization between the DPP steps.
__device__ __forceinline__ float wave_max(float v) {
  int k = f32_ord(v);
  k = max(k, dpp_i32<0xB1>(k));   // lane ^ 1
  k = max(k, dpp_i32<0x4E>(k));   // lane ^ 2
  k = max(k, dpp_i32<0x141>(k));  // mirror in 8 lanes
  k = max(k, dpp_i32<0x140>(k));  // mirror in 16 lanes
  const int a = max(__builtin_amdgcn_readlane(k, 0), __builtin_amdgcn_readlane(k, 16));
  const int b = max(__builtin_amdgcn_readlane(k, 32), __builtin_amdgcn_readlane(k, 48));
  return ord_f32(max(a, b));
}

// lag_argmax2 followed by group_argmax of both results, for a one-wave
// group (L = 64): the same values and keys, fewer instructions.  Each
// series' maximum g is a per-lane max (NaN never wins, as the scan's '>'
// never takes it) then a wave max; its key is the first lag holding g in
// key order, found by one ballot per row visited in key order (rows
// P/2..P-1 then 0..P/2-1, ascending lanes; mirrored when rev, whose keys
// run the other way), so ties keep the smallest key; a row holding g is
// then masked to its lags (tau in [m, F-m] is no lag).  g is first taken
// over all F positions; only if no lag holds it (every lag below some
// non-lag: the non-lags hold the circular wrap, about 0) is it taken again
// over the lags alone.
template <int P>
__device__ __forceinline__ void lag_max_wave(const float2 (&z)[P], int rev, int lane, int m,
                                             int F, float& va, int& ka, float& vb, int& kb) {
  constexpr int L = kWave;
  auto is_lag = [&](int s) {
    const int tau = lane + L * s;
    return tau <= m - 1 || tau >= F - m + 1;
  };
  auto find = [&](float g, bool isx, int& key) -> bool {
    key = 0x7fffffff;
    if (!(g > -INFINITY)) return true;  // no lag taken (uniform)
    auto row = [&](int s) -> bool {
      uint64_t b = __ballot((isx ? z[s].x : z[s].y) == g);
      if (!b) return false;
      b &= __ballot(is_lag(s));  // only on a row holding g
      if (!b) return false;
      const int ln = rev ? 63 - __builtin_clzll(b) : __builtin_ctzll(b);
      key = lag_key(ln + L * s, m, F, rev);
      return true;
    };
    if (!rev) {
#pragma unroll
      for (int s = P / 2; s < P; ++s)
        if (row(s)) return true;
#pragma unroll
      for (int s = 0; s < P / 2; ++s)
        if (row(s)) return true;
    } else {
#pragma unroll
      for (int s = P / 2 - 1; s >= 0; --s)
        if (row(s)) return true;
#pragma unroll
      for (int s = P - 1; s >= P / 2; --s)
        if (row(s)) return true;
    }
    return false;
  };
  float mx = -INFINITY, my = -INFINITY;
#pragma unroll
  for (int s = 0; s < P; ++s) {
    mx = fmaxf(mx, z[s].x);
    my = fmaxf(my, z[s].y);
  }
  va = wave_max(mx);
  vb = wave_max(my);
  if (find(va, true, ka) && find(vb, false, kb)) return;
  // rare: the maximum over the lags alone
  mx = -INFINITY;
  my = -INFINITY;
#pragma unroll
  for (int s = 0; s < P; ++s) {
    const bool ok = is_lag(s);
    mx = ok ? fmaxf(mx, z[s].x) : mx;
    my = ok ? fmaxf(my, z[s].y) : my;
  }
  va = wave_max(mx);
  vb = wave_max(my);
  find(va, true, ka);
  find(vb, false, kb);
}

// lag_max_wave for a group of L = 64 * NW lanes (NW waves, the whole
// block, P values per lane): each wave takes the maximum of its lanes; the
// group maximum g goes through LDS (`red`, 2 * NW words); each wave then
// finds its first lag holding g in key order over its own lanes (rows in
// key order; within a row the keys run with the group lane, so the first
// key of the group is the smallest of the waves' first keys), and the
// smallest key goes through LDS again.  If no lag holds g, the same over
// the lags alone.  Equal to lag_argmax2 followed by group_argmax.
template <int L, int P>
__device__ __forceinline__ void lag_max_group(const float2 (&z)[P], int rev, int glane, int m,
                                              int F, float* red, float& va, int& ka, float& vb,
                                              int& kb) {
  static_assert(L > kWave && L % kWave == 0, "multi-wave group");
  constexpr int NW = L / kWave;
  const int w = glane / kWave, lane = glane % kWave;
  int* ir = reinterpret_cast<int*>(red);
  auto is_lag = [&](int s) {
    const int tau = glane + L * s;
    return tau <= m - 1 || tau >= F - m + 1;
  };
  // the group maxima of the two series from per-lane maxima
  auto group_max = [&](float mx, float my, float& gx, float& gy) {
    const float wx = wave_max(mx), wy = wave_max(my);
    __syncthreads();  // earlier users of red are done
    if (lane == 0) {
      ir[w] = f32_ord(wx);
      ir[NW + w] = f32_ord(wy);
    }
    __syncthreads();
    int ax = ir[0], ay = ir[NW];
#pragma unroll
    for (int u = 1; u < NW; ++u) {
      ax = max(ax, ir[u]);
      ay = max(ay, ir[NW + u]);
    }
    gx = ord_f32(ax);
    gy = ord_f32(ay);
  };
  // this wave's first lag key holding g (INT_MAX if none)
  auto find = [&](float g, bool isx) -> int {
    if (!(g > -INFINITY)) return 0x7fffffff;
    auto row = [&](int s, int& key) -> bool {
      uint64_t b = __ballot((isx ? z[s].x : z[s].y) == g);
      if (!b) return false;
      b &= __ballot(is_lag(s));
      if (!b) return false;
      const int ln = rev ? 63 - __builtin_clzll(b) : __builtin_ctzll(b);
      key = lag_key(ln + kWave * w + L * s, m, F, rev);
      return true;
    };
    int key = 0x7fffffff;
    if (!rev) {
#pragma unroll
      for (int s = P / 2; s < P; ++s)
        if (row(s, key)) return key;
#pragma unroll
      for (int s = 0; s < P / 2; ++s)
        if (row(s, key)) return key;
    } else {
#pragma unroll
      for (int s = P / 2 - 1; s >= 0; --s)
        if (row(s, key)) return key;
#pragma unroll
      for (int s = P - 1; s >= P / 2; --s)
        if (row(s, key)) return key;
    }
    return key;
  };
  // the smallest of the waves' keys
  auto group_min_keys = [&](int kx, int ky, int& ox, int& oy) {
    __syncthreads();
    if (lane == 0) {
      ir[w] = kx;
      ir[NW + w] = ky;
    }
    __syncthreads();
    int ax = ir[0], ay = ir[NW];
#pragma unroll
    for (int u = 1; u < NW; ++u) {
      ax = min(ax, ir[u]);
      ay = min(ay, ir[NW + u]);
    }
    ox = ax;
    oy = ay;
  };
  float mx = -INFINITY, my = -INFINITY;
#pragma unroll
  for (int s = 0; s < P; ++s) {
    mx = fmaxf(mx, z[s].x);
    my = fmaxf(my, z[s].y);
  }
  group_max(mx, my, va, vb);
  group_min_keys(find(va, true), find(vb, false), ka, kb);
  const bool miss_a = va > -INFINITY && ka == 0x7fffffff;
  const bool miss_b = vb > -INFINITY && kb == 0x7fffffff;
  if (miss_a || miss_b) {  // uniform: the maximum over the lags alone
    mx = -INFINITY;
    my = -INFINITY;
#pragma unroll
    for (int s = 0; s < P; ++s) {
      const bool ok = is_lag(s);
      mx = ok ? fmaxf(mx, z[s].x) : mx;
      my = ok ? fmaxf(my, z[s].y) : my;
    }
    group_max(mx, my, va, vb);
    group_min_keys(find(va, true), find(vb, false), ka, kb);
  }
}

}  // namespace tcx

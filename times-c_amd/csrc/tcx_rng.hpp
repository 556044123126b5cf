// Counter-based hashing shared by host and device code.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define TCX_HD __host__ __device__
#else
#define TCX_HD
#endif

namespace tcx {

// SplitMix64 finalizer.
TCX_HD inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Power-iteration start vector entry i of cluster `cluster` at iteration
// `iter`: uniform in [0, 1), like np.random.rand (test_timesC.py:31), but a
// pure function of (seed, iter, cluster, i) so device and host agree.
TCX_HD inline double power_v0(uint64_t seed, int iter, int cluster, int i) {
  uint64_t h = mix64(seed + 0x9E3779B97F4A7C15ULL * (uint64_t)(iter + 1));
  h = mix64(h ^ (0xD1B54A32D192ED03ULL * (uint64_t)(cluster + 1)));
  uint64_t z = mix64(h ^ (0x8CB92BA72F3D8DD7ULL * (uint64_t)(i + 1)));
  return (double)(z >> 11) * 0x1.0p-53;
}

}  // namespace tcx

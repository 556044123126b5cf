// Seeded synthetic workload generator and initial-state draws (host code).
//
// The reference's example dataset is not shipped (its README points to the
// UCR archive), so the benchmark inputs come from this generator:
//   prototype_c[t] = sum_{q<3} A_q sin(2 pi f_q t / m + phi_q)
//                    + exp(-(t - mu)^2 / (2 w^2)),      w = m / 20
//     f_q ~ U{1..8}, phi_q ~ U[0, 2 pi), A_q ~ N(0,1), mu ~ U[0, m)
//   x = a * roll(prototype_c, s) + b + N(0, sigma^2)
//     a ~ U[0.5, 2), b ~ N(0,1), s ~ U{-floor(m/10) .. floor(m/10)},
//     roll circular: out[t] = p[(t - s) mod m]
// All draws come from one SplitMix64 stream in a fixed order, so the same
// seed gives the same FP32 matrix on every host.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "timesc.h"
#include "tcx_rng.hpp"

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

struct Stream {
  uint64_t s;
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ULL;
    return tcx::mix64(s);
  }
  double u01() { return (double)(next() >> 11) * 0x1.0p-53; }
  double normal() {
    double u1 = 1.0 - u01();  // (0, 1]
    double u2 = u01();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(kTwoPi * u2);
  }
  // uniform integer in [lo, hi]
  long uint_range(long lo, long hi) {
    long span = hi - lo + 1;
    long v = lo + (long)(u01() * (double)span);
    return v > hi ? hi : v;
  }
};

}  // namespace

extern "C" int tcx_generate(uint64_t seed, long n, int m, int K, double sigma,
                            const double* class_frac, float* X, int* y) {
  if (n < 1 || m < 1 || K < 1 || !X || !y) return TCX_ERR_ARG;
  Stream rng{seed};
  std::vector<double> proto((size_t)K * m);
  for (int c = 0; c < K; ++c) {
    int f[3];
    double phi[3], A[3];
    for (int q = 0; q < 3; ++q) {
      f[q] = (int)rng.uint_range(1, 8);
      phi[q] = kTwoPi * rng.u01();
      A[q] = rng.normal();
    }
    double mu = rng.u01() * m;
    double w = m / 20.0;
    for (int t = 0; t < m; ++t) {
      double v = 0.0;
      for (int q = 0; q < 3; ++q)
        v += A[q] * std::sin(kTwoPi * f[q] * t / m + phi[q]);
      double d = t - mu;
      v += std::exp(-d * d / (2.0 * w * w));
      proto[(size_t)c * m + t] = v;
    }
  }
  // class sizes: balanced, or floor(frac * n) with the remainder to the last
  std::vector<long> size(K, 0);
  long assigned = 0;
  for (int c = 0; c < K; ++c) {
    if (class_frac)
      size[c] = (long)std::floor(class_frac[c] * (double)n);
    else
      size[c] = n / K + (c < n % K ? 1 : 0);
    assigned += size[c];
  }
  size[K - 1] += n - assigned;
  long pos = 0;
  for (int c = 0; c < K; ++c)
    for (long i = 0; i < size[c] && pos < n; ++i) y[pos++] = c;
  for (long i = n - 1; i > 0; --i) {  // Fisher-Yates
    long j = rng.uint_range(0, i);
    int t = y[i];
    y[i] = y[j];
    y[j] = t;
  }
  long smax = m / 10;
  for (long i = 0; i < n; ++i) {
    const double* p = proto.data() + (size_t)y[i] * m;
    double a = 0.5 + 1.5 * rng.u01();
    double b = rng.normal();
    long s = rng.uint_range(-smax, smax);
    float* x = X + (size_t)i * m;
    for (int t = 0; t < m; ++t) {
      long src = ((t - s) % m + m) % m;
      x[t] = (float)(a * p[src] + b + sigma * rng.normal());
    }
  }
  return TCX_OK;
}

extern "C" int tcx_init_labels(uint64_t seed, long n, int k, int* labels) {
  if (n < 1 || k < 1 || !labels) return TCX_ERR_ARG;
  Stream rng{seed ^ 0x5851F42D4C957F2DULL};
  for (long i = 0; i < n; ++i) labels[i] = (int)(rng.next() % (uint64_t)k);
  return TCX_OK;
}

extern "C" int tcx_init_rows(uint64_t seed, long n, int k, int* rows) {
  if (n < 1 || k < 1 || k > n || !rows) return TCX_ERR_ARG;
  Stream rng{seed ^ 0x14057B7EF767814FULL};
  for (int j = 0; j < k; ++j) {
    for (;;) {
      long r = rng.uint_range(0, n - 1);
      bool dup = false;
      for (int q = 0; q < j; ++q) dup = dup || rows[q] == r;
      if (!dup) {
        rows[j] = (int)r;
        break;
      }
    }
  }
  return TCX_OK;
}

extern "C" void tcx_power_v0(uint64_t seed, int iter, int cluster, int m,
                             double* v0) {
  for (int i = 0; i < m; ++i) v0[i] = tcx::power_v0(seed, iter, cluster, i);
}

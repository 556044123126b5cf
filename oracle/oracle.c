/*
 * oracle.c -- FP64 CPU restatement of the reference's CPU k-Shape path,
 * Code/test/test_timesC.py.  TEST INFRASTRUCTURE ONLY (see oracle.h): the
 * product library never links or calls it.
 *
 * What pins each function (tests/test_oracle.py):
 *   orc_fft            numpy.fft.fft / ifft on random inputs, all F in 2..16384
 *   orc_sum_pairwise   bit-exact against numpy's float64 sum
 *   orc_zscore         closed form mean/std (numpy), constant rows -> 0
 *   orc_roll_zeropad   closed form out[t] = a[t-s], |s| >= m -> zeros
 *   orc_ncc            O(m^2) direct correlation; ncc(x,x)[m-1] = 1; |ncc| <= 1;
 *                      zero norm -> 0
 *   orc_align_shift    known answer: member = roll_zeropad(c, d) gives +d (CPU
 *                      convention) and -d (k-Shape convention)
 *   orc_assign         brute force argmin/argmax over the direct correlation
 *   orc_extract_shape  zscore_ddof1(signfix(top eigenvector of S)) via
 *                      numpy.linalg.eigh; one-member cluster -> zscore_ddof1(member)
 *   orc_top_eig        numpy.linalg.eigh
 *   orc_refine/kshape  composition of the above, plus a line-by-line numpy
 *                      restatement of _kshape on small inputs and the
 *                      known answers "true labels are a fixed point" and
 *                      "noise-free prototypes converge to the ground truth"
 *
 * Numerics follow NumPy's float64 semantics, including NaN propagation
 * (0/0 -> NaN, nan_to_num -> 0).  Summation orders differ from NumPy's
 * pairwise sums and BLAS, i.e. results agree to ~1e-15 relative.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_D 3.14159265358979323846264338327950288

/* ------------------------------------------------------------------ FFT */

/* Twiddle tables cos/sin(2*pi*q/F), q < F/2, one per power of two.  Built
 * on first use under a critical section; callers build them before any
 * parallel region so the table pointers are published once. */
static double* g_tw_cos[32];
static double* g_tw_sin[32];

static int ilog2i(int F) {
  int p = 0;
  while ((1 << p) < F) ++p;
  return p;
}

static void tw_get(int F, const double** c, const double** s) {
  int p = ilog2i(F);
  if (!g_tw_cos[p]) {
#pragma omp critical(orc_twiddle)
    {
      if (!g_tw_cos[p]) {
        int h = F / 2 > 0 ? F / 2 : 1;
        double* cc = (double*)malloc(sizeof(double) * h);
        double* ss = (double*)malloc(sizeof(double) * h);
        for (int q = 0; q < h; ++q) {
          cc[q] = cos(2.0 * PI_D * (double)q / (double)F);
          ss[q] = sin(2.0 * PI_D * (double)q / (double)F);
        }
        g_tw_sin[p] = ss;
        g_tw_cos[p] = cc;
      }
    }
  }
  *c = g_tw_cos[p];
  *s = g_tw_sin[p];
}

int orc_fft_size(int m) {
  /* (2m-1).bit_length() */
  unsigned v = (unsigned)(2 * m - 1);
  int bl = 0;
  while (v) { ++bl; v >>= 1; }
  return 1 << bl;
}

void orc_fft(double* re, double* im, int F, int sign) {
  if (F <= 1) return;
  const double *tc, *ts;
  tw_get(F, &tc, &ts);
  /* bit reversal */
  for (int i = 1, j = 0; i < F; ++i) {
    int bit = F >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = re[i]; re[i] = re[j]; re[j] = t;
      t = im[i]; im[i] = im[j]; im[j] = t;
    }
  }
  for (int len = 2; len <= F; len <<= 1) {
    int half = len >> 1, step = F / len;
    for (int i = 0; i < F; i += len) {
      for (int j = 0; j < half; ++j) {
        double wr = tc[j * step], wi = (double)sign * ts[j * step];
        double xr = re[i + j + half], xi = im[i + j + half];
        double vr = xr * wr - xi * wi, vi = xr * wi + xi * wr;
        double ur = re[i + j], ui = im[i + j];
        re[i + j] = ur + vr; im[i + j] = ui + vi;
        re[i + j + half] = ur - vr; im[i + j + half] = ui - vi;
      }
    }
  }
}

/* ------------------------------------------------------------ helpers */

static double nan_to_num(double v) {
  if (isnan(v)) return 0.0;
  if (isinf(v)) return v > 0 ? DBL_MAX : -DBL_MAX;
  return v;
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

/* zscore (test_timesC.py:42-52): (a - mean) / std_ddof, then nan_to_num. */
void orc_zscore(double* x, long rows, int m, int ddof) {
  for (long r = 0; r < rows; ++r) {
    double* a = x + r * (long)m;
    double s = 0.0;
    for (int t = 0; t < m; ++t) s += a[t];
    double mean = s / (double)m;
    double ss = 0.0;
    for (int t = 0; t < m; ++t) { double d = a[t] - mean; ss += d * d; }
    double sd = sqrt(ss / (double)(m - ddof));
    for (int t = 0; t < m; ++t) a[t] = nan_to_num((a[t] - mean) / sd);
  }
}

/* roll_zeropad (test_timesC.py:55-77): out[t] = a[t - shift], zero fill;
 * |shift| >= m gives all zeros. */
void orc_roll_zeropad(const double* a, int m, int shift, double* out) {
  for (int t = 0; t < m; ++t) {
    long src = (long)t - shift;
    out[t] = (src >= 0 && src < m) ? a[src] : 0.0;
  }
}

/* NumPy's float64 add.reduce of a contiguous array (pairwise summation,
 * blocks of 8 accumulators up to 128 elements, halves above). */
double orc_sum_pairwise(const double* a, long n) {
  if (n < 8) {
    double r = 0.0;
    for (long i = 0; i < n; ++i) r += a[i];
    return r;
  }
  if (n <= 128) {
    double r[8];
    for (int j = 0; j < 8; ++j) r[j] = a[j];
    long i = 8;
    for (; i < n - (n % 8); i += 8)
      for (int j = 0; j < 8; ++j) r[j] += a[i + j];
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res += a[i];
    return res;
  }
  long n2 = n / 2;
  n2 -= n2 % 8;
  return orc_sum_pairwise(a, n2) + orc_sum_pairwise(a + n2, n - n2);
}

static double norm2(const double* a, int m) {
  double s = 0.0;
  for (int t = 0; t < m; ++t) s += a[t] * a[t];
  return sqrt(s);
}

/* Lag-ordered index idx in [0, 2m-2] -> circular position in the length-F
 * correlation (the concatenation cc[-(m-1):] ++ cc[:m], test_timesC.py:90). */
static inline int lag_pos(int idx, int m, int F) {
  int lag = idx - (m - 1);
  return lag < 0 ? F + lag : lag;
}

/* _ncc_c (test_timesC.py:80-91) */
void orc_ncc(const double* x, const double* y, int m, double* out) {
  int F = orc_fft_size(m);
  double* buf = (double*)calloc((size_t)4 * F, sizeof(double));
  double *xr = buf, *xi = buf + F, *yr = buf + 2 * F, *yi = buf + 3 * F;
  memcpy(xr, x, sizeof(double) * m);
  memcpy(yr, y, sizeof(double) * m);
  orc_fft(xr, xi, F, -1);
  orc_fft(yr, yi, F, -1);
  for (int f = 0; f < F; ++f) {
    double ar = xr[f], ai = xi[f], br = yr[f], bi = -yi[f];
    xr[f] = ar * br - ai * bi;
    xi[f] = ar * bi + ai * br;
  }
  orc_fft(xr, xi, F, +1);
  double den = norm2(x, m) * norm2(y, m);
  if (den == 0.0) den = INFINITY;
  for (int idx = 0; idx < 2 * m - 1; ++idx)
    out[idx] = (xr[lag_pos(idx, m, F)] / (double)F) / den;
  free(buf);
}

static int first_argmax(const double* v, int len) {
  int best = 0;
  for (int i = 1; i < len; ++i)
    if (v[i] > v[best]) best = i;
  return best;
}

int orc_align_shift(const double* member, const double* centroid, int m,
                    int convention) {
  double* cc = (double*)malloc(sizeof(double) * (2 * m - 1));
  if (convention == ORC_CONV_KSHAPE)
    orc_ncc(centroid, member, m, cc); /* SBD.cuh:78 NCC_3D(center, member) */
  else
    orc_ncc(member, centroid, m, cc); /* test_timesC.py:129 _ncc_c_2dim(x, c) */
  int idx = first_argmax(cc, 2 * m - 1);
  free(cc);
  return idx + 1 - m; /* test_timesC.py:132, SBD.cuh:61 */
}

/* ---------------------------------------------------------- assignment */

/* distances = (1 - _ncc_c_3dim(x, centroids).max(axis=2)).T; argmin(1)
 * (test_timesC.py:105-114, 258-260).  Two centroids share one inverse FFT
 * (ifft(P_j + i P_j+1) = cc_j + i cc_j+1 for Hermitian P), which is the same
 * arithmetic up to rounding. */
void orc_assign(const double* X, long n, int m, const double* C, int k,
                int* labels, double* dists, int* shifts, double* dmat,
                int nthreads) {
  int F = orc_fft_size(m);
  const double *tc, *ts;
  tw_get(F, &tc, &ts);
  /* centroid spectra and norms */
  double* Cr = (double*)calloc((size_t)k * F, sizeof(double));
  double* Ci = (double*)calloc((size_t)k * F, sizeof(double));
  double* cn = (double*)malloc(sizeof(double) * k);
  for (int j = 0; j < k; ++j) {
    memcpy(Cr + (size_t)j * F, C + (size_t)j * m, sizeof(double) * m);
    orc_fft(Cr + (size_t)j * F, Ci + (size_t)j * F, F, -1);
    cn[j] = norm2(C + (size_t)j * m, m);
  }
  set_threads(nthreads);
#pragma omp parallel
  {
    double* xr = (double*)malloc(sizeof(double) * F);
    double* xi = (double*)malloc(sizeof(double) * F);
    double* zr = (double*)malloc(sizeof(double) * F);
    double* zi = (double*)malloc(sizeof(double) * F);
    double* dist = (double*)malloc(sizeof(double) * k);
    int* amax = (int*)malloc(sizeof(int) * k);
#pragma omp for schedule(dynamic, 16)
    for (long i = 0; i < n; ++i) {
      const double* x = X + i * (long)m;
      memset(xr, 0, sizeof(double) * F);
      memset(xi, 0, sizeof(double) * F);
      memcpy(xr, x, sizeof(double) * m);
      orc_fft(xr, xi, F, -1);
      double xn = norm2(x, m);
      for (int j = 0; j < k; j += 2) {
        int two = (j + 1 < k);
        const double* ar = Cr + (size_t)j * F;
        const double* ai = Ci + (size_t)j * F;
        const double* br = two ? Cr + (size_t)(j + 1) * F : NULL;
        const double* bi = two ? Ci + (size_t)(j + 1) * F : NULL;
        for (int f = 0; f < F; ++f) {
          /* P_j = X conj(C_j) */
          double pr = xr[f] * ar[f] + xi[f] * ai[f];
          double pi = xi[f] * ar[f] - xr[f] * ai[f];
          double qr = 0.0, qi = 0.0;
          if (two) {
            qr = xr[f] * br[f] + xi[f] * bi[f];
            qi = xi[f] * br[f] - xr[f] * bi[f];
          }
          /* Z = P + i Q */
          zr[f] = pr - qi;
          zi[f] = pi + qr;
        }
        orc_fft(zr, zi, F, +1);
        for (int u = 0; u < 1 + two; ++u) {
          const double* cc = u == 0 ? zr : zi;
          double den = xn * cn[j + u];
          if (den == 0.0) den = INFINITY;
          int best = 0;
          double bv = (cc[lag_pos(0, m, F)] / (double)F) / den;
          for (int idx = 1; idx < 2 * m - 1; ++idx) {
            double v = (cc[lag_pos(idx, m, F)] / (double)F) / den;
            if (v > bv) { bv = v; best = idx; }
          }
          dist[j + u] = 1.0 - bv;
          amax[j + u] = best;
        }
      }
      int lab = 0;
      for (int j = 1; j < k; ++j)
        if (dist[j] < dist[lab]) lab = j;
      labels[i] = lab;
      if (dists) dists[i] = dist[lab];
      if (shifts) shifts[i] = amax[lab] + 1 - m;
      if (dmat)
        for (int j = 0; j < k; ++j) dmat[i * (long)k + j] = dist[j];
    }
    free(xr); free(xi); free(zr); free(zi); free(dist); free(amax);
  }
  free(Cr); free(Ci); free(cn);
}

/* ------------------------------------------------------ power-iteration v0 */

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void orc_power_v0(uint64_t seed, int iter, int cluster, int m, double* v0) {
  uint64_t h = mix64(seed + 0x9E3779B97F4A7C15ULL * (uint64_t)(iter + 1));
  h = mix64(h ^ (0xD1B54A32D192ED03ULL * (uint64_t)(cluster + 1)));
  for (int i = 0; i < m; ++i) {
    uint64_t z = mix64(h ^ (0x8CB92BA72F3D8DD7ULL * (uint64_t)(i + 1)));
    v0[i] = (double)(z >> 11) * 0x1.0p-53;
  }
}

/* ------------------------------------------------- top eigenpair (eigh) */

/* LU with partial pivoting of the tridiagonal T - sigma I and one solve. */
static void tri_solve(int n, const double* d, const double* e, double sigma,
                      double tiny, double* b) {
  double* dl = (double*)malloc(sizeof(double) * n);
  double* dd = (double*)malloc(sizeof(double) * n);
  double* du = (double*)malloc(sizeof(double) * n);
  double* du2 = (double*)calloc(n, sizeof(double));
  int* piv = (int*)malloc(sizeof(int) * n);
  for (int i = 0; i < n; ++i) {
    dd[i] = d[i] - sigma;
    if (i + 1 < n) { dl[i] = e[i + 1]; du[i] = e[i + 1]; }
  }
  for (int i = 0; i + 1 < n; ++i) {
    if (fabs(dd[i]) >= fabs(dl[i])) {
      if (dd[i] == 0.0) dd[i] = tiny;
      double f = dl[i] / dd[i];
      dl[i] = f;
      dd[i + 1] -= f * du[i];
      du2[i] = 0.0;
      piv[i] = i;
    } else {
      double f = dd[i] / dl[i];
      dd[i] = dl[i];
      dl[i] = f;
      double t = du[i];
      du[i] = dd[i + 1];
      dd[i + 1] = t - f * dd[i + 1];
      if (i + 2 < n) {
        du2[i] = du[i + 1];
        du[i + 1] = -f * du[i + 1];
      }
      piv[i] = i + 1;
    }
  }
  if (dd[n - 1] == 0.0) dd[n - 1] = tiny;
  for (int i = 0; i + 1 < n; ++i) {
    if (piv[i] == i) {
      b[i + 1] -= dl[i] * b[i];
    } else {
      double t = b[i];
      b[i] = b[i + 1];
      b[i + 1] = t - dl[i] * b[i];
    }
  }
  b[n - 1] /= dd[n - 1];
  if (n >= 2) b[n - 2] = (b[n - 2] - du[n - 2] * b[n - 1]) / dd[n - 2];
  for (int i = n - 3; i >= 0; --i)
    b[i] = (b[i] - du[i] * b[i + 1] - du2[i] * b[i + 2]) / dd[i];
  free(dl); free(dd); free(du); free(du2); free(piv);
}

/* number of eigenvalues of the tridiagonal (d, e) that are < x */
static int sturm_count(int n, const double* d, const double* e, double x,
                       double tiny) {
  int cnt = 0;
  double q = d[0] - x;
  if (q < 0) ++cnt;
  for (int i = 1; i < n; ++i) {
    if (q == 0.0) q = tiny;
    q = d[i] - x - e[i] * e[i] / q;
    if (q < 0) ++cnt;
  }
  return cnt;
}

int orc_top_eig(const double* Min, int n, double* lambda, double* vec) {
  if (n == 1) { *lambda = Min[0]; vec[0] = 1.0; return 0; }
  double* A = (double*)malloc(sizeof(double) * (size_t)n * n);
  memcpy(A, Min, sizeof(double) * (size_t)n * n);
  double* U = (double*)calloc((size_t)n * n, sizeof(double)); /* Householder u_k in row k */
  double* d = (double*)malloc(sizeof(double) * n);
  double* e = (double*)calloc(n, sizeof(double));
  double* p = (double*)malloc(sizeof(double) * n);
  /* Householder tridiagonalization: T = Q^T A Q, Q = H_0 ... H_{n-3} */
  for (int k = 0; k + 2 < n; ++k) {
    int len = n - k - 1;
    double* x = p; /* reuse as scratch */
    double s = 0.0;
    for (int i = 0; i < len; ++i) { x[i] = A[(size_t)(k + 1 + i) * n + k]; s += x[i] * x[i]; }
    double nx = sqrt(s);
    double* u = U + (size_t)k * n;
    if (nx == 0.0) { e[k + 1] = 0.0; continue; }
    double alpha = x[0] > 0 ? -nx : nx;
    for (int i = 0; i < len; ++i) u[k + 1 + i] = x[i];
    u[k + 1] -= alpha;
    double nu = 0.0;
    for (int i = k + 1; i < n; ++i) nu += u[i] * u[i];
    nu = sqrt(nu);
    if (nu == 0.0) { e[k + 1] = A[(size_t)(k + 1) * n + k]; continue; }
    for (int i = k + 1; i < n; ++i) u[i] /= nu;
    /* p = B u, K = u^T p, q = p - K u, B -= 2 (u q^T + q u^T) */
    double K = 0.0;
    for (int i = k + 1; i < n; ++i) {
      double acc = 0.0;
      const double* Ai = A + (size_t)i * n;
      for (int j = k + 1; j < n; ++j) acc += Ai[j] * u[j];
      p[i] = acc;
    }
    for (int i = k + 1; i < n; ++i) K += u[i] * p[i];
    for (int i = k + 1; i < n; ++i) p[i] -= K * u[i];
    for (int i = k + 1; i < n; ++i) {
      double* Ai = A + (size_t)i * n;
      for (int j = k + 1; j < n; ++j) Ai[j] -= 2.0 * (u[i] * p[j] + p[i] * u[j]);
    }
    e[k + 1] = alpha;
    for (int i = k + 1; i < n; ++i) { A[(size_t)i * n + k] = 0.0; A[(size_t)k * n + i] = 0.0; }
    A[(size_t)(k + 1) * n + k] = alpha;
    A[(size_t)k * n + k + 1] = alpha;
  }
  for (int i = 0; i < n; ++i) d[i] = A[(size_t)i * n + i];
  e[n - 1] = A[(size_t)(n - 1) * n + n - 2];
  if (n >= 3) { /* e[i] = T[i][i-1] for i < n-1 was set above */ }
  /* Gershgorin bounds, bisection for the largest eigenvalue */
  double lo = DBL_MAX, hi = -DBL_MAX, anorm = 0.0;
  for (int i = 0; i < n; ++i) {
    double r = (i > 0 ? fabs(e[i]) : 0.0) + (i + 1 < n ? fabs(e[i + 1]) : 0.0);
    if (d[i] - r < lo) lo = d[i] - r;
    if (d[i] + r > hi) hi = d[i] + r;
    if (fabs(d[i]) + r > anorm) anorm = fabs(d[i]) + r;
  }
  double tiny = (anorm > 0 ? anorm : 1.0) * DBL_EPSILON * 1e-3;
  for (int it = 0; it < 200; ++it) {
    double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (sturm_count(n, d, e, mid, tiny) >= n) hi = mid; /* all < mid */
    else lo = mid;
  }
  double lam = hi;
  /* inverse iteration on T */
  double* z = (double*)malloc(sizeof(double) * n);
  for (int i = 0; i < n; ++i) z[i] = 1.0 + 0.1 * (double)((i * 7919) % 13) / 13.0;
  double shift = lam + (anorm > 0 ? anorm : 1.0) * DBL_EPSILON * 4.0;
  for (int it = 0; it < 4; ++it) {
    tri_solve(n, d, e, shift, tiny, z);
    double nz = 0.0;
    for (int i = 0; i < n; ++i) nz += z[i] * z[i];
    nz = sqrt(nz);
    if (!(nz > 0) || !isfinite(nz)) { free(A); free(U); free(d); free(e); free(p); free(z); return 1; }
    for (int i = 0; i < n; ++i) z[i] /= nz;
  }
  /* back-transform: vec = H_0 (H_1 (... H_{n-3} z)) */
  for (int k = n - 3; k >= 0; --k) {
    const double* u = U + (size_t)k * n;
    double dot = 0.0;
    for (int i = k + 1; i < n; ++i) dot += u[i] * z[i];
    for (int i = k + 1; i < n; ++i) z[i] -= 2.0 * dot * u[i];
  }
  memcpy(vec, z, sizeof(double) * n);
  *lambda = lam;
  free(A); free(U); free(d); free(e); free(p); free(z);
  return 0;
}

/* ---------------------------------------------------- shape extraction */

/* S = y^T y (upper triangle mirrored), y row-major count x m. */
static void syrk_upper(const double* y, long count, int m, double* S) {
  double* yt = (double*)malloc(sizeof(double) * (size_t)m * count);
  for (long p = 0; p < count; ++p)
    for (int i = 0; i < m; ++i) yt[(size_t)i * count + p] = y[p * (long)m + i];
  const int TB = 32;
  int nb = (m + TB - 1) / TB;
#pragma omp parallel for schedule(dynamic, 1)
  for (int bb = 0; bb < nb * nb; ++bb) {
    int bi = bb / nb, bj = bb % nb;
    if (bj < bi) continue;
    for (int i = bi * TB; i < m && i < (bi + 1) * TB; ++i)
      for (int j = bj * TB; j < m && j < (bj + 1) * TB; ++j) {
        if (j < i) continue;
        const double* a = yt + (size_t)i * count;
        const double* b = yt + (size_t)j * count;
        double acc = 0.0;
        for (long p = 0; p < count; ++p) acc += a[p] * b[p];
        S[(size_t)i * m + j] = acc;
        S[(size_t)j * m + i] = acc;
      }
  }
  free(yt);
}

static void matvec(const double* M, int m, const double* v, double* w) {
#pragma omp parallel for schedule(static) if (m >= 256)
  for (int i = 0; i < m; ++i) {
    const double* Mi = M + (size_t)i * m;
    double acc = 0.0;
    for (int j = 0; j < m; ++j) acc += Mi[j] * v[j];
    w[i] = acc;
  }
}

static void sign_fix(const double* a0, double* c, int m, int* flipped) {
  /* test_timesC.py:163-166 / 187-191 */
  double s1 = 0.0, s2 = 0.0;
  for (int t = 0; t < m; ++t) {
    double d1 = a0[t] - c[t], d2 = a0[t] + c[t];
    s1 += d1 * d1;
    s2 += d2 * d2;
  }
  if (sqrt(s1) >= sqrt(s2)) {
    for (int t = 0; t < m; ++t) c[t] = -c[t];
    if (flipped) *flipped = 1;
  }
}

int orc_extract_shape(const double* A, long count, int m, const double* v0,
                      int exact_steps, double* centroid, orc_shape_info* info,
                      int cal_select) {
  orc_shape_info inf;
  memset(&inf, 0, sizeof(inf));
  if (m < 1) return 1;
  if (count == 0) { /* test_timesC.py:139-140 */
    memset(centroid, 0, sizeof(double) * m);
    inf.empty = 1;
    if (info) *info = inf;
    return 0;
  }
  double* y = (double*)malloc(sizeof(double) * (size_t)count * m);
  memcpy(y, A, sizeof(double) * (size_t)count * m);
  orc_zscore(y, count, m, 1); /* :142 */
  double* M = (double*)malloc(sizeof(double) * (size_t)m * m);
  syrk_upper(y, count, m, M); /* :143 */
  free(y);
  /* :145-149  M = P S P, P = I - 11^T/m, in its exact rank-1 form
   * S - (r 1^T + 1 r^T)/m + (sigma/m^2) 1 1^T with r = S 1, sigma = 1^T S 1 */
  double* r = (double*)malloc(sizeof(double) * m);
  double sigma = 0.0;
  for (int i = 0; i < m; ++i) {
    double acc = 0.0;
    for (int j = 0; j < m; ++j) acc += M[(size_t)i * m + j];
    r[i] = acc;
    sigma += acc;
  }
  double inv_m = 1.0 / (double)m;
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j)
      M[(size_t)i * m + j] += -(r[i] + r[j]) * inv_m + sigma * inv_m * inv_m;
  free(r);
  /* :152-157 power_iteration(M, min(m, 1000)) (:30-38); cal_select = 1
   * skips it and takes the eigh branch directly (the CUDA path's
   * get_class_centers_a, shapeExtract.cuh:366-468; the commented-out block
   * test_timesC.py:194-201) */
  int iters = cal_select == 1 ? 0 : (m > 1000 ? 1000 : m);
  double* v = (double*)malloc(sizeof(double) * m);
  double* w = (double*)malloc(sizeof(double) * m);
  double nv = 0.0;
  for (int i = 0; i < m; ++i) nv += v0[i] * v0[i];
  nv = sqrt(nv);
  for (int i = 0; i < m; ++i) v[i] = v0[i] / nv;
  int steps = 0;
  for (int t = 0; t < iters; ++t) {
    matvec(M, m, v, w);
    double nw = 0.0;
    for (int i = 0; i < m; ++i) nw += w[i] * w[i];
    nw = sqrt(nw);
    double dmax = 0.0;
    for (int i = 0; i < m; ++i) {
      double nvv = w[i] / nw;
      double dd = fabs(nvv - v[i]);
      if (dd > dmax || isnan(dd)) dmax = isnan(dd) ? INFINITY : dd;
      v[i] = nvv;
    }
    ++steps;
    if (!exact_steps && dmax <= 1e-14) break;
  }
  matvec(M, m, v, w);
  double num = 0.0, den = 0.0;
  for (int i = 0; i < m; ++i) { num += w[i] * v[i]; den += v[i] * v[i]; }
  double lambda = num / den;
  double* c = v;
  const double* a0 = A; /* a[0]: first aligned member, before z-norm */
  int fallback = 0;
  double sumdiff = 0.0;
  if (cal_select == 1) {
    fallback = 1;
  } else if (lambda == 0.0) {
    /* :159-161 flag=False leaves vecright unbound in the reference (an
     * UnboundLocalError); treated here as the eigh fallback. */
    fallback = 1;
  } else {
    sign_fix(a0, c, m, &inf.flipped); /* :163-166 */
    double* vr = (double*)malloc(sizeof(double) * m);
    double* vl = w;
    for (int i = 0; i < m; ++i) vr[i] = lambda * c[i]; /* :168 */
    matvec(M, m, c, vl);                                /* :169 */
    int pos = 0;                                        /* :170-174 */
    for (int i = 0; i < m; ++i)
      if (vl[i] != 0.0 && vr[i] != 0.0) { pos = i; break; }
    if (vl[pos] < 0 && vr[pos] > 0) { /* :175-177 */
      lambda = -lambda;
      for (int i = 0; i < m; ++i) vr[i] = -vl[i];
    }
    double sd = 0.0; /* :178-179, Python's sequential sum */
    for (int i = 0; i < m; ++i) sd += vr[i] - vl[i];
    sumdiff = fabs(sd);
    free(vr);
    if (lambda <= 0 || sumdiff > (double)m * 0.01) fallback = 1; /* :180-183 */
  }
  if (fallback) { /* :184-192 */
    double lam;
    if (orc_top_eig(M, m, &lam, c) != 0) {
      free(M); free(v); free(w);
      return 2;
    }
    inf.flipped = 0;
    sign_fix(a0, c, m, &inf.flipped);
    lambda = lam;
  }
  memcpy(centroid, c, sizeof(double) * m);
  orc_zscore(centroid, 1, m, 1); /* :204 */
  inf.lambda = lambda;
  inf.steps = steps;
  inf.fallback = fallback;
  inf.sumdiff = sumdiff;
  if (info) *info = inf;
  free(M); free(v); free(w);
  return 0;
}

/* ----------------------------------------------------------- refinement */

void orc_refine(const double* X, long n, int m, int k, const int* labels,
                double* C, int convention, uint64_t seed, int iter,
                int exact_steps, int* counts, orc_shape_info* infos,
                int nthreads, int skip_rule, int cal_select) {
  orc_refine_shifts(X, n, m, k, labels, C, convention, seed, iter, exact_steps,
                    counts, infos, nthreads, skip_rule, cal_select, NULL);
}

void orc_refine_shifts(const double* X, long n, int m, int k, const int* labels,
                       double* C, int convention, uint64_t seed, int iter,
                       int exact_steps, int* counts, orc_shape_info* infos,
                       int nthreads, int skip_rule, int cal_select,
                       const int* shifts) {
  set_threads(nthreads);
  int F = orc_fft_size(m);
  const double *tc, *ts;
  tw_get(F, &tc, &ts);
  /* group rows by label in index order (test_timesC.py:228-243) */
  long* cnt = (long*)calloc(k, sizeof(long));
  for (long i = 0; i < n; ++i) cnt[labels[i]]++;
  long* off = (long*)calloc((size_t)k + 1, sizeof(long));
  for (int j = 0; j < k; ++j) off[j + 1] = off[j] + cnt[j];
  long* members = (long*)malloc(sizeof(long) * (n > 0 ? n : 1));
  long* fill = (long*)calloc(k, sizeof(long));
  for (long i = 0; i < n; ++i) {
    int j = labels[i];
    members[off[j] + fill[j]++] = i;
  }
  double* newC = (double*)malloc(sizeof(double) * (size_t)k * m);
  double* v0 = (double*)malloc(sizeof(double) * m);
  for (int j = 0; j < k; ++j) {
    long c = cnt[j];
    double* A = (double*)malloc(sizeof(double) * (size_t)(c > 0 ? c : 1) * m);
    for (long p = 0; p < c; ++p)
      memcpy(A + p * (long)m, X + members[off[j] + p] * (long)m, sizeof(double) * m);
    const double* cj = C + (size_t)j * m;
    /* test_timesC.py:128 skips alignment when cur_center.sum() == 0.
     * ORC_SKIP_ZERO: only an all-zero centroid skips (initial / empty
     * cluster).  ORC_SKIP_NUMPY_SUM: the literal rule, NumPy's pairwise
     * float64 sum compared with 0.0. */
    int align = 0;
    if (skip_rule == ORC_SKIP_NUMPY_SUM) {
      align = orc_sum_pairwise(cj, m) != 0.0;
    } else {
      for (int t = 0; t < m; ++t) align |= cj[t] != 0.0;
    }
    if (align && c > 0) {
#pragma omp parallel
      {
        double* tmp = (double*)malloc(sizeof(double) * m);
#pragma omp for schedule(dynamic, 8)
        for (long p = 0; p < c; ++p) {
          double* a = A + p * (long)m;
          int s = shifts ? shifts[members[off[j] + p]]
                         : orc_align_shift(a, cj, m, convention);
          orc_roll_zeropad(a, m, s, tmp);
          memcpy(a, tmp, sizeof(double) * m);
        }
        free(tmp);
      }
    }
    orc_power_v0(seed, iter, j, m, v0);
    orc_shape_info info;
    orc_extract_shape(A, c, m, v0, exact_steps, newC + (size_t)j * m, &info, cal_select);
    if (infos) infos[j] = info;
    if (counts) counts[j] = (int)c;
    free(A);
  }
  memcpy(C, newC, sizeof(double) * (size_t)k * m);
  free(newC); free(v0); free(cnt); free(off); free(members); free(fill);
}

/* ----------------------------------------------------------- the loop */

/* sum_class == 1 || sum_class <= k/2 (kshape.cuh:248, 454) */
static int orc_collapsed(const int* counts, int k) {
  int nonempty = 0;
  for (int j = 0; j < k; ++j) nonempty += counts[j] > 0;
  return nonempty == 1 || nonempty <= k / 2;
}

int orc_kshape(const double* X, long n, int m, int k, int init_mode,
               const int* init_labels, const int* init_rows, int max_iter,
               int convention, uint64_t seed, int exact_steps,
               int* labels_hist, double* centroids, int nthreads, int skip_rule,
               int cal_select, int stop_on_collapse, int* collapsed) {
  int* lab = (int*)malloc(sizeof(int) * n);
  int* old = (int*)malloc(sizeof(int) * n);
  int* counts = (int*)malloc(sizeof(int) * k);
  double* C = (double*)calloc((size_t)k * m, sizeof(double));
  int it = 0, stop = 0;
  if (collapsed) *collapsed = 0;
  if (init_mode == 0) {
    /* test_timesC.py:217-264: refine, assign, stop when labels repeat */
    memcpy(lab, init_labels, sizeof(int) * n);
    for (it = 0; it < max_iter; ++it) {
      memcpy(old, lab, sizeof(int) * n);
      orc_refine(X, n, m, k, lab, C, convention, seed, it, exact_steps, counts,
                 NULL, nthreads, skip_rule, cal_select);
      /* the GPU path's break after extract_shape (kshape.cuh:248-249): no
       * assignment, the labels stay those of the previous one */
      if (stop_on_collapse && orc_collapsed(counts, k)) { stop = 1; break; }
      orc_assign(X, n, m, C, k, lab, NULL, NULL, NULL, nthreads);
      memcpy(labels_hist + (size_t)it * n, lab, sizeof(int) * n);
      if (memcmp(old, lab, sizeof(int) * n) == 0) { ++it; break; }
    }
  } else {
    /* centre initialisation, loop order of k_shape_tsl (kshape.cuh:376-461):
     * assign, stop test, refine (then the collapse break, :454-455) */
    for (int j = 0; j < k; ++j)
      memcpy(C + (size_t)j * m, X + (long)init_rows[j] * m, sizeof(double) * m);
    for (it = 0; it < max_iter; ++it) {
      orc_assign(X, n, m, C, k, lab, NULL, NULL, NULL, nthreads);
      memcpy(labels_hist + (size_t)it * n, lab, sizeof(int) * n);
      if (it > 0 && memcmp(old, lab, sizeof(int) * n) == 0) { ++it; break; }
      memcpy(old, lab, sizeof(int) * n);
      orc_refine(X, n, m, k, lab, C, convention, seed, it, exact_steps, counts,
                 NULL, nthreads, skip_rule, cal_select);
      if (stop_on_collapse && orc_collapsed(counts, k)) { stop = 1; ++it; break; }
    }
  }
  if (collapsed) *collapsed = stop;
  if (centroids) memcpy(centroids, C, sizeof(double) * (size_t)k * m);
  free(lab); free(old); free(counts); free(C);
  return it;
}

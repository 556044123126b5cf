/*
 * oracle.h -- FP64 CPU restatement of the reference's CPU k-Shape path
 * (Code/test/test_timesC.py).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is used by tests/ (the parity reference), by
 * __graft_entry__.smoke() (to check one small run) and by bench.py's
 * cpu_baseline leg (timed on the host, never as the measured GPU value).
 *
 * Every function cites the reference line range it restates.  Pinning (what
 * checks each function against something other than itself) is listed in the
 * header comment of oracle.c.
 */
#ifndef TIMESC_ORACLE_H
#define TIMESC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Alignment sign conventions.
 * ORC_CONV_CPU:    cc(member, centroid), shift = argmax + 1 - m
 *                  (test_timesC.py:129-132) -- the parity default.
 * ORC_CONV_KSHAPE: cc(centroid, member), shift = argmax + 1 - m
 *                  (SBD.cuh:61,78) -- classic k-Shape alignment. */
#define ORC_CONV_CPU 0
#define ORC_CONV_KSHAPE 1

/* Alignment-skip rules for `cur_center.sum() != 0` (test_timesC.py:128).
 * ORC_SKIP_ZERO (default): skip only for an all-zero centroid.
 * ORC_SKIP_NUMPY_SUM: literal -- NumPy's pairwise float64 sum == 0.0, which
 * also fires for non-zero z-scored centroids whose rounded sum is exactly 0
 * (about 6% of centroids at m=512, 14% at m=32). */
#define ORC_SKIP_ZERO 0
#define ORC_SKIP_NUMPY_SUM 1

/* NumPy's float64 sum of a contiguous array (pairwise summation). */
double orc_sum_pairwise(const double* a, long n);

typedef struct {
  double lambda;      /* Rayleigh quotient returned by the power iteration   */
  int steps;          /* power-iteration steps actually run                  */
  int fallback;       /* 1 if the eigh branch (test_timesC.py:183-192) ran   */
  int flipped;        /* 1 if the sign fix negated the vector                */
  int empty;          /* 1 if the cluster was empty (zero centroid)          */
  double sumdiff;     /* |sum(lambda*c - M*c)| of test_timesC.py:178-179     */
} orc_shape_info;

/* 1 << (2m-1).bit_length()  (test_timesC.py:110) */
int orc_fft_size(int m);

/* In-place complex FFT (sign=-1 forward, +1 inverse, inverse unscaled). */
void orc_fft(double* re, double* im, int F, int sign);

/* zscore along rows with nan_to_num (test_timesC.py:42-52). */
void orc_zscore(double* x, long rows, int m, int ddof);

/* roll_zeropad (test_timesC.py:55-77). */
void orc_roll_zeropad(const double* a, int m, int shift, double* out);

/* _ncc_c(x, y): 2m-1 normalized cross-correlation values (test_timesC.py:80-91). */
void orc_ncc(const double* x, const double* y, int m, double* out);

/* Alignment shift of one member against one centroid
 * (test_timesC.py:129-132 for ORC_CONV_CPU). */
int orc_align_shift(const double* member, const double* centroid, int m, int convention);

/* Assignment (test_timesC.py:258-260): dist = 1 - max_lag ncc(x_i, c_j),
 * label = first argmin over j.  shifts[i] = argmax+1-m of ncc(x_i, c_label)
 * (the shift the next refinement applies to x_i).  dmat (n*k, may be NULL)
 * receives the full distance matrix.  nthreads<=0: OpenMP default. */
void orc_assign(const double* X, long n, int m, const double* C, int k,
                int* labels, double* dists, int* shifts, double* dmat,
                int nthreads);

/* Deterministic power-iteration start vector shared with the GPU engine
 * (the reference draws np.random.rand(m), test_timesC.py:31). */
void orc_power_v0(uint64_t seed, int iter, int cluster, int m, double* v0);

/* _extract_shape (test_timesC.py:136-204) on `count` aligned members A
 * (count x m).  v0 is the start vector (m values, not yet normalized).
 * exact_steps != 0 runs exactly min(m,1000) steps; otherwise it stops once a
 * step changes v by <= 1e-14 (infinity norm).  cal_select = 1 takes the
 * top eigenvector (the eigh branch) without power iteration.  Returns 0 on
 * success. */
int orc_extract_shape(const double* A, long count, int m, const double* v0,
                      int exact_steps, double* centroid, orc_shape_info* info,
                      int cal_select);

/* One refinement (test_timesC.py:228-255): group rows by label in index
 * order, align each cluster to C[j] unless C[j] sums to 0, extract shapes.
 * C is read (old centroids) and overwritten (new centroids).
 * counts (k, may be NULL) receives cluster sizes; infos (k, may be NULL). */
void orc_refine(const double* X, long n, int m, int k, const int* labels,
                double* C, int convention, uint64_t seed, int iter,
                int exact_steps, int* counts, orc_shape_info* infos,
                int nthreads, int skip_rule, int cal_select);

/* orc_refine with the alignment shift of every row given (shifts, n values;
 * NULL: computed as _preSBD_extract_shape does).  Lets a test refine
 * from the exact shifts an FP32 engine chose where two lags tie. */
void orc_refine_shifts(const double* X, long n, int m, int k, const int* labels,
                       double* C, int convention, uint64_t seed, int iter,
                       int exact_steps, int* counts, orc_shape_info* infos,
                       int nthreads, int skip_rule, int cal_select,
                       const int* shifts);

/* Full k-Shape loop (test_timesC.py:213-265) on z-normalized X.
 * init_mode 0: init_labels (n) are the starting labels, centroids start at 0.
 * init_mode 1: init_rows (k) are row indices whose (z-normalized) rows are
 *              the starting centroids; loop order assign -> stop test ->
 *              refine (the k_shape_tsl order, kshape.cuh:376-461).
 * labels_hist (max_iter*n) receives the labels after each iteration's
 * assignment; centroids (k*m) the final centroids; returns iterations run
 * (assignments made).  stop_on_collapse != 0 adds the GPU path's break after
 * a refinement leaving one non-empty cluster or at most k/2
 * (kshape.cuh:248-249, 454-455); *collapsed (may be NULL) reports it. */
int orc_kshape(const double* X, long n, int m, int k, int init_mode,
               const int* init_labels, const int* init_rows, int max_iter,
               int convention, uint64_t seed, int exact_steps,
               int* labels_hist, double* centroids, int nthreads,
               int skip_rule, int cal_select, int stop_on_collapse,
               int* collapsed);

/* Top eigenpair of a symmetric m x m matrix (FP64): Householder
 * tridiagonalization, bisection for the largest eigenvalue, inverse
 * iteration, back-transformation.  Stands in for numpy.linalg.eigh's last
 * column (test_timesC.py:185-186).  Returns 0 on success. */
int orc_top_eig(const double* M, int m, double* lambda, double* vec);

#ifdef __cplusplus
}
#endif
#endif

/*
 * timesc.h -- C ABI of the MI355X k-Shape engine (libtimesc_hip.so).
 *
 * The engine replaces the GPU path of rbowS/Times-C.  Each entry point below
 * names the reference interface it stands in for (paths relative to the
 * reference repository).  All entry points:
 *   - take plain pointers and sizes ("d_" = device memory, "h_" = host);
 *   - return an int status, 0 = TCX_OK (the reference prints and calls
 *     exit(-1) instead, check.cuh:7-19; timesc_ref.hpp restores that
 *     behaviour for drop-in callers);
 *   - are synchronous with respect to the host unless stated otherwise and
 *     run on the context's stream;
 *   - follow the reference's CPU path semantics (Code/test/test_timesC.py)
 *     for every result: FFT length F = 1 << (2m-1).bit_length(), first-index
 *     argmax over lags, first-index argmin over clusters, alignment sign of
 *     _preSBD_extract_shape, zscore with nan_to_num.
 * Series are rows of a row-major [n][m] FP32 matrix.
 */
#ifndef TIMESC_H
#define TIMESC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define TCX_OK 0
#define TCX_ERR_ARG 1         /* invalid argument (sizes, k > n, null pointer) */
#define TCX_ERR_HIP 2         /* a HIP runtime call failed                     */
#define TCX_ERR_NOMEM 3       /* device allocation failed                      */
#define TCX_ERR_BUDGET 4      /* request exceeds the context's HBM budget      */
#define TCX_ERR_IO 5          /* file missing, unreadable or malformed         */
#define TCX_ERR_UNSUPPORTED 6 /* option not supported by this build            */
#define TCX_ERR_NUMERIC 7     /* kernel reported a numerical failure           */

/* alignment sign conventions (see tcx_sbd_align) */
#define TCX_CONV_CPU 0    /* cc(member, centroid): test_timesC.py:129-132 (default) */
#define TCX_CONV_KSHAPE 1 /* cc(centroid, member): SBD.cuh:61,78                   */

/* precision of S = Y^T Y in shape extraction */
#define TCX_S_SPLIT_BF16 0 /* hi+lo BF16 MFMA, FP32 accumulation (default)      */
#define TCX_S_BF16 1       /* single BF16 MFMA product, FP32 accumulation        */
#define TCX_S_AUTO 2       /* split-BF16 until a refinement's BF16 centroids are
                              within s_budget of its split ones, then BF16 for
                              the rest of the fit (checked at iterations 2..4)  */

typedef struct tcx_ctx tcx_ctx;

/* Message of the last failing call on this thread ("" if none). */
const char* tcx_last_error(void);

/* Context: device, stream, cached device workspace, FFT twiddle tables.
 * hbm_budget = 0 means no limit beyond the device's memory.  No reference
 * counterpart: the reference allocates and frees per call
 * (kshape.cuh:209-216, 317-323; ncc.cuh:484-488). */
int tcx_create(int device, size_t hbm_budget, tcx_ctx** out);
int tcx_destroy(tcx_ctx* ctx);
/* Bytes of workspace the context currently holds. */
size_t tcx_workspace_bytes(const tcx_ctx* ctx);
/* Upper bound of the workspace bytes one fit of n series of length m into k
 * clusters grows on this context's device (what an HBM budget must leave
 * for it).  0 for invalid sizes.  No reference counterpart. */
size_t tcx_fit_workspace_bytes(const tcx_ctx* ctx, int n, int m, int k);

/* Device memory helpers for hosts without their own allocator (tests,
 * the CLI).  Stand in for the cudaMalloc/cudaMemcpy/cudaFree calls of
 * test_kshape.cu:35-36, 74. */
int tcx_malloc(tcx_ctx* ctx, void** d_ptr, size_t bytes);
int tcx_free(tcx_ctx* ctx, void* d_ptr);
int tcx_memcpy_h2d(tcx_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
int tcx_memcpy_d2h(tcx_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);
int tcx_memset(tcx_ctx* ctx, void* d_ptr, int value, size_t bytes);
int tcx_synchronize(tcx_ctx* ctx);

/* ------------------------------------------------------------ primitives */

/* In-place z-normalization of each row: (x - mean) / std_ddof, FP64
 * accumulation, constant rows -> 0.
 * Replaces z_norm_gpu_x<float> (Code/include/normalization.cuh:149-177);
 * semantics of zscore (Code/test/test_timesC.py:42-52). */
int tcx_znorm(tcx_ctx* ctx, float* d_x, int rows, int cols, int ddof);

/* Materialized normalized cross-correlation d_out[x][y][2m-1]
 * (lags -(m-1)..m-1), linear (non-aliasing) FFT length.
 * Replaces NCC_3D<float> (Code/include/ncc.cuh:467-517);
 * values of _ncc_c_3dim (Code/test/test_timesC.py:105-114). */
int tcx_ncc(tcx_ctx* ctx, const float* d_x, const float* d_y, float* d_out,
            int x_rows, int y_rows, int cols);

/* In-place alignment of `rows` members to one centroid: shift each member
 * by the lag of its maximum NCC with zero fill.  A centroid that is all
 * zeros leaves the members unchanged (test_timesC.py:128).
 * Replaces sbd3D<float> (Code/include/SBD.cuh:68-85); semantics of
 * _preSBD_extract_shape (Code/test/test_timesC.py:127-133). */
int tcx_sbd_align(tcx_ctx* ctx, float* d_members, const float* d_center,
                  int rows, int cols, int convention);

/* Fused SBD assignment: dist[i] = min_j (1 - max_lag NCC(x_i, c_j)),
 * computed exactly; without d_dmat, centroids whose spectral upper bound
 * on the NCC cannot beat the best distance found are skipped,
 * label[i] = first argmin (zero-based), shift[i] = lag of the maximum for
 * the chosen centroid (the shift the next refinement applies).  d_dmat
 * (n*k, may be NULL) receives all distances; d_dist and d_shift may be NULL.
 * Replaces NCC_3D + get_pos_idx + the inertia reduce
 * (Code/include/kshape.cuh:257-271, 35-96); semantics of
 * test_timesC.py:258-260. */
int tcx_assign(tcx_ctx* ctx, const float* d_x, const float* d_centroids,
               int n, int m, int k, int* d_label, float* d_dist, int* d_shift,
               float* d_dmat);

/* tcx_assign without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No reference counterpart.  Same arguments and
 * results as tcx_assign, on the context's stream; the centroid spectra and
 * the FFT twiddle tables live in a static device array of 3.3 MB per GPU
 * (held until the process ends) instead of context workspace: k <= 1024 and
 * k * F <= 262144 (F = 1 << (2m-1).bit_length()), else TCX_ERR_ARG.  The
 * first call on a device writes the twiddle tables there.  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (other threads,
 * other contexts) race and must be serialized by the caller. */
int tcx_assign_noalloc(tcx_ctx* ctx, const float* d_x, const float* d_centroids,
                       int n, int m, int k, int* d_label, float* d_dist,
                       int* d_shift, float* d_dmat);

/* Shape extraction: group rows by d_label (zero-based), align each member
 * by d_shift[i] (ignored for clusters whose old centroid is all zero; a
 * NULL d_shift means: compute each member's shift against its own old
 * centroid, as _preSBD_extract_shape does),
 * z-normalize members (ddof=1), S_j = Y_j^T Y_j on MFMA, centre with
 * rank-1 updates, power iteration with an FP64 Rayleigh-quotient test,
 * sign fix by the lowest-index member, z-normalize (ddof=1).
 * d_centroids: in = old centroids (k*m), out = new centroids; an empty
 * cluster gets a zero centroid.  `iter` selects the power-iteration start
 * vectors.  cal_select 0: power iteration (I = min(m, 1000) steps,
 * evaluated by Lanczos), with clusters that fail its tests taking the top
 * eigenvector; 1: the top eigenvector of every M_j (the eigh branch,
 * test_timesC.py:184-192; shapeExtract.cuh:366-468).  *sum_class (may be
 * NULL) = number of non-empty clusters.
 * Replaces extract_shape<float> (Code/include/shapeExtract.cuh:1057-1181);
 * semantics of _preSBD_extract_shape + _extract_shape
 * (Code/test/test_timesC.py:127-204). */
int tcx_extract_shape(tcx_ctx* ctx, const int* d_label, const float* d_x,
                      const int* d_shift, float* d_centroids, int n, int m,
                      int k, int iter, int cal_select, int* sum_class);

/* ------------------------------------------------------------------- fit */

typedef struct {
  int max_iter;          /* iterations cap (test_timesC.py:221 uses 10)      */
  int init_mode;         /* 0: random labels, 1: random centres (init_chose)  */
  uint64_t seed;         /* initial labels / centres and power start vectors */
  const int* init_labels;/* optional host labels (n, zero-based), mode 0     */
  const int* init_rows;  /* optional host row indices (k), mode 1            */
  int convention;        /* TCX_CONV_CPU (default) or TCX_CONV_KSHAPE         */
  int s_precision;       /* TCX_S_SPLIT_BF16 (default), TCX_S_BF16, TCX_S_AUTO */
  int cal_select;        /* 0: power iteration, 1: top eigenvector (eigh)      */
  double power_tol;      /* relative Rayleigh-quotient tolerance, 0 -> 1e-12 */
  int* labels_per_iter;  /* optional host buffer (max_iter*n): labels after
                            each iteration's assignment                      */
  int stop_on_collapse;  /* 1: stop right after a refinement that leaves
                            sum_class == 1 or sum_class <= k/2 non-empty
                            clusters, without assigning, keeping the labels
                            of the previous assignment (the GPU path's break,
                            kshape.cuh:248-249, 454-455); 0: the CPU path,
                            which has no such stop (default)                 */
  double s_budget;       /* TCX_S_AUTO: the largest relative L2 difference
                            between a refinement's BF16 and split-BF16
                            centroids that switches S to BF16; 0 -> 1e-4    */
} tcx_fit_opts;

typedef struct {
  int iterations;        /* iterations run                                  */
  int converged;         /* 1 if labels stopped changing                    */
  int sum_class;         /* non-empty clusters after the last refinement    */
  int power_steps_max;   /* largest power-iteration step count of the run   */
  int fallbacks;         /* eigen-fallback branches taken                    */
  double ms_total;       /* wall time of the fit (host clock)               */
  double ms_refine;      /* summed refinement time (HIP events)             */
  double ms_assign;      /* summed assignment time (HIP events)             */
  long long sbd_evals;   /* inverse transforms run by the pruned assignment
                            (each covers up to two series x one centroid)   */
  long long sbd_pairs;   /* series x centroid pairs assigned                */
  double ms_assign_bounds; /* part of ms_assign spent forming the pruning
                              bounds of all series (HIP events)              */
  int collapsed;         /* 1 if the fit stopped on collapse (stop_on_collapse) */
  int s_bf16_from;       /* iteration (0-based) from which S ran in BF16; -1:
                            never (TCX_S_AUTO decides it, TCX_S_BF16 gives 0) */
  double s_check_err;    /* TCX_S_AUTO: the last BF16 - split centroid
                            difference measured (relative L2; 0: none)      */
} tcx_fit_stats;

/* Fill opts with defaults (max_iter 10, seed 0, CPU convention,
 * split-BF16, power_tol 1e-12). */
void tcx_fit_opts_default(tcx_fit_opts* opts);

/* Full k-Shape fit on z-normalized device data.  h_labels (n) receives
 * zero-based labels, h_centroids (k*m, may be NULL) the final centroids,
 * stats may be NULL.  The loop is _kshape (test_timesC.py:213-265) for
 * init_mode 0 and the k_shape_tsl order (kshape.cuh:376-461) for 1. */
int tcx_kshape_fit(tcx_ctx* ctx, const float* d_x, int n, int m, int k,
                   const tcx_fit_opts* opts, int* h_labels, float* h_centroids,
                   tcx_fit_stats* stats);

/* Stepwise form of the same fit (one context runs one fit at a time):
 * tcx_fit_begin sets up the initial state, each tcx_fit_step runs one
 * iteration (refine + assign, or assign + stop test + refine for
 * init_mode 1) and reports in *changed (may be NULL) whether any label
 * changed; the caller decides when to stop.  A step may return while its
 * last refine still runs on the context's stream; the next call on the
 * context orders after it.  tcx_fit_end copies out the result (h_labels,
 * h_centroids and stats may each be NULL). */
int tcx_fit_begin(tcx_ctx* ctx, const float* d_x, int n, int m, int k,
                  const tcx_fit_opts* opts);
int tcx_fit_step(tcx_ctx* ctx, int* changed);
int tcx_fit_end(tcx_ctx* ctx, int* h_labels, float* h_centroids,
                tcx_fit_stats* stats);
/* Statistics of the fit in progress (no reference counterpart). */
int tcx_fit_query(tcx_ctx* ctx, tcx_fit_stats* stats);
/* State of the fit in progress, copied to host buffers (each may be NULL):
 * labels, SBD distances and lags of the last assignment (n each; h_dist
 * needs one assignment done) and the current centroids (k*m): for
 * init_mode 0 those the last assignment used, for init_mode 1 those of the
 * refinement that followed it.  No reference counterpart (the reference
 * frees its centroids, kshape.cuh:321); for diagnostics and parity checks
 * of a fit step by step. */
int tcx_fit_snapshot(tcx_ctx* ctx, int* h_labels, float* h_dist, int* h_shift,
                     float* h_centroids);

/* Reference-shaped fit: one-based labels into h_idx, *iter_num in = cap,
 * out = iterations run (left unchanged when the fit stops on collapse, as
 * the reference leaves it), *sum_class out = non-empty clusters after the
 * last refinement.  Stops on collapse (stop_on_collapse = 1), like the
 * reference's loop (kshape.cuh:248-249).
 * Replaces k_shape<float> (Code/include/kshape.cuh:197-324). */
int tcx_kshape(tcx_ctx* ctx, const float* d_x, int* h_idx, int n, int m,
               int k, int* iter_num, int* sum_class, int cal_select);

/* Centre-initialized fit (init_chose=1).  h_x is accepted for signature
 * parity and unused: the centres are rows of the z-normalized d_x.  Stops
 * on collapse after a refinement (kshape.cuh:454-455); *iter_num as for
 * tcx_kshape.
 * Replaces k_shape_tsl<float> (Code/include/kshape.cuh:332-474). */
int tcx_kshape_tsl(tcx_ctx* ctx, const float* d_x, const float* h_x,
                   int* h_idx, int n, int m, int k, int* iter_num,
                   int* sum_class, int cal_select);

/* ------------------------------------------------------ batched fits */
/* A pool of worker threads, each owning one context (stream + workspace)
 * on the same device, that runs many independent fits concurrently.
 * No reference counterpart (the reference runs one fit per process,
 * test_kshape.cu:13-80). */
typedef struct tcx_batch tcx_batch;

typedef struct {
  /* in */
  float* d_x;            /* device [n][m] FP32 series                        */
  int n, m, k;
  int znorm_ddof;        /* >= 0: z-normalize d_x in place first (ddof);
                            < 0: d_x is already normalized                  */
  tcx_fit_opts opts;     /* as for tcx_kshape_fit                            */
  int* h_labels;         /* out: n zero-based labels (required)              */
  float* h_centroids;    /* out: k*m final centroids (may be NULL)           */
  /* out */
  int status;            /* TCX_OK or the job's error code                   */
  int worker;            /* worker that ran the job                          */
  tcx_fit_stats stats;
  double ms_queued;      /* batch start -> job start (host clock)            */
  double ms_run;         /* job start -> job end                             */
  double ms_latency;     /* batch start -> job end                           */
  char message[160];     /* error message of a failed job ("" if none)       */
} tcx_job;

/* Job order of a batch run. */
#define TCX_ORDER_GIVEN 0
#define TCX_ORDER_SHORTEST_FIRST 1 /* by n*k*F log F, smallest first      */
#define TCX_ORDER_LONGEST_FIRST 2  /* largest first (shortest makespan)   */

/* workers >= 1 contexts on `device`, each limited to hbm_budget bytes of
 * workspace (0 = no limit). */
int tcx_batch_create(int device, int workers, size_t hbm_budget, tcx_batch** out);
/* Runs all jobs and returns when every job has finished.  Returns TCX_OK
 * when every job succeeded, else the status of the first failed job (each
 * job's own status is in jobs[i].status).  One batch runs at a time. */
int tcx_batch_run(tcx_batch* b, tcx_job* jobs, int njobs, int order);
/* The same jobs through the grouped executor: each phase of an iteration
 * (grouping, align + pack, S = Y^T Y, power stage, centroid spectra, bound
 * product, pruned assignment) runs as one launch over every job in
 * progress, with three host synchronizations per iteration for the whole
 * batch, on the first worker's stream.  Jobs are admitted while their
 * device memory fits the batch's hbm_budget (0: 80% of the free device
 * memory) and join as others finish.  A job's results do not depend on the
 * other jobs; they equal those of tcx_kshape_fit to rounding (the power
 * stage is the one-block-per-cluster Lanczos up to m = 768 and the
 * two-launch form above; S takes one K chunk per tile).  Jobs with
 * init_mode 1, cal_select 1 or k beyond the pruned assignment's limit run on
 * the worker pool as in tcx_batch_run.  Returns as tcx_batch_run; ms_refine
 * and ms_assign of a grouped job are the wall times of the grouped phases
 * it took part in (shared with the other jobs).  No reference counterpart. */
int tcx_batch_run_grouped(tcx_batch* b, tcx_job* jobs, int njobs);
int tcx_batch_destroy(tcx_batch* b);

/* -------------------------------------------------------- serving loop */
/* A long-running server on one device: fit requests and assignment
 * queries arrive asynchronously (tcx_server_submit), are admitted under an
 * HBM budget and run on worker contexts -- fit workers take fits (and
 * queries when no fit waits), query workers take only queries, so short
 * queries are not stuck behind long fits.  Queries assign series to a
 * resident model: centroids whose spectra stay in HBM.  No reference
 * counterpart. */
typedef struct tcx_server tcx_server;

#define TCX_REQ_FIT 0
#define TCX_REQ_ASSIGN 1

typedef struct {
  int kind;              /* TCX_REQ_FIT or TCX_REQ_ASSIGN                     */
  const float* h_x;      /* host [n][m] series (copied at admission)          */
  int n, m;
  int znorm_ddof;        /* >= 0: z-normalize the series first; < 0: as is    */
  /* fit */
  int k;
  tcx_fit_opts opts;
  int publish_model;     /* 1: register the fitted centroids as a model       */
  /* assign */
  int model;             /* model id to assign against                        */
  /* outputs (host buffers owned by the caller) */
  int* h_labels;         /* n labels (required)                               */
  float* h_centroids;    /* fit: k*m centroids (may be NULL)                  */
  float* h_dist;         /* assign: n SBD distances (may be NULL)             */
  int* h_shift;          /* assign: n lags of the maxima (may be NULL)        */
  int model_out;         /* fit with publish_model: the new model id          */
  tcx_fit_stats stats;   /* fit statistics                                    */
  int status;            /* TCX_OK or the request's error code                */
  int done;              /* set to 1 when the request has finished            */
  double ms_queued;      /* submit -> start (host clock)                      */
  double ms_run;         /* start -> end                                      */
  double ms_latency;     /* submit -> end                                     */
  char message[160];     /* error message ("" if none)                        */
} tcx_request;

/* fit_workers + query_workers >= 1 worker contexts on `device`.
 * hbm_budget (0 = none) bounds the device bytes of the resident models plus
 * every worker's reservation (a running request's staging buffers and, for
 * a fit, tcx_fit_workspace_bytes of its size) or held workspace (idle).  A
 * request is admitted when it fits; idle workers release their workspace
 * when that is what blocks it.  A request that cannot fit beside the
 * resident models is refused (TCX_ERR_BUDGET). */
int tcx_server_create(int device, int fit_workers, int query_workers, size_t hbm_budget,
                      tcx_server** out);
/* Uploads k centroids (k*m, host) as a resident model; *model_id out.
 * TCX_ERR_BUDGET when the model does not fit the budget left now (idle
 * workers are asked to release their workspace; a retry may succeed). */
int tcx_server_add_model(tcx_server* s, const float* h_centroids, int k, int m, int* model_id);
/* Queues a request (the struct and its buffers must stay valid until it is
 * done).  Returns at once; errors found at submission are returned here
 * (TCX_ERR_BUDGET: the request needs more than the budget leaves beside the
 * resident models; a queued request that a later model makes so fails with
 * it). */
int tcx_server_submit(tcx_server* s, tcx_request* req);
/* Blocks until req is done; returns its status. */
int tcx_server_wait(tcx_server* s, tcx_request* req);
/* Fails the requests still queued (TCX_ERR_ARG, done = 1), lets the running
 * ones finish, stops the workers and frees the models. */
int tcx_server_destroy(tcx_server* s);

/* ------------------------------------------------------------- host I/O */

/* Labeled CSV "label,v1,...,vm" per line, parsed by several threads
 * (same values as a sequential atof/stoi parse).  *data and *labels are
 * allocated with malloc (free with tcx_free_host).
 * Replaces readFile<float> (Code/include/readfile.cuh:12-67). */
int tcx_read_labeled_csv(const char* path, float** data, int** labels,
                         int* rows, int* cols);
/* Unlabeled CSV "v1,...,vm" per line.
 * Replaces read_unlabeled_File<float> (Code/include/readfile.cuh:72-116). */
int tcx_read_unlabeled_csv(const char* path, float** data, int* rows,
                           int* cols);
/* "l,l,...,l," on one line.
 * Replaces writeFile<int> (Code/include/readfile.cuh:120-129). */
int tcx_write_labels(const char* path, const int* labels, int n);
/* Zero-based labels one per line, the CPU path's output
 * (np.savetxt(..., fmt='%.0f'), test_timesC.py:324-325, 342-343). */
int tcx_write_labels_lines(const char* path, const int* labels, int n);
/* Binary matrix: "TCXB0001", int64 rows, cols, labeled flag, then int32
 * labels (when labeled) and float32 values, little-endian.  labels may be
 * NULL on write (unlabeled) and on read (labels ignored); *labels is set to
 * NULL for an unlabeled file.  No reference counterpart (I/O fast path). */
int tcx_write_binary(const char* path, const float* data, const int* labels, int rows,
                     int cols);
int tcx_read_binary(const char* path, float** data, int** labels, int* rows, int* cols);
void tcx_free_host(void* p);

/* ---------------------------------------------------------- autotuner */

/* Launch shapes that do not change results, measured on the caller's data
 * and kept in a process-wide table that every context and the grouped
 * executor consult: the pruned assignment's grid cap per (F, ceil(k/16),
 * floor(log2 n)) (TCX_TUNE_PRUNE_CAP, keys F, k/16, log2 n) and the
 * align + pack rows per block per m rounded up to 32, m <= 1024
 * (TCX_TUNE_PACK_ROWS, key m; the other keys 0).  No counterpart in the
 * reference (its launch shapes are fixed). */
#define TCX_TUNE_PRUNE_CAP 0
#define TCX_TUNE_PACK_ROWS 1
typedef struct {
  int fft;               /* F of the tuned shape                             */
  int prune_cap;         /* chosen grid cap (0: the built-in one was fastest) */
  int pack_rows;         /* chosen rows per block (0: the built-in choice)   */
  double ms_default;     /* best of reps 3-iteration fits with the built-in shapes */
  double ms_tuned;       /* the same with the chosen shapes                  */
} tcx_tune_report;
/* Times 3-iteration fits of d_x (n x m, z-normalized, k clusters, seed 1)
 * under each candidate, best of `reps`, and keeps the fastest; every
 * candidate must reproduce the labels and centroids of the built-in shapes
 * (else TCX_ERR_NUMERIC, table unchanged). */
int tcx_autotune(tcx_ctx* ctx, const float* d_x, int n, int m, int k, int reps,
                 tcx_tune_report* report);
int tcx_tune_set(int kind, int key0, int key1, int key2, int value);
int tcx_tune_get(int kind, int key0, int key1, int key2, int* value);
void tcx_tune_reset(void);
/* Table file: lines "prune <F> <k/16> <log2 n> <cap>" and "pack <m32>
 * <rows>", '#' comments; load checks every entry before applying any
 * (TCX_ERR_IO). */
int tcx_tune_save(const char* path);
int tcx_tune_load(const char* path);

/* --------------------------------------------- synthetic data (host) */

/* Seeded generator of the benchmark workloads: per class a prototype
 * sum_{q<3} A_q sin(2 pi f_q t/m + phi_q) + exp(-(t-mu)^2/(2 (m/20)^2));
 * series a*roll(proto, s) + b + N(0, sigma^2).  class_frac (K, may be NULL
 * for balanced classes) gives class proportions.  X (n*m) and y (n) out.
 * No reference counterpart (the reference's dataset is not available). */
int tcx_generate(uint64_t seed, long n, int m, int K, double sigma,
                 const double* class_frac, float* X, int* y);
/* Initial labels in [0,k) and k distinct initial centre rows. */
int tcx_init_labels(uint64_t seed, long n, int k, int* labels);
int tcx_init_rows(uint64_t seed, long n, int k, int* rows);
/* Power-iteration start vector of (iter, cluster), FP64 (length m). */
void tcx_power_v0(uint64_t seed, int iter, int cluster, int m, double* v0);

#ifdef __cplusplus
}
#endif
#endif

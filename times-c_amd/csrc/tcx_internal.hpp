// Engine entry points shared between the library's own translation units
// (not part of the C ABI): resident centroid spectra for the serving loop.
#pragma once
#include <stddef.h>

#include <chrono>

#include "timesc.h"

namespace tcx {

// Launch shapes chosen by the autotuner (0: the built-in choice): the
// pruned assignment's grid cap for (F, k, n), the align + pack rows per
// block for m.  Neither changes results.
int tuned_prune_cap(int F, int k, long n);
int tuned_pack_rows(int m);
constexpr int kTunePackMaxM = 1024;          // align + pack from LDS up to this m
constexpr int kTunePackMaxBytes = 72 * 1024;  // its LDS per block at most

// FFT length of series of length m: 1 << (2m-1).bit_length().
int fft_length(int m);
// Spectra of k centroid rows d_cent (k*m) into caller-owned device buffers:
// spec k*F complex FP32 (2 floats each), mag k*F, norm k, zero k.
// Asynchronous on the context's stream.
int model_spectra(tcx_ctx* c, const float* d_cent, int k, int m, void* spec, float* mag,
                  float* norm, int* zero);
// Pruned (or, beyond the pruned kernel's k limit, full) SBD assignment of
// n series against resident spectra; asynchronous on the context's stream.
int assign_resident(tcx_ctx* c, const float* d_x, long n, int m, int k, const void* spec,
                    const float* mag, const float* norm, int convention, int* label, float* dist,
                    int* shift);
// Waits for the context's stream.
int stream_sync(tcx_ctx* c);
// Device bytes the context holds: workspace slots, twiddle tables and
// tcx_malloc allocations.
size_t held_bytes(const tcx_ctx* c);
// Free device memory of the context's GPU (0 when it cannot be read).
size_t device_free_bytes(tcx_ctx* c);
// Grow-only buffers the context keeps between calls of the grouped executor
// (device memory, or page-locked host memory when pinned), so that repeated
// batches allocate and free nothing; buffer `id` < kCtxBuffers holds at
// least `bytes` afterwards.  Contents are not kept when a buffer grows, and
// growing frees the old one: only while nothing in flight uses it.  Device
// bytes count as workspace (held_bytes); trim_workspace and tcx_destroy
// release them.
constexpr int kCtxBuffers = 16;
int ctx_buffer(tcx_ctx* c, int id, size_t bytes, bool pinned, void** out);
// Device bytes the context's kept buffers hold (reusable by the next call).
size_t ctx_buffer_device_bytes(const tcx_ctx* c);
// A second stream of the context, created on first use (default priority,
// non-blocking) and destroyed with the context.
int ctx_aux_stream(tcx_ctx* c, void** out);
// Frees every workspace slot and twiddle table of an idle context (no fit
// in progress); the next call grows them again.
int trim_workspace(tcx_ctx* c);
// Upper bound of the workspace one fit of (n, m, k) grows on a device with
// `cus` compute units.
size_t fit_workspace_bound(long n, int m, int k, int cus);
// The context's stream (a hipStream_t), device and compute units, and its
// twiddle table of FFT length F (a device float2 array, grown on first use).
void* stream_of(tcx_ctx* c);
int device_of(const tcx_ctx* c);
int cus_of(const tcx_ctx* c);
int twiddle_table(tcx_ctx* c, int F, const void** tw);

// Grouped executor of a batch (host_multi.hip): whether a job can run
// there (why: the reason when not), and the run of jobs[idx[0..count)] on
// the context's stream, admitting jobs while their device memory fits
// hbm_budget (0: most of the free device memory).  Every job gets its
// status, results and timings (from t0); returns TCX_OK unless the run
// itself failed.  second_stream (may be NULL: one is created) carries the
// block-form Gram Lanczos beside the context's stream.
bool grouped_supported(const tcx_job& j, const char** why);
int grouped_fit(tcx_ctx* c, tcx_job* jobs, const int* idx, int count, size_t hbm_budget,
                std::chrono::steady_clock::time_point t0, void* second_stream = nullptr);

}  // namespace tcx

/*
 * gdb_kernels.h — fixed device kernels of libgdbhip.so (no plan, no JIT).
 *
 * None of these has a counterpart in the reference (it has no learned
 * estimator, no device-side generator and no bandwidth calibration); they
 * serve the additions BASELINE's north star asks for:
 *   - the group-count estimator: an FP32 MLP 16 -> 64 -> 64 -> 1 (ReLU) with
 *     inference, per-sample backward and a deterministic gradient reduction
 *     fused with Adam, all on the GPU;
 *   - device generators for the estimator's training runs;
 *   - a streaming copy used to calibrate FETCH_SIZE/WRITE_SIZE counters and
 *     to measure the achievable HBM bandwidth.
 * Conventions as gdb_hip.h (gdb_status, asynchronous on the given stream).
 */
#ifndef GDB_KERNELS_H
#define GDB_KERNELS_H

#include "gdb_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GDB_MLP_IN 16
#define GDB_MLP_HIDDEN 64
#define GDB_MLP_NPARAMS 5313 /* W1 64x16, b1 64, W2 64x64, b2 64, W3 64, b3 1 */

const char* gdb_kernels_last_error(void);

/* Y[s] = MLP(params, X[s][0..15]) for s < n. */
gdb_status gdb_mlp_forward(const float* params, const float* X, float* Y, int n, gdb_stream s);
/* Per-sample gradients of scale * sum_s (Y[s] - T[s])^2: G[s][GDB_MLP_NPARAMS],
 * and L[s] = scale * (Y[s] - T[s])^2. */
gdb_status gdb_mlp_backward(const float* params, const float* X, const float* T, float* G, float* L, int n,
                            float scale, gdb_stream s);
/* grad[p] = sum_s G[s][p] (sample order, deterministic; grad may be NULL);
 * when lr > 0, one Adam step (bias-corrected, step >= 1) on params/m/v. */
gdb_status gdb_mlp_reduce_adam(float* params, float* m, float* v, const float* G, float* grad, int n, int step,
                               float lr, float beta1, float beta2, float eps, gdb_stream s);

/* Allocation-free training step: gdb_mlp_backward then gdb_mlp_reduce_adam
 * with the per-sample gradient rows held in the library's static device
 * workspace instead of a caller buffer G. No counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops). At most GDB_MLP_NOALLOC_MAX_SAMPLES samples
 * per call (larger batches are rejected); no hipMalloc/hipFree. Calls that
 * overlap on one GPU share its workspace and race. L[n] receives the
 * per-sample losses; grad (NULL allowed) the summed gradient; m, v may be
 * NULL when lr <= 0. */
#define GDB_MLP_NOALLOC_MAX_SAMPLES 256
gdb_status gdb_mlp_train_step_noalloc(float* params, float* m, float* v, const float* X, const float* T, float* L,
                                     float* grad, int n, float scale, int step, float lr, float beta1, float beta2,
                                     float eps, gdb_stream s);

/* Device column generators (counter-based; estimator training runs). */
gdb_status gdb_gen_device_zipf_i32(int32_t* out, uint64_t n, double s, int64_t nkeys, uint64_t seed, gdb_stream st);
gdb_status gdb_gen_device_uniform_i64(int64_t* out, uint64_t n, int64_t lo, int64_t hi, uint64_t seed,
                                      gdb_stream st);

/* dst[i * words + w] = src[i][w] for i < n, w < words: src is a device array
 * of n device pointers (e.g. the status words of many runners, read back
 * with one copy instead of one synchronising read each). No counterpart in
 * the reference (its runner reads status per query from shared memory). */
gdb_status gdb_gather_words(const uint64_t* const* src, int n, int words, uint64_t* dst, gdb_stream st);

/* Streaming device copy (16-byte aligned pointers and size). */
gdb_status gdb_copy_async(void* dst, const void* src, size_t bytes, gdb_stream st);

#ifdef __cplusplus
}
#endif

#endif /* GDB_KERNELS_H */

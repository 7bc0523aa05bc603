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

/* Device column generators (counter-based; estimator training runs). */
gdb_status gdb_gen_device_zipf_i32(int32_t* out, uint64_t n, double s, int64_t nkeys, uint64_t seed, gdb_stream st);
gdb_status gdb_gen_device_uniform_i64(int64_t* out, uint64_t n, int64_t lo, int64_t hi, uint64_t seed,
                                      gdb_stream st);

/* Streaming device copy (16-byte aligned pointers and size). */
gdb_status gdb_copy_async(void* dst, const void* src, size_t bytes, gdb_stream st);

#ifdef __cplusplus
}
#endif

#endif /* GDB_KERNELS_H */

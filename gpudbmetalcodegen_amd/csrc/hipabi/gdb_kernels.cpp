// gdb_kernels.cpp — fixed (ahead-of-time compiled) device kernels of the
// C-ABI library: the group-count estimator's MLP (inference, per-sample
// backward, deterministic gradient reduction + Adam), device-side column
// generators for the estimator's training runs, and a streaming copy used to
// calibrate HBM counters. Generated query kernels are compiled at run time;
// nothing here depends on a plan.

#include "gdb/gdb_hip.h"
#include "gdb/gdb_kernels.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

namespace {

thread_local std::string g_kerr;

gdb_status kfail(gdb_status code, const std::string& m) {
    g_kerr = m;
    return code;
}

gdb_status launchCheck(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return kfail(GDB_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return GDB_OK;
}

// ---------------------------------------------------------------------------
// MLP 16 -> 64 -> 64 -> 1, ReLU, FP32. Parameter layout (floats):
//   W1[64][16] @0, b1[64] @1024, W2[64][64] @1088, b2[64] @5184,
//   W3[64] @5248, b3 @5312.
// One wave per sample; lane j owns hidden unit j. Reductions over lanes use
// xor shuffles in a fixed order, so results are deterministic.
// ---------------------------------------------------------------------------
constexpr int IN = GDB_MLP_IN, H = GDB_MLP_HIDDEN, NP = GDB_MLP_NPARAMS;
constexpr int OW1 = 0, OB1 = 1024, OW2 = 1088, OB2 = 5184, OW3 = 5248, OB3 = 5312;
constexpr int SPB = 4;  // samples (waves) per block

__device__ __forceinline__ float waveSum(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

struct Fwd {
    float a1, h1, a2, h2, y;
};

__device__ __forceinline__ Fwd forwardLane(const float* __restrict__ P, const float* sx, float* sh1, int j) {
    Fwd f;
    float a1 = P[OB1 + j];
#pragma unroll
    for (int i = 0; i < IN; ++i) a1 = fmaf(P[OW1 + j * IN + i], sx[i], a1);
    f.a1 = a1;
    f.h1 = a1 > 0.0f ? a1 : 0.0f;
    sh1[j] = f.h1;
    __syncthreads();
    float a2 = P[OB2 + j];
#pragma unroll 8
    for (int k = 0; k < H; ++k) a2 = fmaf(P[OW2 + j * H + k], sh1[k], a2);
    f.a2 = a2;
    f.h2 = a2 > 0.0f ? a2 : 0.0f;
    f.y = waveSum(P[OW3 + j] * f.h2) + P[OB3];
    return f;
}

__global__ void __launch_bounds__(256) mlpForwardKernel(const float* __restrict__ P, const float* __restrict__ X,
                                                         float* __restrict__ Y, int n) {
    __shared__ float sx[SPB][IN];
    __shared__ float sh1[SPB][H];
    const int w = threadIdx.x >> 6, j = threadIdx.x & 63;
    const int s = blockIdx.x * SPB + w;
    const bool valid = s < n;
    if (j < IN) sx[w][j] = valid ? X[(size_t)s * IN + j] : 0.0f;
    __syncthreads();
    Fwd f = forwardLane(P, sx[w], sh1[w], j);
    if (valid && j == 0) Y[s] = f.y;
}

// Per-sample gradient of L = scale * sum_s (y_s - t_s)^2 into G[s][NP].
__global__ void __launch_bounds__(256) mlpBackwardKernel(const float* __restrict__ P, const float* __restrict__ X,
                                                          const float* __restrict__ T, float* __restrict__ G,
                                                          float* __restrict__ L, int n, float scale) {
    __shared__ float sx[SPB][IN];
    __shared__ float sh1[SPB][H];
    __shared__ float sd2[SPB][H];
    __shared__ float sd1[SPB][H];
    const int w = threadIdx.x >> 6, j = threadIdx.x & 63;
    const int s = blockIdx.x * SPB + w;
    const bool valid = s < n;
    if (j < IN) sx[w][j] = valid ? X[(size_t)s * IN + j] : 0.0f;
    __syncthreads();
    Fwd f = forwardLane(P, sx[w], sh1[w], j);
    const float diff = f.y - (valid ? T[s] : f.y);
    const float d3 = 2.0f * scale * diff;
    float* g = G + (size_t)(valid ? s : 0) * NP;
    // layer 3
    if (valid) {
        g[OW3 + j] = d3 * f.h2;
        if (j == 0) {
            g[OB3] = d3;
            L[s] = scale * diff * diff;
        }
    }
    const float d2 = f.a2 > 0.0f ? d3 * P[OW3 + j] : 0.0f;
    sd2[w][j] = d2;
    __syncthreads();
    if (valid) {
        g[OB2 + j] = d2;
        for (int r = 0; r < H; ++r) g[OW2 + r * H + j] = sd2[w][r] * sh1[w][j];
    }
    float acc = 0.0f;
#pragma unroll 8
    for (int r = 0; r < H; ++r) acc = fmaf(P[OW2 + r * H + j], sd2[w][r], acc);
    const float d1 = f.a1 > 0.0f ? acc : 0.0f;
    sd1[w][j] = d1;
    __syncthreads();
    if (valid) {
        g[OB1 + j] = d1;
#pragma unroll
        for (int t = 0; t < (H * IN) / 64; ++t) {
            const int idx = j + 64 * t;
            g[OW1 + idx] = sd1[w][idx / IN] * sx[w][idx % IN];
        }
    }
}

// grad[p] = sum_s G[s][p] in sample order; then one Adam step when lr > 0.
__global__ void __launch_bounds__(256) mlpReduceAdamKernel(float* __restrict__ P, float* __restrict__ M,
                                                            float* __restrict__ V, const float* __restrict__ G,
                                                            float* __restrict__ grad, int n, float lr, float b1,
                                                            float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NP) return;
    float g = 0.0f;
    for (int s = 0; s < n; ++s) g += G[(size_t)s * NP + p];
    if (grad) grad[p] = g;
    if (lr > 0.0f) {
        const float m = b1 * M[p] + (1.0f - b1) * g;
        const float v = b2 * V[p] + (1.0f - b2) * g * g;
        M[p] = m;
        V[p] = v;
        P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
    }
}

// ---------------------------------------------------------------------------
// Device generators (estimator training runs). Counter-based splitmix64 as
// in the host generators; Zipf by rejection-inversion (Hoermann &
// Derflinger 1996) in double precision.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long dmix(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ unsigned long long drng(unsigned long long seed, unsigned long long stream,
                                                   unsigned long long i) {
    const unsigned long long st = dmix(seed * 0xD1B54A32D192ED03ull + dmix(stream));
    return dmix(st + i * 0x9E3779B97F4A7C15ull);
}

struct DZipf {
    double s, n, hX1, hN, sv;
    __device__ static double helper1(double x) {
        return fabs(x) > 1e-8 ? log1p(x) / x : 1.0 - x * (0.5 - x * (1.0 / 3.0 - 0.25 * x));
    }
    __device__ static double helper2(double x) {
        return fabs(x) > 1e-8 ? expm1(x) / x : 1.0 + x * 0.5 * (1.0 + x * (1.0 / 3.0) * (1.0 + 0.25 * x));
    }
    __device__ double h(double x) const { return exp(-s * log(x)); }
    __device__ double Hf(double x) const {
        const double lx = log(x);
        return helper2((1.0 - s) * lx) * lx;
    }
    __device__ double Hinv(double x) const {
        double t = x * (1.0 - s);
        if (t < -1.0) t = -1.0;
        return exp(helper1(t) * x);
    }
};

__global__ void __launch_bounds__(256) zipfKernel(int* __restrict__ out, unsigned long long n, double s, double nk,
                                                  unsigned long long seed) {
    DZipf z;
    z.s = s;
    z.n = nk;
    z.hX1 = z.Hf(1.5) - 1.0;
    z.hN = z.Hf(nk + 0.5);
    z.sv = 2.0 - z.Hinv(z.Hf(2.5) - z.h(2.0));
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        double k = 1.0;
        for (unsigned long long a = 0; a < 64; ++a) {
            const double r = (double)(drng(seed, 41ull + (a << 40), i) >> 11) * (1.0 / 9007199254740992.0);
            const double u = z.hN + r * (z.hX1 - z.hN);
            const double x = z.Hinv(u);
            k = floor(x + 0.5);
            if (k < 1.0) k = 1.0;
            if (k > nk) k = nk;
            if (k - x <= z.sv || u >= z.Hf(k + 0.5) - z.h(k)) break;
        }
        out[i] = (int)k;
    }
}

__global__ void __launch_bounds__(256) uniformI64Kernel(long long* __restrict__ out, unsigned long long n, long long lo,
                                                        long long hi, unsigned long long seed) {
    const unsigned long long span = (unsigned long long)(hi - lo) + 1ull;
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        const unsigned long long r = drng(seed, 42ull, i);
        out[i] = lo + (long long)__umul64hi(r, span);
    }
}

// dst[i * words + w] = src[i][w]: status words of many runners gathered into
// one buffer, so the host reads them with one copy.
__global__ void __launch_bounds__(256) gatherWordsKernel(const unsigned long long* const* __restrict__ src, int n,
                                                         int words, unsigned long long* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n * words) dst[i] = src[i / words][i % words];
}

// Streaming copy, 16 bytes per lane per iteration.
__global__ void __launch_bounds__(256) copyKernel(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                                  unsigned long long n16) {
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n16;
         i += (unsigned long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Workspace of the allocation-free training step: the per-sample gradient
// rows G[s][NP] of up to GDB_MLP_NOALLOC_MAX_SAMPLES samples (5.4 MB), held by
// every GPU the library runs on until the process ends. Host code never
// passes `scratch` itself to a kernel (in host code it names a host-side
// stand-in); it passes the device address that hipGetSymbolAddress returned
// for the current device, looked up once per device.
constexpr int kMaxDevices = 64;
__device__ __align__(256) float scratch[(size_t)GDB_MLP_NOALLOC_MAX_SAMPLES * NP];
std::once_flag g_scratch_once[kMaxDevices];
float* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

gdb_status scratchFor(float** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return kfail(GDB_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return kfail(GDB_ERR_INVALID, "scratch: device index out of range");
    std::call_once(g_scratch_once[dev], [dev]() {
        void* p = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (g_scratch_err[dev] == hipSuccess) g_scratch_ptr[dev] = static_cast<float*>(p);
    });
    if (g_scratch_err[dev] != hipSuccess || !g_scratch_ptr[dev])
        return kfail(GDB_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(g_scratch_err[dev]));
    *out = g_scratch_ptr[dev];
    return GDB_OK;
}

unsigned gridFor(unsigned long long n, unsigned cap = 8192) {
    unsigned long long g = (n + 255) / 256;
    if (g > cap) g = cap;
    return g == 0 ? 1u : (unsigned)g;
}

}  // namespace

extern "C" {

const char* gdb_kernels_last_error(void) { return g_kerr.c_str(); }

gdb_status gdb_mlp_forward(const float* params, const float* X, float* Y, int n, gdb_stream s) {
    if (n < 0) return kfail(GDB_ERR_INVALID, "gdb_mlp_forward: negative batch");
    if (n == 0) return GDB_OK;
    hipLaunchKernelGGL(mlpForwardKernel, dim3((n + SPB - 1) / SPB), dim3(256), 0, (hipStream_t)s, params, X, Y, n);
    return launchCheck("mlpForwardKernel");
}

gdb_status gdb_mlp_backward(const float* params, const float* X, const float* T, float* G, float* L, int n,
                            float scale, gdb_stream s) {
    if (n <= 0) return kfail(GDB_ERR_INVALID, "gdb_mlp_backward: empty batch");
    hipLaunchKernelGGL(mlpBackwardKernel, dim3((n + SPB - 1) / SPB), dim3(256), 0, (hipStream_t)s, params, X, T, G,
                       L, n, scale);
    return launchCheck("mlpBackwardKernel");
}

gdb_status gdb_mlp_reduce_adam(float* params, float* m, float* v, const float* G, float* grad, int n, int step,
                               float lr, float beta1, float beta2, float eps, gdb_stream s) {
    if (n <= 0) return kfail(GDB_ERR_INVALID, "gdb_mlp_reduce_adam: empty batch");
    const float c1 = 1.0f - std::pow(beta1, (float)step);
    const float c2 = 1.0f - std::pow(beta2, (float)step);
    hipLaunchKernelGGL(mlpReduceAdamKernel, dim3((NP + 255) / 256), dim3(256), 0, (hipStream_t)s, params, m, v, G,
                       grad, n, lr, beta1, beta2, eps, c1 > 0 ? c1 : 1.0f, c2 > 0 ? c2 : 1.0f);
    return launchCheck("mlpReduceAdamKernel");
}

gdb_status gdb_mlp_train_step_noalloc(float* params, float* m, float* v, const float* X, const float* T, float* L,
                                     float* grad, int n, float scale, int step, float lr, float beta1, float beta2,
                                     float eps, gdb_stream s) {
    if (n <= 0) return kfail(GDB_ERR_INVALID, "gdb_mlp_train_step_noalloc: empty batch");
    if (n > GDB_MLP_NOALLOC_MAX_SAMPLES)
        return kfail(GDB_ERR_INVALID, "gdb_mlp_train_step_noalloc: batch of " + std::to_string(n) +
                                          " samples exceeds the static workspace (" +
                                          std::to_string(GDB_MLP_NOALLOC_MAX_SAMPLES) + " samples)");
    if (!params || !X || !T || !L) return kfail(GDB_ERR_INVALID, "gdb_mlp_train_step_noalloc: null argument");
    if (lr > 0.0f && (!m || !v)) return kfail(GDB_ERR_INVALID, "gdb_mlp_train_step_noalloc: Adam moments missing");
    float* G = nullptr;
    const gdb_status st = scratchFor(&G);
    if (st != GDB_OK) return st;
    // the backward kernel writes every element of rows 0..n-1: nothing to clear
    hipLaunchKernelGGL(mlpBackwardKernel, dim3((n + SPB - 1) / SPB), dim3(256), 0, (hipStream_t)s, params, X, T, G, L,
                       n, scale);
    const gdb_status b = launchCheck("mlpBackwardKernel");
    if (b != GDB_OK) return b;
    const float c1 = 1.0f - std::pow(beta1, (float)step);
    const float c2 = 1.0f - std::pow(beta2, (float)step);
    hipLaunchKernelGGL(mlpReduceAdamKernel, dim3((NP + 255) / 256), dim3(256), 0, (hipStream_t)s, params, m, v, G,
                       grad, n, lr, beta1, beta2, eps, c1 > 0 ? c1 : 1.0f, c2 > 0 ? c2 : 1.0f);
    return launchCheck("mlpReduceAdamKernel");
}

gdb_status gdb_gen_device_zipf_i32(int32_t* out, uint64_t n, double s, int64_t nkeys, uint64_t seed, gdb_stream st) {
    if (nkeys < 1 || s < 0.0) return kfail(GDB_ERR_INVALID, "gdb_gen_device_zipf_i32: bad parameters");
    if (n == 0) return GDB_OK;
    hipLaunchKernelGGL(zipfKernel, dim3(gridFor(n)), dim3(256), 0, (hipStream_t)st, out, (unsigned long long)n, s,
                       (double)nkeys, (unsigned long long)seed);
    return launchCheck("zipfKernel");
}

gdb_status gdb_gen_device_uniform_i64(int64_t* out, uint64_t n, int64_t lo, int64_t hi, uint64_t seed, gdb_stream st) {
    if (hi < lo) return kfail(GDB_ERR_INVALID, "gdb_gen_device_uniform_i64: hi < lo");
    if (n == 0) return GDB_OK;
    hipLaunchKernelGGL(uniformI64Kernel, dim3(gridFor(n)), dim3(256), 0, (hipStream_t)st, (long long*)out,
                       (unsigned long long)n, (long long)lo, (long long)hi, (unsigned long long)seed);
    return launchCheck("uniformI64Kernel");
}

gdb_status gdb_gather_words(const uint64_t* const* src, int n, int words, uint64_t* dst, gdb_stream st) {
    if (n < 0 || words < 1 || (n && (!src || !dst))) return kfail(GDB_ERR_INVALID, "gdb_gather_words: bad arguments");
    if (n == 0) return GDB_OK;
    const long long total = (long long)n * words;
    if (total > (1ll << 30)) return kfail(GDB_ERR_INVALID, "gdb_gather_words: too many words");
    hipLaunchKernelGGL(gatherWordsKernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)st,
                       (const unsigned long long* const*)src, n, words, (unsigned long long*)dst);
    return launchCheck("gatherWordsKernel");
}

gdb_status gdb_copy_async(void* dst, const void* src, size_t bytes, gdb_stream st) {
    if (bytes % 16 || ((uintptr_t)dst % 16) || ((uintptr_t)src % 16))
        return kfail(GDB_ERR_INVALID, "gdb_copy_async: 16-byte aligned sizes and pointers only");
    if (!bytes) return GDB_OK;
    const unsigned long long n16 = bytes / 16;
    hipLaunchKernelGGL(copyKernel, dim3(gridFor(n16, 16384)), dim3(256), 0, (hipStream_t)st, (uint4*)dst,
                       (const uint4*)src, n16);
    return launchCheck("copyKernel");
}

}  // extern "C"

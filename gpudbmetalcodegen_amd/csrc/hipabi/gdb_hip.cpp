// gdb_hip.cpp — implementation of the C-ABI in include/gdb/gdb_hip.h.
//
// Everything HIP-specific in the project lives in this file (plus the fixed
// kernels next to it): the engine above talks to the device only through the
// gdb_* functions. hipRTC compiles generated sources for the context's gfx
// architecture; code objects are cached in memory and on disk under a key that
// hashes the source, the options, the architecture and the hipRTC version.

#include "gdb/gdb_hip.h"

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_err;

gdb_status fail(gdb_status code, const std::string& msg) {
    g_err = msg;
    return code;
}

gdb_status hipFail(hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
    (void)hipGetLastError();  // clear the sticky-free error state for the next call
    return fail(e == hipErrorOutOfMemory ? GDB_ERR_OOM : GDB_ERR_HIP, m);
}

#define GDB_HIP_TRY(call, what)                       \
    do {                                              \
        hipError_t _e = (call);                       \
        if (_e != hipSuccess) return hipFail(_e, what); \
    } while (0)

// 64-bit FNV-1a, used for the plan hash (cache key) and payload checksums.
uint64_t fnv1a(const void* data, size_t len, uint64_t h = 1469598103934665603ull) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < len; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

std::string shortArch(const char* gcnArchName) {
    std::string a(gcnArchName);
    auto pos = a.find(':');
    return pos == std::string::npos ? a : a.substr(0, pos);
}

std::string hiprtcVersionString() {
    int maj = 0, min = 0;
    hiprtcVersion(&maj, &min);
    return std::to_string(maj) + "." + std::to_string(min);
}

uint64_t planHash(const std::string& arch, const char* src, size_t len, const char** opts, int nopts) {
    uint64_t h = fnv1a(src, len);
    for (int i = 0; i < nopts; ++i) {
        h = fnv1a("\x1f", 1, h);
        h = fnv1a(opts[i], std::strlen(opts[i]), h);
    }
    h = fnv1a("\x1e", 1, h);
    h = fnv1a(arch.data(), arch.size(), h);
    std::string v = hiprtcVersionString();
    h = fnv1a(v.data(), v.size(), h);
    return h;
}

// On-disk code-object record: magic, key, payload length, payload checksum,
// payload. A record whose header or checksum does not match is treated as a
// miss and rewritten.
constexpr uint64_t kCoMagic = 0x31304F4342444754ull;  // "TGDBCO01"

std::string cachePath(const std::string& dir, uint64_t key) {
    char buf[32];
    std::snprintf(buf, sizeof(buf), "%016llx.co", (unsigned long long)key);
    return dir + "/" + buf;
}

bool readCache(const std::string& dir, uint64_t key, std::vector<char>& out) {
    if (dir.empty()) return false;
    std::ifstream f(cachePath(dir, key), std::ios::binary);
    if (!f) return false;
    uint64_t hdr[4];
    if (!f.read(reinterpret_cast<char*>(hdr), sizeof(hdr))) return false;
    if (hdr[0] != kCoMagic || hdr[1] != key || hdr[2] == 0 || hdr[2] > (1ull << 30)) return false;
    out.resize(hdr[2]);
    if (!f.read(out.data(), static_cast<std::streamsize>(hdr[2]))) return false;
    return fnv1a(out.data(), out.size()) == hdr[3];
}

void writeCache(const std::string& dir, uint64_t key, const std::vector<char>& code) {
    if (dir.empty()) return;
    ::mkdir(dir.c_str(), 0755);
    std::string path = cachePath(dir, key);
    std::string tmp = path + ".tmp" + std::to_string(::getpid());
    {
        std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
        if (!f) return;
        uint64_t hdr[4] = {kCoMagic, key, code.size(), fnv1a(code.data(), code.size())};
        f.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
        f.write(code.data(), static_cast<std::streamsize>(code.size()));
        if (!f) return;
    }
    std::rename(tmp.c_str(), path.c_str());
}

gdb_status rtcCompile(const std::string& arch, const char* src, size_t len, const char** opts,
                      int nopts, std::vector<char>& code, double& ms) {
    auto t0 = std::chrono::steady_clock::now();
    std::string source(src, len);
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, source.c_str(), "gdb_plan.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        return fail(GDB_ERR_COMPILE, "hiprtcCreateProgram failed");
    std::vector<std::string> all;
    all.push_back("--offload-arch=" + arch);
    all.push_back("-O3");
    all.push_back("-std=c++17");
    // Packed FP32 instructions are disabled for all generated code.
    all.push_back("-Xclang");
    all.push_back("-target-feature");
    all.push_back("-Xclang");
    all.push_back("-packed-fp32-ops");
    for (int i = 0; i < nopts; ++i) all.push_back(opts[i]);
    std::vector<const char*> argv;
    for (auto& s : all) argv.push_back(s.c_str());
    hiprtcResult r = hiprtcCompileProgram(prog, static_cast<int>(argv.size()), argv.data());
    size_t logSize = 0;
    hiprtcGetProgramLogSize(prog, &logSize);
    std::string log(logSize, '\0');
    if (logSize) hiprtcGetProgramLog(prog, &log[0]);
    if (r != HIPRTC_SUCCESS) {
        // Numbered source dump, as the reference's compile-error path prints.
        std::ostringstream os;
        os << "hipRTC compile failed (" << hiprtcGetErrorString(r) << ")\n" << log << "\n--- source ---\n";
        std::istringstream is(source);
        std::string line;
        int n = 1;
        while (std::getline(is, line)) os << n++ << ": " << line << "\n";
        hiprtcDestroyProgram(&prog);
        return fail(GDB_ERR_COMPILE, os.str());
    }
    size_t codeSize = 0;
    hiprtcGetCodeSize(prog, &codeSize);
    code.resize(codeSize);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return GDB_OK;
}

// Fill kernel: repeats a 256-byte pattern (already expanded by the host) over
// the destination, 8 bytes per store; a byte tail handles odd sizes.
struct FillPattern {
    unsigned long long w[32];
};

__global__ void __launch_bounds__(256) gdb_fill_kernel(unsigned long long* __restrict__ dst,
                                                       size_t nwords, FillPattern pat,
                                                       unsigned char* __restrict__ tail,
                                                       unsigned tail_bytes, unsigned tail_phase) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < nwords; i += stride) dst[i] = pat.w[i & 31];
    if (blockIdx.x == 0 && threadIdx.x < tail_bytes) {
        const unsigned char* pb = reinterpret_cast<const unsigned char*>(pat.w);
        tail[threadIdx.x] = pb[(tail_phase + threadIdx.x) & 255];
    }
}

// Multi-buffer fill: one launch initialises up to kFillMax buffers, each with
// a pattern of 1..4 words repeated; blockIdx.y selects the buffer. Used to
// issue all of a run's buffer initialisations at once (48 descriptors of 48
// bytes: 2.3 KB of kernel arguments; the pattern period, in words, rides in
// the top two bits of the word count).
constexpr int kFillMax = 48;
struct FillDesc {
    unsigned long long* dst;
    unsigned long long nwords;  // words | (period - 1) << 62
    unsigned long long w[4];
};
struct FillDescs {
    FillDesc d[kFillMax];
};

// Words i = start, start + stride, ... of one buffer set to w[i % period]:
// the phase advances by stride % period per step (one conditional subtract).
__device__ __forceinline__ void gdb_fill_words(unsigned long long* __restrict__ dst, size_t n, unsigned period,
                                               const unsigned long long* w) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (period == 1) {
        const unsigned long long v = w[0];
        for (; i < n; i += stride) dst[i] = v;
        return;
    }
    unsigned ph = (unsigned)(i % period);
    const unsigned step = (unsigned)(stride % period);
    for (; i < n; i += stride) {
        dst[i] = w[ph];
        ph += step;
        if (ph >= period) ph -= period;
    }
}

__global__ void __launch_bounds__(256) gdb_fill_multi_kernel(FillDescs ds) {
    const FillDesc& f = ds.d[blockIdx.y];
    unsigned long long w[4] = {f.w[0], f.w[1], f.w[2], f.w[3]};
    gdb_fill_words(f.dst, (size_t)(f.nwords & ((1ull << 62) - 1ull)), (unsigned)(f.nwords >> 62) + 1u, w);
}

__global__ void __launch_bounds__(256) gdb_fill_table_kernel(const gdb_fill_desc* __restrict__ t) {
    const gdb_fill_desc f = t[blockIdx.y];
    unsigned long long w[4] = {f.pattern[0], f.pattern[1], f.pattern[2], f.pattern[3]};
    gdb_fill_words(static_cast<unsigned long long*>(f.dst), (size_t)f.words, (unsigned)f.period, w);
}

// A fill pattern of 1, 2 or 4 bytes widened to one word, or 1..4 whole words:
// its period in words, or 0 when the buffer cannot take it.
int fillPeriod(size_t bytes, int pl, const void* dst) {
    if (bytes % 8 != 0 || reinterpret_cast<uintptr_t>(dst) % 8 != 0 || pl < 1) return 0;
    if (pl == 1 || pl == 2 || pl == 4) return 1;
    if (pl % 8 == 0 && pl <= 32 && bytes % static_cast<size_t>(pl) == 0) return pl / 8;
    return 0;
}

}  // namespace

struct gdb_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string arch;
    std::string cacheDir;
    std::mutex mu;
    std::unordered_map<uint64_t, std::vector<char>> memCache;
    uint64_t hits = 0, misses = 0;
    double compileMs = 0.0;
    int cuCount = 0;
};

struct gdb_module_s {
    gdb_ctx ctx = nullptr;
    hipModule_t mod = nullptr;
    uint64_t key = 0;
    std::mutex mu;
    std::unordered_map<std::string, gdb_kernel> kernels;
};

struct gdb_kernel_s {
    hipFunction_t fn = nullptr;
    std::string name;
};

struct gdb_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

const char* gdb_last_error(void) { return g_err.c_str(); }

gdb_status gdb_device_count(int* count) {
    if (!count) return fail(GDB_ERR_INVALID, "gdb_device_count: null output");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *count = n;
    return GDB_OK;
}

gdb_status gdb_init(int device, gdb_ctx* out) {
    if (!out) return fail(GDB_ERR_INVALID, "gdb_init: null output");
    int n = 0;
    gdb_device_count(&n);
    if (n <= 0) return fail(GDB_ERR_NO_DEVICE, "gdb_init: no GPU is visible to this process");
    if (device < 0 || device >= n) return fail(GDB_ERR_INVALID, "gdb_init: device index out of range");
    GDB_HIP_TRY(hipSetDevice(device), "hipSetDevice");
    hipDeviceProp_t prop;
    GDB_HIP_TRY(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    auto* c = new gdb_ctx_s();
    c->device = device;
    c->arch = shortArch(prop.gcnArchName);
    c->cuCount = prop.multiProcessorCount;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hipFail(e, "hipStreamCreateWithFlags");
    }
    const char* dir = std::getenv("GDB_CO_CACHE");
    if (dir && *dir) c->cacheDir = dir;
    *out = c;
    return GDB_OK;
}

gdb_status gdb_shutdown(gdb_ctx ctx) {
    if (!ctx) return GDB_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return GDB_OK;
}

gdb_status gdb_device_info(gdb_ctx ctx, gdb_device_info_t* out) {
    if (!ctx || !out) return fail(GDB_ERR_INVALID, "gdb_device_info: null argument");
    hipDeviceProp_t prop;
    GDB_HIP_TRY(hipGetDeviceProperties(&prop, ctx->device), "hipGetDeviceProperties");
    std::memset(out, 0, sizeof(*out));
    out->device = ctx->device;
    out->cu_count = prop.multiProcessorCount;
    out->lds_per_block = static_cast<int>(prop.sharedMemPerBlock);
    out->warp_size = prop.warpSize;
    out->clock_khz = prop.clockRate;
    out->mem_clock_khz = prop.memoryClockRate;
    out->mem_bus_width = prop.memoryBusWidth;
    out->l2_bytes = prop.l2CacheSize;
    out->hbm_bytes = static_cast<long long>(prop.totalGlobalMem);
    std::snprintf(out->arch, sizeof(out->arch), "%s", prop.gcnArchName);
    std::snprintf(out->name, sizeof(out->name), "%s", prop.name);
    return GDB_OK;
}

gdb_stream gdb_default_stream(gdb_ctx ctx) { return ctx ? static_cast<gdb_stream>(ctx->stream) : nullptr; }

gdb_status gdb_alloc(gdb_ctx ctx, size_t bytes, void** dptr) {
    if (!ctx || !dptr) return fail(GDB_ERR_INVALID, "gdb_alloc: null argument");
    *dptr = nullptr;
    if (bytes == 0) bytes = 256;
    GDB_HIP_TRY(hipMalloc(dptr, bytes), "hipMalloc");
    return GDB_OK;
}

gdb_status gdb_free(gdb_ctx ctx, void* dptr) {
    (void)ctx;
    if (!dptr) return GDB_OK;
    GDB_HIP_TRY(hipFree(dptr), "hipFree");
    return GDB_OK;
}

gdb_status gdb_mem_info(gdb_ctx ctx, size_t* free_bytes, size_t* total_bytes) {
    if (!ctx) return fail(GDB_ERR_INVALID, "gdb_mem_info: null context");
    size_t f = 0, t = 0;
    GDB_HIP_TRY(hipMemGetInfo(&f, &t), "hipMemGetInfo");
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return GDB_OK;
}

gdb_status gdb_host_alloc_pinned(size_t bytes, void** hptr) {
    if (!hptr) return fail(GDB_ERR_INVALID, "gdb_host_alloc_pinned: null output");
    GDB_HIP_TRY(hipHostMalloc(hptr, bytes ? bytes : 64, hipHostMallocDefault), "hipHostMalloc");
    return GDB_OK;
}

gdb_status gdb_host_free_pinned(void* hptr) {
    if (!hptr) return GDB_OK;
    GDB_HIP_TRY(hipHostFree(hptr), "hipHostFree");
    return GDB_OK;
}

gdb_status gdb_memcpy_h2d_async(void* dst, const void* src, size_t bytes, gdb_stream s) {
    if (!bytes) return GDB_OK;
    GDB_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(s)),
                "hipMemcpyAsync(H2D)");
    return GDB_OK;
}

gdb_status gdb_memcpy_d2h_async(void* dst, const void* src, size_t bytes, gdb_stream s) {
    if (!bytes) return GDB_OK;
    GDB_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, static_cast<hipStream_t>(s)),
                "hipMemcpyAsync(D2H)");
    return GDB_OK;
}

gdb_status gdb_memcpy_d2d_async(void* dst, const void* src, size_t bytes, gdb_stream s) {
    if (!bytes) return GDB_OK;
    GDB_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(s)),
                "hipMemcpyAsync(D2D)");
    return GDB_OK;
}

gdb_status gdb_fill_multi_async(int n, void* const* dst, const size_t* bytes, const void* const* patterns,
                                const int* pattern_len, gdb_stream s) {
    if (n <= 0) return GDB_OK;
    if (!dst || !bytes || !patterns || !pattern_len) return fail(GDB_ERR_INVALID, "gdb_fill_multi_async: null argument");
    for (int i = 0; i < n; ++i) {
        if (!dst[i] || !patterns[i] || fillPeriod(bytes[i], pattern_len[i], dst[i]) == 0)
            return fail(GDB_ERR_INVALID, "gdb_fill_multi_async: buffer " + std::to_string(i) +
                                             " needs 8-byte alignment and size and a pattern of 1, 2, 4 bytes or "
                                             "1..4 words dividing the size");
    }
    for (int b = 0; b < n; b += kFillMax) {
        const int m = std::min(kFillMax, n - b);
        FillDescs ds{};
        size_t maxWords = 0;
        for (int j = 0; j < m; ++j) {
            FillDesc& f = ds.d[j];
            const int pl = pattern_len[b + j];
            const int period = fillPeriod(bytes[b + j], pl, dst[b + j]);
            f.dst = static_cast<unsigned long long*>(dst[b + j]);
            f.nwords = (bytes[b + j] / 8) | (static_cast<unsigned long long>(period - 1) << 62);
            unsigned char* pb = reinterpret_cast<unsigned char*>(f.w);
            for (int k = 0; k < 8 * period; ++k) pb[k] = static_cast<const unsigned char*>(patterns[b + j])[k % pl];
            maxWords = std::max<size_t>(maxWords, bytes[b + j] / 8);
        }
        size_t blocks = (maxWords + 255) / 256;
        if (blocks > 4096) blocks = 4096;
        if (blocks == 0) blocks = 1;
        hipLaunchKernelGGL(gdb_fill_multi_kernel, dim3(static_cast<unsigned>(blocks), static_cast<unsigned>(m)), dim3(256),
                           0, static_cast<hipStream_t>(s), ds);
        GDB_HIP_TRY(hipGetLastError(), "gdb_fill_multi_kernel launch");
    }
    return GDB_OK;
}

gdb_status gdb_fill_table_async(const gdb_fill_desc* device_table, int n, uint64_t max_words, gdb_stream s) {
    if (n <= 0) return GDB_OK;
    if (!device_table) return fail(GDB_ERR_INVALID, "gdb_fill_table_async: null table");
    if (n > 65535) return fail(GDB_ERR_INVALID, "gdb_fill_table_async: more than 65535 buffers");
    size_t blocks = (static_cast<size_t>(max_words) + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(gdb_fill_table_kernel, dim3(static_cast<unsigned>(blocks), static_cast<unsigned>(n)), dim3(256), 0,
                       static_cast<hipStream_t>(s), device_table);
    GDB_HIP_TRY(hipGetLastError(), "gdb_fill_table_kernel launch");
    return GDB_OK;
}

gdb_status gdb_fill_async(void* dst, size_t bytes, const void* pattern, int pattern_len, gdb_stream s) {
    if (!bytes) return GDB_OK;
    if (!dst || !pattern) return fail(GDB_ERR_INVALID, "gdb_fill_async: null argument");
    if (pattern_len < 1 || pattern_len > 256 || (256 % pattern_len) != 0)
        return fail(GDB_ERR_INVALID, "gdb_fill_async: pattern_len must divide 256");
    if (bytes % static_cast<size_t>(pattern_len) != 0)
        return fail(GDB_ERR_INVALID, "gdb_fill_async: size is not a multiple of the pattern");
    if (reinterpret_cast<uintptr_t>(dst) % 8 != 0)
        return fail(GDB_ERR_INVALID, "gdb_fill_async: destination must be 8-byte aligned");
    FillPattern pat;
    unsigned char* pb = reinterpret_cast<unsigned char*>(pat.w);
    for (int i = 0; i < 256; ++i) pb[i] = static_cast<const unsigned char*>(pattern)[i % pattern_len];
    const size_t nwords = bytes / 8;
    const unsigned tail = static_cast<unsigned>(bytes % 8);
    unsigned char* tailPtr = static_cast<unsigned char*>(dst) + nwords * 8;
    const unsigned phase = static_cast<unsigned>((nwords * 8) % 256);
    size_t blocks = (nwords + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(gdb_fill_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                       static_cast<hipStream_t>(s), static_cast<unsigned long long*>(dst), nwords, pat,
                       tailPtr, tail, phase);
    GDB_HIP_TRY(hipGetLastError(), "gdb_fill_kernel launch");
    return GDB_OK;
}

gdb_status gdb_memset_async(void* dst, int value, size_t bytes, gdb_stream s) {
    unsigned char b = static_cast<unsigned char>(value);
    return gdb_fill_async(dst, bytes, &b, 1, s);
}

gdb_status gdb_module_load(gdb_ctx ctx, const char* src, size_t len, const char** opts, int nopts,
                           uint64_t* plan_hash, gdb_module* out) {
    if (!ctx || !src || !out) return fail(GDB_ERR_INVALID, "gdb_module_load: null argument");
    const uint64_t key = planHash(ctx->arch, src, len, opts, nopts);
    if (plan_hash) *plan_hash = key;
    std::vector<char> code;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        auto it = ctx->memCache.find(key);
        if (it != ctx->memCache.end()) {
            code = it->second;
            ctx->hits++;
        } else if (readCache(ctx->cacheDir, key, code)) {
            ctx->memCache[key] = code;
            ctx->hits++;
        }
    }
    if (code.empty()) {
        double ms = 0.0;
        gdb_status st = rtcCompile(ctx->arch, src, len, opts, nopts, code, ms);
        if (st != GDB_OK) return st;
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->misses++;
        ctx->compileMs += ms;
        ctx->memCache[key] = code;
        writeCache(ctx->cacheDir, key, code);
    }
    GDB_HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    hipModule_t mod = nullptr;
    GDB_HIP_TRY(hipModuleLoadData(&mod, code.data()), "hipModuleLoadData");
    auto* m = new gdb_module_s();
    m->ctx = ctx;
    m->mod = mod;
    m->key = key;
    *out = m;
    return GDB_OK;
}

gdb_status gdb_module_precompile(const char* cache_dir, const char* arch, const char* src, size_t len,
                                 const char** opts, int nopts, uint64_t* plan_hash, int* was_cached) {
    if (!cache_dir || !arch || !src) return fail(GDB_ERR_INVALID, "gdb_module_precompile: null argument");
    const uint64_t key = planHash(arch, src, len, opts, nopts);
    if (plan_hash) *plan_hash = key;
    std::vector<char> code;
    if (readCache(cache_dir, key, code)) {
        if (was_cached) *was_cached = 1;
        return GDB_OK;
    }
    if (was_cached) *was_cached = 0;
    double ms = 0.0;
    gdb_status st = rtcCompile(arch, src, len, opts, nopts, code, ms);
    if (st != GDB_OK) return st;
    writeCache(cache_dir, key, code);
    return GDB_OK;
}

gdb_status gdb_module_unload(gdb_module m) {
    if (!m) return GDB_OK;
    for (auto& kv : m->kernels) delete kv.second;
    (void)hipModuleUnload(m->mod);
    delete m;
    return GDB_OK;
}

gdb_status gdb_kernel_get(gdb_module m, const char* name, gdb_kernel* out) {
    if (!m || !name || !out) return fail(GDB_ERR_INVALID, "gdb_kernel_get: null argument");
    std::lock_guard<std::mutex> lk(m->mu);
    auto it = m->kernels.find(name);
    if (it != m->kernels.end()) {
        *out = it->second;
        return GDB_OK;
    }
    hipFunction_t fn = nullptr;
    hipError_t e = hipModuleGetFunction(&fn, m->mod, name);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(GDB_ERR_NOT_FOUND, std::string("kernel not found in module: ") + name);
    }
    auto* k = new gdb_kernel_s();
    k->fn = fn;
    k->name = name;
    m->kernels[name] = k;
    *out = k;
    return GDB_OK;
}

gdb_status gdb_kernel_occupancy(gdb_kernel k, int block, size_t dyn_lds, int* blocks_per_cu) {
    if (!k || !blocks_per_cu) return fail(GDB_ERR_INVALID, "gdb_kernel_occupancy: null argument");
    int nb = 0;
    GDB_HIP_TRY(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k->fn, block, dyn_lds),
                "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
    *blocks_per_cu = nb;
    return GDB_OK;
}

gdb_status gdb_kernel_attributes(gdb_kernel k, int* num_regs, int* static_lds, int* max_threads,
                                 int* spill_bytes) {
    if (!k) return fail(GDB_ERR_INVALID, "gdb_kernel_attributes: null kernel");
    int v = 0;
    if (num_regs) {
        GDB_HIP_TRY(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_NUM_REGS, k->fn), "hipFuncGetAttribute");
        *num_regs = v;
    }
    if (static_lds) {
        GDB_HIP_TRY(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, k->fn), "hipFuncGetAttribute");
        *static_lds = v;
    }
    if (max_threads) {
        GDB_HIP_TRY(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, k->fn),
                    "hipFuncGetAttribute");
        *max_threads = v;
    }
    if (spill_bytes) {
        GDB_HIP_TRY(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, k->fn), "hipFuncGetAttribute");
        *spill_bytes = v;
    }
    return GDB_OK;
}

gdb_status gdb_launch(gdb_kernel k, unsigned grid_x, unsigned block_x, unsigned dyn_lds, void* kernarg,
                      size_t kernarg_size, gdb_stream s) {
    if (!k) return fail(GDB_ERR_INVALID, "gdb_launch: null kernel");
    if (grid_x == 0 || block_x == 0 || block_x > 1024)
        return fail(GDB_ERR_INVALID, "gdb_launch: bad launch shape for " + k->name);
    size_t size = kernarg_size;
    void* config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, kernarg, HIP_LAUNCH_PARAM_BUFFER_SIZE, &size,
                      HIP_LAUNCH_PARAM_END};
    GDB_HIP_TRY(hipModuleLaunchKernel(k->fn, grid_x, 1, 1, block_x, 1, 1, dyn_lds,
                                      static_cast<hipStream_t>(s), nullptr, config),
                ("hipModuleLaunchKernel(" + k->name + ")").c_str());
    return GDB_OK;
}

gdb_status gdb_cache_set_dir(gdb_ctx ctx, const char* dir) {
    if (!ctx) return fail(GDB_ERR_INVALID, "gdb_cache_set_dir: null context");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->cacheDir = dir ? dir : "";
    return GDB_OK;
}

gdb_status gdb_cache_stats(gdb_ctx ctx, uint64_t* hits, uint64_t* misses, double* compile_ms) {
    if (!ctx) return fail(GDB_ERR_INVALID, "gdb_cache_stats: null context");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (hits) *hits = ctx->hits;
    if (misses) *misses = ctx->misses;
    if (compile_ms) *compile_ms = ctx->compileMs;
    return GDB_OK;
}

gdb_status gdb_stream_create(gdb_ctx ctx, gdb_stream* out) {
    if (!ctx || !out) return fail(GDB_ERR_INVALID, "gdb_stream_create: null argument");
    hipStream_t s = nullptr;
    GDB_HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    GDB_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    *out = s;
    return GDB_OK;
}

gdb_status gdb_stream_sync(gdb_stream s) {
    GDB_HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(s)), "hipStreamSynchronize");
    return GDB_OK;
}

gdb_status gdb_stream_query(gdb_stream s, int* done) {
    hipError_t e = hipStreamQuery(static_cast<hipStream_t>(s));
    if (e == hipSuccess) {
        *done = 1;
        return GDB_OK;
    }
    if (e == hipErrorNotReady) {
        (void)hipGetLastError();
        *done = 0;
        return GDB_OK;
    }
    return hipFail(e, "hipStreamQuery");
}

gdb_status gdb_stream_destroy(gdb_stream s) {
    if (!s) return GDB_OK;
    GDB_HIP_TRY(hipStreamDestroy(static_cast<hipStream_t>(s)), "hipStreamDestroy");
    return GDB_OK;
}

gdb_status gdb_stream_wait_event(gdb_stream s, gdb_event e) {
    GDB_HIP_TRY(hipStreamWaitEvent(static_cast<hipStream_t>(s), static_cast<hipEvent_t>(e), 0),
                "hipStreamWaitEvent");
    return GDB_OK;
}

gdb_status gdb_event_create(gdb_event* out) {
    if (!out) return fail(GDB_ERR_INVALID, "gdb_event_create: null output");
    hipEvent_t e = nullptr;
    GDB_HIP_TRY(hipEventCreate(&e), "hipEventCreate");
    *out = e;
    return GDB_OK;
}

gdb_status gdb_event_record(gdb_event e, gdb_stream s) {
    GDB_HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(e), static_cast<hipStream_t>(s)), "hipEventRecord");
    return GDB_OK;
}

gdb_status gdb_event_sync(gdb_event e) {
    GDB_HIP_TRY(hipEventSynchronize(static_cast<hipEvent_t>(e)), "hipEventSynchronize");
    return GDB_OK;
}

gdb_status gdb_event_query(gdb_event e, int* done) {
    hipError_t r = hipEventQuery(static_cast<hipEvent_t>(e));
    if (r == hipSuccess) {
        *done = 1;
        return GDB_OK;
    }
    if (r == hipErrorNotReady) {
        (void)hipGetLastError();
        *done = 0;
        return GDB_OK;
    }
    return hipFail(r, "hipEventQuery");
}

gdb_status gdb_event_elapsed_ms(gdb_event start, gdb_event stop, float* ms) {
    GDB_HIP_TRY(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)),
                "hipEventElapsedTime");
    return GDB_OK;
}

gdb_status gdb_event_destroy(gdb_event e) {
    if (!e) return GDB_OK;
    GDB_HIP_TRY(hipEventDestroy(static_cast<hipEvent_t>(e)), "hipEventDestroy");
    return GDB_OK;
}

gdb_status gdb_device_sync(gdb_ctx ctx) {
    if (ctx) (void)hipSetDevice(ctx->device);
    GDB_HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return GDB_OK;
}

gdb_status gdb_graph_begin_capture(gdb_stream s) {
    GDB_HIP_TRY(hipStreamBeginCapture(static_cast<hipStream_t>(s), hipStreamCaptureModeThreadLocal),
                "hipStreamBeginCapture");
    return GDB_OK;
}

gdb_status gdb_graph_end_capture(gdb_stream s, gdb_graph* out) {
    if (!out) return fail(GDB_ERR_INVALID, "gdb_graph_end_capture: null output");
    hipGraph_t g = nullptr;
    GDB_HIP_TRY(hipStreamEndCapture(static_cast<hipStream_t>(s), &g), "hipStreamEndCapture");
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hipFail(e, "hipGraphInstantiate");
    }
    auto* gg = new gdb_graph_s();
    gg->graph = g;
    gg->exec = ex;
    *out = gg;
    return GDB_OK;
}

gdb_status gdb_graph_launch(gdb_graph g, gdb_stream s) {
    if (!g) return fail(GDB_ERR_INVALID, "gdb_graph_launch: null graph");
    GDB_HIP_TRY(hipGraphLaunch(g->exec, static_cast<hipStream_t>(s)), "hipGraphLaunch");
    return GDB_OK;
}

gdb_status gdb_graph_destroy(gdb_graph g) {
    if (!g) return GDB_OK;
    (void)hipGraphExecDestroy(g->exec);
    (void)hipGraphDestroy(g->graph);
    delete g;
    return GDB_OK;
}

}  // extern "C"

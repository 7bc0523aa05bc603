/*
 * gdb_hip.h — thin C-ABI over the HIP runtime, hipRTC and hipGraph for gfx950.
 *
 * This is the device boundary of the engine: the C++ host (plan builder,
 * operator emitters, runner) never includes a HIP header; every device action
 * goes through the functions below. Each entry point replaces the Metal call
 * named in its comment (paths relative to the reference repository root).
 *
 * Conventions
 *   - Every function returns gdb_status (0 = GDB_OK). On failure the message is
 *     available from gdb_last_error(), which is per host thread.
 *   - No C++ exception crosses this boundary.
 *   - Launches, copies and fills are asynchronous on the given stream; a NULL
 *     stream is the context's default stream.
 *   - The context owns compiled modules and the code-object cache (guarded by
 *     a mutex); the caller owns device buffers it allocates.
 */
#ifndef GDB_HIP_H
#define GDB_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int gdb_status;
enum {
    GDB_OK = 0,
    GDB_ERR_HIP = 1,        /* a HIP runtime call failed */
    GDB_ERR_COMPILE = 2,    /* hipRTC rejected the generated source */
    GDB_ERR_INVALID = 3,    /* bad argument (shape, alignment, unknown name) */
    GDB_ERR_OOM = 4,        /* device allocation failed */
    GDB_ERR_OVERFLOW = 5,   /* a kernel reported a capacity overflow */
    GDB_ERR_NOT_FOUND = 6,  /* unknown kernel / symbol / table */
    GDB_ERR_ADMISSION = 7,  /* request exceeds the memory budget */
    GDB_ERR_NO_DEVICE = 8   /* no GPU visible */
};

typedef struct gdb_ctx_s* gdb_ctx;
typedef struct gdb_module_s* gdb_module;
typedef struct gdb_kernel_s* gdb_kernel;
typedef void* gdb_stream;   /* hipStream_t */
typedef void* gdb_event;    /* hipEvent_t */
typedef struct gdb_graph_s* gdb_graph;

typedef struct {
    int device;
    int cu_count;
    int lds_per_block;      /* bytes */
    int warp_size;
    int clock_khz;
    int mem_clock_khz;
    int mem_bus_width;
    int l2_bytes;
    long long hbm_bytes;
    char arch[64];          /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    char name[128];
} gdb_device_info_t;

/* Per-thread message of the last failure on this thread ("" if none). */
const char* gdb_last_error(void);

/* Number of visible GPUs (0 when none). Replaces CreateSystemDefaultDevice
 * (codegen/codegen_main.cpp:1905-1911). */
gdb_status gdb_device_count(int* count);

/* Create a context on `device`: sets the device, creates the default stream.
 * Replaces CreateSystemDefaultDevice + newCommandQueue
 * (codegen/codegen_main.cpp:1905-1911). */
gdb_status gdb_init(int device, gdb_ctx* out);
gdb_status gdb_shutdown(gdb_ctx ctx);
gdb_status gdb_device_info(gdb_ctx ctx, gdb_device_info_t* out);
gdb_stream gdb_default_stream(gdb_ctx ctx);

/* Device memory. Replaces newBuffer(len, Shared)
 * (codegen/execution/metal_generic_executor.cpp:121-122, :142-143). */
gdb_status gdb_alloc(gdb_ctx ctx, size_t bytes, void** dptr);
gdb_status gdb_free(gdb_ctx ctx, void* dptr);
gdb_status gdb_mem_info(gdb_ctx ctx, size_t* free_bytes, size_t* total_bytes);
/* Pinned host staging. Replaces newBuffer(ptr, len, Shared, nullptr)
 * zero-copy columns (src/infra.h:894-897). */
gdb_status gdb_host_alloc_pinned(size_t bytes, void** hptr);
gdb_status gdb_host_free_pinned(void* hptr);

/* Copies. Replace host memcpy/contents() access on shared buffers
 * (src/infra.h:1039; codegen/execution/metal_result_collector.cpp:170-171). */
gdb_status gdb_memcpy_h2d_async(void* dst, const void* src, size_t bytes, gdb_stream s);
gdb_status gdb_memcpy_d2h_async(void* dst, const void* src, size_t bytes, gdb_stream s);
gdb_status gdb_memcpy_d2d_async(void* dst, const void* src, size_t bytes, gdb_stream s);
/* Fill `bytes` (a multiple of pattern_len) with a repeating pattern of
 * 1..256 bytes (a divisor of 256) using a fill kernel (never hipMemset, so it is safe inside a
 * captured graph). Replaces the host memset zero/0xFF init
 * (codegen/execution/metal_generic_executor.cpp:164-175). */
gdb_status gdb_fill_async(void* dst, size_t bytes, const void* pattern, int pattern_len,
                          gdb_stream s);
/* Fill n buffers in one launch (no counterpart in the reference, whose buffer
 * initialisation is a host memset per buffer, metal_generic_executor.cpp:164-175):
 * buffer i gets patterns[i] repeated, a pattern of 1, 2 or 4 bytes or of 1..4
 * whole 8-byte words (8, 16, 24 or 32 bytes) that divides the size; sizes and
 * addresses must be multiples of 8. Graph-safe like gdb_fill_async. */
gdb_status gdb_fill_multi_async(int n, void* const* dst, const size_t* bytes, const void* const* patterns,
                                const int* pattern_len, gdb_stream s);
/* One buffer of a fill table: `words` 8-byte words at dst (8-byte aligned)
 * set to pattern[i % period], period 1..4. */
typedef struct gdb_fill_desc {
    void* dst;
    uint64_t words;
    uint64_t period;
    uint64_t pattern[4];
} gdb_fill_desc;
/* Fill every buffer of a table of n descriptors that is already in device
 * memory, in one launch (no counterpart in the reference; for callers that
 * initialise the same buffers on every replay, e.g. a batch of pipelines
 * captured in a graph). The table must stay allocated until the launch has
 * run; max_words (the largest `words`) sizes the grid. Graph-safe. */
gdb_status gdb_fill_table_async(const gdb_fill_desc* device_table, int n, uint64_t max_words, gdb_stream s);
gdb_status gdb_memset_async(void* dst, int value, size_t bytes, gdb_stream s);

/* JIT. Compile HIP C++ for the context's architecture with hipRTC, or take the
 * code object from the cache keyed by hash(source, options, arch, hipRTC
 * version). `plan_hash` (may be NULL) receives that key. Replaces
 * newLibrary(src, opts) (codegen/execution/runtime_compiler.cpp:20-42).
 * On GDB_ERR_COMPILE, gdb_last_error() holds the hipRTC log. */
gdb_status gdb_module_load(gdb_ctx ctx, const char* src, size_t len, const char** opts,
                           int nopts, uint64_t* plan_hash, gdb_module* out);
/* Compile into the on-disk cache only (no device needed). Used to warm the
 * cache at build time. `arch` e.g. "gfx950". */
gdb_status gdb_module_precompile(const char* cache_dir, const char* arch, const char* src,
                                 size_t len, const char** opts, int nopts,
                                 uint64_t* plan_hash, int* was_cached);
gdb_status gdb_module_unload(gdb_module m);
/* Look up a kernel. Replaces newFunction + newComputePipelineState
 * (codegen/execution/runtime_compiler.cpp:44-67). */
gdb_status gdb_kernel_get(gdb_module m, const char* name, gdb_kernel* out);
/* Max resident blocks per CU for a block size and dynamic LDS, and the
 * kernel's static resource use. Replaces maxTotalThreadsPerThreadgroup
 * (codegen/execution/metal_generic_executor.cpp:245-274). */
gdb_status gdb_kernel_occupancy(gdb_kernel k, int block, size_t dyn_lds, int* blocks_per_cu);
gdb_status gdb_kernel_attributes(gdb_kernel k, int* num_regs, int* static_lds,
                                 int* max_threads, int* spill_bytes);

/* Launch with a packed kernel-argument buffer (natural alignment per
 * argument). Replaces setComputePipelineState/setBuffer/setBytes/
 * dispatchThreadgroups (codegen/execution/metal_generic_executor.cpp:238-275). */
gdb_status gdb_launch(gdb_kernel k, unsigned grid_x, unsigned block_x, unsigned dyn_lds,
                      void* kernarg, size_t kernarg_size, gdb_stream s);

/* Code-object cache controls. */
gdb_status gdb_cache_set_dir(gdb_ctx ctx, const char* dir);
gdb_status gdb_cache_stats(gdb_ctx ctx, uint64_t* hits, uint64_t* misses, double* compile_ms);

/* Streams and events. Replace commandBuffer/commit/waitUntilCompleted and
 * GPUStartTime/GPUEndTime (codegen/execution/metal_generic_executor.cpp:189-390).
 * All streams are created at default priority. */
gdb_status gdb_stream_create(gdb_ctx ctx, gdb_stream* out);
gdb_status gdb_stream_sync(gdb_stream s);
gdb_status gdb_stream_query(gdb_stream s, int* done);
gdb_status gdb_stream_destroy(gdb_stream s);
gdb_status gdb_stream_wait_event(gdb_stream s, gdb_event e);
gdb_status gdb_event_create(gdb_event* out);
gdb_status gdb_event_record(gdb_event e, gdb_stream s);
gdb_status gdb_event_sync(gdb_event e);
gdb_status gdb_event_query(gdb_event e, int* done);
gdb_status gdb_event_elapsed_ms(gdb_event start, gdb_event stop, float* ms);
gdb_status gdb_event_destroy(gdb_event e);
gdb_status gdb_device_sync(gdb_ctx ctx);

/* Graphs: capture a stream's work once and replay it. No Metal counterpart
 * (the reference encodes one command buffer per run,
 * codegen/execution/metal_generic_executor.cpp:338-390). */
gdb_status gdb_graph_begin_capture(gdb_stream s);
gdb_status gdb_graph_end_capture(gdb_stream s, gdb_graph* out);
gdb_status gdb_graph_launch(gdb_graph g, gdb_stream s);
gdb_status gdb_graph_destroy(gdb_graph g);

#ifdef __cplusplus
}
#endif

#endif /* GDB_HIP_H */

/*
 * gdb_engine.h — C-ABI over the engine's host surfaces (plan builder,
 * emitters, runtime compiler, runner, result collector) and the seeded
 * column generators.
 *
 * The engine itself is C++ (csrc/planning, csrc/operators, csrc/core,
 * csrc/execution) and reaches the device only through gdb_hip.h. These entry
 * points expose the reference's four host surfaces to callers that cannot
 * link C++ (tests, benchmarks, a Python/ctypes or Go/cgo front end):
 *   Plan API     buildMetalPlan / generateFromPlan   (codegen/planning/metal_plan_builder.h:73-78)
 *   Runner API   RuntimeCompiler + MetalGenericExecutor
 *                (codegen/execution/runtime_compiler.h:9-41, metal_generic_executor.h:46-164)
 *   Result API   MetalResultCollector / GenericResult (codegen/execution/metal_result_collector.h:19-60)
 *   Column API   loadQueryColumns (src/infra.h:930-1062) -> seeded generators + device columns
 * Same conventions as gdb_hip.h: gdb_status return, gdb_last_error() text,
 * no exception crosses the boundary.
 */
#ifndef GDB_ENGINE_H
#define GDB_ENGINE_H

#include "gdb_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gdb_plan_s* gdb_plan;
typedef struct gdb_executor_s* gdb_executor;
typedef struct gdb_result_s* gdb_result;

/* Message of the last engine failure on this thread. */
const char* gdb_engine_last_error(void);

/* ---- Column generators (replace the .tbl/.colbin loaders, src/infra.h:141-782) ---- */
gdb_status gdb_gen_uniform_i32(uint64_t seed, uint64_t stream, uint64_t n, int64_t lo, int64_t hi, int32_t* out);
gdb_status gdb_gen_uniform_i64(uint64_t seed, uint64_t stream, uint64_t n, int64_t lo, int64_t hi, int64_t* out);
gdb_status gdb_gen_categorical_u8(uint64_t seed, uint64_t stream, uint64_t n, const uint8_t* values,
                                  const uint32_t* weights, int nvalues, uint8_t* out);
gdb_status gdb_gen_permutation_i32(uint64_t seed, uint64_t stream, uint64_t n, int64_t base, int32_t* out);
gdb_status gdb_gen_fanout_total(uint64_t seed, uint64_t stream, uint64_t nparent, int cmin, int cmax,
                                uint64_t* total);
gdb_status gdb_gen_fanout_fkey_i32(uint64_t seed, uint64_t stream, uint64_t nparent, int cmin, int cmax,
                                   uint64_t nchild, int32_t* out);
gdb_status gdb_gen_zipf_i32(uint64_t seed, uint64_t stream, uint64_t n, double s, int64_t nkeys, int32_t* out);
gdb_status gdb_gen_vocabulary(uint64_t seed, int nwords, int minlen, int maxlen, char* out /* nwords*16 */);
gdb_status gdb_gen_text(uint64_t seed, uint64_t stream, uint64_t n, int width, const char* vocab, int nwords,
                        char* out /* n*width */);

/* ---- Plan API ---- */
/* Build a plan by name ("Q1", "Q3", "GROUPBY", ...) with shape options
 * "key=value,..." and run the emitters. Replaces buildMetalPlan +
 * generateFromPlan (codegen/planning/metal_plan_builder.h:73-78). */
gdb_status gdb_plan_build(const char* name, const char* options, gdb_plan* out);
/* Build a plan from SQL text: a SELECT over TPC-H tables joined in a tree
 * of key equalities, [NOT] EXISTS / [NOT] IN subqueries as semi- /
 * anti-joins, with
 * WHERE, GROUP BY, SUM / AVG / MIN / MAX / COUNT, HAVING, ORDER BY, LIMIT, compiled
 * to exact fixed-point HIP expressions. Replaces analyzeSQL + the
 * analysis-based buildMetalPlan dispatch (codegen/planning/query_analyzer.h:80-85,
 * metal_plan_builder.cpp:2035-2076) for those queries; the others go through
 * gdb_plan_build by name. Options: groups=N (dense group capacity;
 * default what fits register accumulators, at most 8; execute reports
 * GDB_ERR_OVERFLOW beyond it, word 3 of buffer d_sql_groups_status holding
 * the exact group count).
 * GDB_ERR_INVALID with gdb_engine_last_error() naming the construct otherwise. */
gdb_status gdb_plan_build_sql(const char* sql, const char* options, gdb_plan* out);
/* JSON layout of a SQL plan's result: table, group keys (aggregate column or
 * string column attached from the representative row), select list as
 * expressions over the aggregate columns, ORDER BY, LIMIT. Two-call
 * pattern as gdb_plan_describe. No counterpart in the reference (its
 * post-processing is per query in C++). */
gdb_status gdb_plan_sql_layout(gdb_plan p, char* buf, size_t cap, size_t* need);
/* Plan tree (the reference's --print-plan) and generated HIP source
 * (--dump-msl). Two-call pattern: need receives the size incl. NUL. */
gdb_status gdb_plan_describe(gdb_plan p, char* buf, size_t cap, size_t* need);
gdb_status gdb_plan_source(gdb_plan p, char* buf, size_t cap, size_t* need);
gdb_status gdb_plan_phase_count(gdb_plan p, int* n);
gdb_status gdb_plan_phase_name(gdb_plan p, int idx, char* buf, size_t cap);
/* Launch configuration of one phase (block size, ITEMS variant); the
 * reference's getPhasesMutable() used by --autotune-tg
 * (codegen/codegen_main.cpp:837-929). */
gdb_status gdb_plan_set_launch(gdb_plan p, const char* phase, int block, int items);
gdb_status gdb_plan_get_launch(gdb_plan p, const char* phase, int* block, int* items);
/* Compile through the context's RuntimeCompiler (hipRTC + plan-hash cache).
 * Replaces RuntimeCompiler::compile + getPipeline
 * (codegen/execution/runtime_compiler.cpp:20-67). */
gdb_status gdb_plan_compile(gdb_ctx ctx, gdb_plan p, double* ms, int* cache_hit, uint64_t* plan_hash);
/* Warm the on-disk code-object cache without a device. */
gdb_status gdb_plan_precompile(gdb_plan p, const char* cache_dir, const char* arch, int* was_cached,
                               uint64_t* plan_hash);
gdb_status gdb_plan_free(gdb_plan p);

/* ---- Runner API (MetalGenericExecutor, metal_generic_executor.h:46-164) ---- */
gdb_status gdb_executor_create(gdb_ctx ctx, gdb_stream s, gdb_executor* out);
gdb_status gdb_executor_destroy(gdb_executor ex);
gdb_status gdb_executor_register_table(gdb_executor ex, const char* table, uint64_t rows);
gdb_status gdb_executor_register_column(gdb_executor ex, const char* table, const char* column,
                                        const void* dptr);
gdb_status gdb_executor_register_buffer(gdb_executor ex, const char* name, void* dptr, size_t bytes);
gdb_status gdb_executor_register_symbol(gdb_executor ex, const char* name, uint64_t value);
gdb_status gdb_executor_get_symbol(gdb_executor ex, const char* name, uint64_t* value, int* found);
gdb_status gdb_executor_register_scalar(gdb_executor ex, const char* name, int64_t value);
gdb_status gdb_executor_set_skip_init(gdb_executor ex, int skip);
// on = 1 (default): gdb_exec_timing holds every phase's device time (an event
// pair around each phase); on = 0: only the whole run's (one event at each
// end), for runs that need no breakdown. No counterpart in the reference,
// whose timings come from command-buffer timestamps
// (codegen/execution/metal_generic_executor.cpp:338-390).
gdb_status gdb_executor_set_phase_timing(gdb_executor ex, int on);

#define GDB_MAX_PHASES 16
typedef struct {
    float total_ms;                    /* measured run, all phases */
    float alloc_ms;                    /* buffer allocation before the runs */
    int nphases;
    float phase_ms[GDB_MAX_PHASES];
    unsigned grid[GDB_MAX_PHASES];
    unsigned block[GDB_MAX_PHASES];
    unsigned items[GDB_MAX_PHASES];
    uint64_t status[GDB_MAX_PHASES][4]; /* [0] overflow flag, [2] keys inserted, [3] LDS spills */
} gdb_exec_timing;

/* Warm up, run `measured` times, time phases with events; then check the
 * phases' status words. A hash-table overflow returns GDB_ERR_OVERFLOW with
 * the timing still filled. Replaces execute(compiled, cg, warmup, measured)
 * (codegen/execution/metal_generic_executor.cpp:289-399). */
gdb_status gdb_executor_execute(gdb_executor ex, gdb_plan p, int warmup, int measured, gdb_exec_timing* out);
/* Phases [first, last) only (last = -1: to the end). Replaces the
 * phase-range execute (metal_generic_executor.h:86-93). */
gdb_status gdb_executor_execute_range(gdb_executor ex, gdb_plan p, int first, int last, int warmup, int measured,
                                      gdb_exec_timing* out);
/* Allocate, then launch without synchronising (graph capture, overlap). */
/* Out-of-core run of a chunkable plan: table `table` (`rows` rows) is read
 * from host memory (ncols columns: names, host pointers, element bytes) in
 * chunks of chunk_rows rows, double-buffered into HBM on a copy stream while
 * the previous chunk's kernels run; phases before the stream table's run
 * once, its phases once per chunk, later phases once. Other tables, scalars
 * and symbols are registered on the executor as for execute. Host columns
 * should be pinned (gdb_host_alloc_pinned) for the copies to overlap.
 * Replaces the chunk loop of codegen/codegen_main.cpp:686-808 over
 * ChunkedColbinTable (codegen/execution/chunked_colbin_loader.h:12-60). */
typedef struct {
    double total_ms;     /* host wall time of the whole run, synchronised */
    double stream_ms;    /* device time, first chunk copy -> last stream-phase kernel */
    uint64_t chunks;
    uint64_t bytes_copied;
    int first_stream_phase, last_stream_phase;
} gdb_chunk_stats;
gdb_status gdb_executor_execute_streamed(gdb_executor ex, gdb_plan p, const char* table, uint64_t rows, int ncols,
                                         const char* const* names, const void* const* host_cols,
                                         const int* elem_bytes, uint64_t chunk_rows, gdb_chunk_stats* out);
/* The same with the stream table read from its .colbin file (a TPC-H table
 * of the catalog): each chunk of the columns the stream phases read is
 * converted from the file's format into the engine's (FLOAT -> int64
 * hundredths) in pinned staging while the previous chunk's kernels run, so
 * the table is never loaded whole into host memory. A plan whose later
 * stream phase needs the whole table first streams it once per pass
 * (chunks counts every pass). Replaces ChunkedColbinTable's per-chunk reads
 * (codegen/execution/chunked_colbin_loader.cpp:167-194). */
gdb_status gdb_executor_execute_streamed_colbin(gdb_executor ex, gdb_plan p, const char* table, const char* colbin_path,
                                                uint64_t chunk_rows, gdb_chunk_stats* out);
/* 1 when the plan may be streamed in chunks (metal_plan_builder.cpp:2078-2107). */
gdb_status gdb_plan_chunkable(gdb_plan p, int* chunkable);
gdb_status gdb_executor_prepare(gdb_executor ex, gdb_plan p);
gdb_status gdb_executor_enqueue(gdb_executor ex, gdb_plan p);
gdb_status gdb_executor_check_status(gdb_executor ex, gdb_plan p);
gdb_status gdb_executor_read_buffer(gdb_executor ex, const char* name, size_t offset, void* host, size_t bytes);
gdb_status gdb_executor_buffer_ptr(gdb_executor ex, const char* name, void** dptr, size_t* bytes);
gdb_status gdb_executor_owned_bytes(gdb_executor ex, size_t* bytes);
gdb_status gdb_executor_release(gdb_executor ex);

/* ---- Result API (GenericResult, metal_result_collector.h:19-60) ---- */
gdb_status gdb_executor_collect(gdb_executor ex, gdb_plan p, gdb_result* out);
gdb_status gdb_result_shape(gdb_result r, uint64_t* rows, int* cols);
/* type: 0 = int64, 1 = int128 (two int64 per row: low word, high word), 2 = string */
gdb_status gdb_result_column(gdb_result r, int idx, char* name, size_t cap, int* type);
gdb_status gdb_result_i64(gdb_result r, int idx, int64_t* out);
gdb_status gdb_result_i128(gdb_result r, int idx, int64_t* out);
gdb_status gdb_result_str(gdb_result r, int idx, char* out, size_t width);
/* Display format of an integer column: scale = fixed-point digits (the
 * value is v / 10^scale), is_date = YYYYMMDD shown as YYYY-MM-DD. */
gdb_status gdb_result_column_format(gdb_result r, int idx, int* scale, int* is_date);
/* Canonical CSV: header, then one line per row; integers exact (scaled
 * columns as exact decimals), dates YYYY-MM-DD, strings quoted when they hold
 * ',' '"' or a line break (GenericResult::toCanonical,
 * metal_result_collector.cpp:85-122, which prints doubles with %.4f). */
gdb_status gdb_result_canonical(gdb_result r, char* buf, size_t cap, size_t* need);
gdb_status gdb_result_free(gdb_result r);

/* ---- TPC-H ingest: .tbl -> .colbin v2, and the golden check ----
 * Catalog of the TPC-H tables (codegen/planning/tpch_schema.h:47-166):
 * dtype 0 INT (int32), 1 FLOAT (float32), 2 DATE (int32 YYYYMMDD), 3 CHAR1
 * (one byte), 4 CHAR_FIXED (width bytes, NUL-padded). */
gdb_status gdb_tpch_column(const char* table, const char* column, int* index, int* dtype, int* width);
/* Column names of a table, comma-separated, in file order. */
gdb_status gdb_tpch_table_columns(const char* table, char* buf, size_t cap, size_t* need);
/* Parse <dir>/<table>.tbl and write <dir>/<table>.colbin unless it is up to
 * date (force = 1 rebuilds); *built = 1 when a file was written. threads <= 0
 * uses the host's hardware threads. Replaces tools/tbl_to_colbin.cpp:106-172
 * for one table (the .colbin layout of src/infra.h:511-771). */
gdb_status gdb_tbl_to_colbin(const char* tbl_path, const char* table, int force, int threads, int* built);
/* Memory-mapped .colbin reader (loadColumnsFromBinary, src/infra.h:579-664).
 * tbl_path may be NULL; when it names an existing file the recorded source
 * size and mtime must match it. */
typedef struct gdb_colbin_s* gdb_colbin;
gdb_status gdb_colbin_open(const char* colbin_path, const char* tbl_path, gdb_colbin* out);
gdb_status gdb_colbin_shape(gdb_colbin f, uint64_t* rows, int* ncols);
/* Column by its TPC-H column index; data stays valid until gdb_colbin_close. */
gdb_status gdb_colbin_column(gdb_colbin f, int column_index, const void** data, uint64_t* bytes, int* dtype,
                             int* width);
void gdb_colbin_close(gdb_colbin f);
/* Compare a canonical CSV with a golden one (compareCanonical,
 * codegen/codegen_main.cpp:63-186): GDB_OK with an empty message on a match,
 * GDB_OK with the first difference in msg otherwise. */
gdb_status gdb_compare_canonical(const char* got, const char* expected, double abs_tol, double rel_tol, char* msg,
                                 size_t cap, size_t* need);

/* ---- Autotuner: sweep block x ITEMS x (optional) runtime scalar for one
 * phase; 1 untimed + `trials` timed runs per candidate, slowest dropped, p50
 * wins (the reference's --autotune-tg loop, codegen/codegen_main.cpp:837-929).
 * The plan's phase is left at the winner; `table` (may be NULL) receives
 * ncand rows of (block, items, scalar, p50 ms). */
gdb_status gdb_autotune_phase(gdb_executor ex, gdb_plan p, const char* phase, const int* blocks, int nblocks,
                              const int* items, int nitems, const char* scalar_name, const int64_t* scalar_values,
                              int nscalars, int trials, int* best_block, int* best_items, int64_t* best_scalar,
                              float* best_ms, double* table, int* ncand);

/* ---- Batched persistent launch of many pipelines ----
 * No counterpart in the reference, which runs one command buffer per phase
 * (codegen/execution/metal_generic_executor.cpp:338-390); callers are the
 * batched workload and the request loop. Each pipeline is a plan (one or
 * more phases) with its own runner holding its registrations; build() emits
 * one persistent kernel per phase level over all of them (work items from an
 * atomic ticket; `items` is the first phase's ITEMS, later phases use their
 * defaults; work_items counts every level, grid_out is the first level's),
 * capture() records fills + kernels in a graph, run() launches (graph replay
 * when use_graph) and checks every pipeline's status. Results are collected
 * per pipeline with gdb_executor_collect on that pipeline's runner. */
typedef struct gdb_batch_s* gdb_batch;
gdb_status gdb_batch_create(gdb_ctx ctx, gdb_batch* out);
gdb_status gdb_batch_add(gdb_batch b, gdb_plan p, gdb_executor ex, int items, int* index);
gdb_status gdb_batch_build(gdb_batch b, int block, uint64_t chunk_rows, int grid, unsigned* grid_out,
                           uint64_t* work_items);
// Rows per work item of phase level `level` for the next build (0: chunk_rows
// as for the other levels), e.g. smaller items for a fold level over
// partial tables, which has few rows per pipeline.
gdb_status gdb_batch_set_level_chunk(gdb_batch b, int level, uint64_t rows);
gdb_status gdb_batch_source(gdb_batch b, char* buf, size_t cap, size_t* need);
// Source of the batched kernel for pipelines plans[i] at items[i], without a
// device (offline compilation with gdb_module_precompile, inspection).
gdb_status gdb_batch_kernel_source(const gdb_plan* plans, const int* items, int n, char* buf, size_t cap,
                                   size_t* need);
gdb_status gdb_batch_capture(gdb_batch b);
gdb_status gdb_batch_run(gdb_batch b, int use_graph, float* ms);
gdb_status gdb_batch_destroy(gdb_batch b);

/* ---- Learned group-count estimator ----
 * No counterpart in the reference, which sizes tables only by an exact
 * max-key pre-pass (codegen/execution/max_key_symbols.cpp:188-235); callers
 * are the group-by runner (table size, LDS vs global variant) and its tests.
 * Features: 16 floats from plan statistics and a row sample; target:
 * log2(groups) - log2(upper_bound), upper_bound = min(qualifying rows, key range). */
typedef struct gdb_estimator_s* gdb_estimator;
#define GDB_EST_FEATURES 16

gdb_status gdb_est_features(uint64_t n_rows, int64_t key_min, int64_t key_max, int64_t val_cut,
                            const int32_t* sample_keys, const int64_t* sample_vals, uint64_t sample_n, float* out16,
                            double* upper_bound);
/* Copy a sample of sample_n rows (8 evenly spaced chunks) of device columns to the host. */
gdb_status gdb_est_sample(gdb_ctx ctx, const void* dkeys, const void* dvals, uint64_t n, uint64_t sample_n,
                          int32_t* hkeys, int64_t* hvals);
gdb_status gdb_est_create(gdb_ctx ctx, gdb_estimator* out);
gdb_status gdb_est_destroy(gdb_estimator e);
gdb_status gdb_est_init(gdb_estimator e, uint64_t seed);
/* Train on the GPU (kernels of gdb_kernels.h); losses[epochs] optional. */
gdb_status gdb_est_train(gdb_estimator e, const float* X, const float* y, int n, int epochs, int batch, float lr,
                         uint64_t seed, float* losses);
gdb_status gdb_est_predict(gdb_estimator e, const float* X, int n, float* out);
gdb_status gdb_est_save(gdb_estimator e, const char* path);
gdb_status gdb_est_load(gdb_estimator e, const char* path);
gdb_status gdb_est_get_state(gdb_estimator e, float* params, float* mean16, float* std16);
gdb_status gdb_est_set_state(gdb_estimator e, const float* params, const float* mean16, const float* std16);
/* Log `nruns` training runs: device-generated Zipf keys (n, key range, skew,
 * selectivity drawn from the given ranges), features from a sample, and the
 * exact group count from the engine's GB_COUNT plan. spec[r*4..] = n, skew,
 * key range, selectivity. */
gdb_status gdb_est_collect_runs(gdb_ctx ctx, int nruns, uint64_t seed, uint64_t nmin, uint64_t nmax, int64_t kmin,
                                int64_t kmax, double smax, double selmin, float* X, double* ub, double* groups,
                                double* spec);

#ifdef __cplusplus
}
#endif

#endif /* GDB_ENGINE_H */

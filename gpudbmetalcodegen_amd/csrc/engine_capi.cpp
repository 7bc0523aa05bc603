// engine_capi.cpp — C-ABI wrappers of the C++ engine (include/gdb/gdb_engine.h).
// Exceptions are caught here and turned into gdb_status + a per-thread message.

#include "gdb/gdb_engine.h"

#include "execution/hip_generic_executor.h"
#include "execution/runtime_compiler.h"
#include "infra/canonical_compare.h"
#include "infra/colbin.h"
#include "infra/generators.h"
#include "planning/hip_plan_builder.h"
#include "planning/sql_frontend.h"
#include "runtime/autotuner.h"
#include "runtime/batcher.h"
#include "runtime/streaming.h"
#include "runtime/estimator.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>

using namespace gdb;
using namespace gdb::codegen;

namespace {

thread_local std::string g_engineErr;

gdb_status setErr(gdb_status code, const std::string& msg) {
    g_engineErr = msg;
    return code;
}

template <typename F>
gdb_status guard(F&& f) {
    try {
        return f();
    } catch (const GdbOverflowError& e) {
        return setErr(GDB_ERR_OVERFLOW, e.what());
    } catch (const std::exception& e) {
        const std::string m = e.what();
        if (m.find("hipRTC compile failed") != std::string::npos) return setErr(GDB_ERR_COMPILE, m);
        if (m.find("not registered") != std::string::npos || m.find("unknown") != std::string::npos)
            return setErr(GDB_ERR_NOT_FOUND, m);
        if (m.find("hipErrorOutOfMemory") != std::string::npos) return setErr(GDB_ERR_OOM, m);
        return setErr(GDB_ERR_INVALID, m);
    } catch (...) {
        return setErr(GDB_ERR_INVALID, "unknown C++ exception");
    }
}

gdb_status copyOut(const std::string& s, char* buf, size_t cap, size_t* need) {
    if (need) *need = s.size() + 1;
    if (buf && cap > 0) {
        const size_t n = std::min(cap - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
    }
    return GDB_OK;
}

// One RuntimeCompiler per context (module cache shared by all plans).
std::mutex g_rcMu;
std::map<gdb_ctx, std::unique_ptr<RuntimeCompiler>> g_compilers;

RuntimeCompiler& compilerFor(gdb_ctx ctx) {
    std::lock_guard<std::mutex> lk(g_rcMu);
    auto& rc = g_compilers[ctx];
    if (!rc) rc = std::make_unique<RuntimeCompiler>(ctx);
    return *rc;
}

}  // namespace

struct gdb_plan_s {
    HipQueryPlan plan;
    HipCodegen cg;
    std::string source;
    gdb_ctx ctx = nullptr;
    RuntimeCompiler::CompiledQuery cq;
    bool compiled = false;
    void regenerate() {
        cg = generateFromPlan(plan);
        source = cg.print();
        compiled = false;
    }
};

struct gdb_executor_s {
    std::unique_ptr<HipGenericExecutor> ex;
    gdb_ctx ctx = nullptr;
};

struct gdb_result_s {
    GenericResult r;
};

extern "C" {

const char* gdb_engine_last_error(void) { return g_engineErr.c_str(); }

// ---- generators ----
gdb_status gdb_gen_uniform_i32(uint64_t seed, uint64_t stream, uint64_t n, int64_t lo, int64_t hi, int32_t* out) {
    return guard([&]() -> gdb_status {
        infra::genUniformI32(seed, stream, n, lo, hi, out);
        return GDB_OK;
    });
}
gdb_status gdb_gen_uniform_i64(uint64_t seed, uint64_t stream, uint64_t n, int64_t lo, int64_t hi, int64_t* out) {
    return guard([&]() -> gdb_status {
        infra::genUniformI64(seed, stream, n, lo, hi, out);
        return GDB_OK;
    });
}
gdb_status gdb_gen_categorical_u8(uint64_t seed, uint64_t stream, uint64_t n, const uint8_t* values,
                                  const uint32_t* weights, int nvalues, uint8_t* out) {
    return guard([&]() -> gdb_status {
        infra::genCategoricalU8(seed, stream, n, values, weights, nvalues, out);
        return GDB_OK;
    });
}
gdb_status gdb_gen_permutation_i32(uint64_t seed, uint64_t stream, uint64_t n, int64_t base, int32_t* out) {
    return guard([&]() -> gdb_status {
        infra::genPermutationI32(seed, stream, n, base, out);
        return GDB_OK;
    });
}
gdb_status gdb_gen_fanout_total(uint64_t seed, uint64_t stream, uint64_t nparent, int cmin, int cmax,
                                uint64_t* total) {
    return guard([&]() -> gdb_status {
        *total = infra::fanoutTotal(seed, stream, nparent, cmin, cmax);
        return GDB_OK;
    });
}
gdb_status gdb_gen_fanout_fkey_i32(uint64_t seed, uint64_t stream, uint64_t nparent, int cmin, int cmax,
                                   uint64_t nchild, int32_t* out) {
    return guard([&]() -> gdb_status {
        infra::genFanoutFkeyI32(seed, stream, nparent, cmin, cmax, nchild, out);
        return GDB_OK;
    });
}
gdb_status gdb_gen_zipf_i32(uint64_t seed, uint64_t stream, uint64_t n, double s, int64_t nkeys, int32_t* out) {
    return guard([&]() -> gdb_status {
        infra::genZipfI32(seed, stream, n, s, nkeys, out);
        return GDB_OK;
    });
}
gdb_status gdb_gen_vocabulary(uint64_t seed, int nwords, int minlen, int maxlen, char* out) {
    return guard([&]() -> gdb_status {
        infra::genVocabulary(seed, nwords, minlen, maxlen, out);
        return GDB_OK;
    });
}
gdb_status gdb_gen_text(uint64_t seed, uint64_t stream, uint64_t n, int width, const char* vocab, int nwords,
                        char* out) {
    return guard([&]() -> gdb_status {
        infra::genText(seed, stream, n, width, vocab, nwords, out);
        return GDB_OK;
    });
}

// ---- plans ----
gdb_status gdb_plan_build(const char* name, const char* options, gdb_plan* out) {
    return guard([&]() -> gdb_status {
        if (!name || !out) return setErr(GDB_ERR_INVALID, "gdb_plan_build: null argument");
        auto plan = buildHipPlan(name, parsePlanOptions(options ? options : ""));
        if (!plan) return setErr(GDB_ERR_NOT_FOUND, std::string("gdb_plan_build: unknown plan ") + name);
        auto* p = new gdb_plan_s();
        p->plan = std::move(*plan);
        p->regenerate();
        *out = p;
        return GDB_OK;
    });
}

gdb_status gdb_plan_build_sql(const char* sql, const char* options, gdb_plan* out) {
    return guard([&]() -> gdb_status {
        if (!sql || !out) return setErr(GDB_ERR_INVALID, "gdb_plan_build_sql: null argument");
        auto* p = new gdb_plan_s();
        try {
            p->plan = buildSqlPlan(sql, parsePlanOptions(options ? options : ""));
            p->regenerate();
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
        return GDB_OK;
    });
}

gdb_status gdb_plan_sql_layout(gdb_plan p, char* buf, size_t cap, size_t* need) {
    return guard([&]() -> gdb_status { return copyOut(p->plan.sqlLayout, buf, cap, need); });
}

gdb_status gdb_plan_describe(gdb_plan p, char* buf, size_t cap, size_t* need) {
    return guard([&]() -> gdb_status { return copyOut(p->plan.describe(), buf, cap, need); });
}

gdb_status gdb_plan_source(gdb_plan p, char* buf, size_t cap, size_t* need) {
    return guard([&]() -> gdb_status { return copyOut(p->source, buf, cap, need); });
}

gdb_status gdb_plan_phase_count(gdb_plan p, int* n) {
    *n = static_cast<int>(p->plan.phases.size());
    return GDB_OK;
}

gdb_status gdb_plan_phase_name(gdb_plan p, int idx, char* buf, size_t cap) {
    return guard([&]() -> gdb_status {
        if (idx < 0 || idx >= static_cast<int>(p->plan.phases.size()))
            return setErr(GDB_ERR_INVALID, "gdb_plan_phase_name: index out of range");
        return copyOut(p->plan.phases[static_cast<size_t>(idx)].name, buf, cap, nullptr);
    });
}

gdb_status gdb_plan_set_launch(gdb_plan p, const char* phase, int block, int items) {
    return guard([&]() -> gdb_status {
        for (auto& ph : p->cg.getPhasesMutable()) {
            if (ph.name != phase) continue;
            bool known = false;
            for (int v : ph.itemsVariants) known |= (v == items);
            if (!known) return setErr(GDB_ERR_INVALID, "gdb_plan_set_launch: no ITEMS variant " + std::to_string(items));
            if (block < 64 || block > ph.launchBound || block % 64)
                return setErr(GDB_ERR_INVALID, "gdb_plan_set_launch: block size outside [64, launch bound]");
            ph.threadgroupSize = block;
            ph.items = items;
            for (auto& pp : p->plan.phases)
                if (pp.name == phase) {
                    pp.threadgroupSize = block;
                    pp.items = items;
                }
            return GDB_OK;
        }
        return setErr(GDB_ERR_NOT_FOUND, std::string("gdb_plan_set_launch: no phase ") + phase);
    });
}

gdb_status gdb_plan_get_launch(gdb_plan p, const char* phase, int* block, int* items) {
    for (const auto& ph : p->cg.getPhases())
        if (ph.name == phase) {
            *block = ph.threadgroupSize;
            *items = ph.items;
            return GDB_OK;
        }
    return setErr(GDB_ERR_NOT_FOUND, std::string("gdb_plan_get_launch: no phase ") + phase);
}

gdb_status gdb_plan_compile(gdb_ctx ctx, gdb_plan p, double* ms, int* cache_hit, uint64_t* plan_hash) {
    return guard([&]() -> gdb_status {
        auto& rc = compilerFor(ctx);
        p->cq = rc.compile(p->source);
        p->ctx = ctx;
        p->compiled = true;
        if (ms) *ms = p->cq.compileMs;
        if (cache_hit) *cache_hit = p->cq.cacheHit ? 1 : 0;
        if (plan_hash) *plan_hash = p->cq.planHash;
        return GDB_OK;
    });
}

gdb_status gdb_plan_precompile(gdb_plan p, const char* cache_dir, const char* arch, int* was_cached,
                               uint64_t* plan_hash) {
    gdb_status st = gdb_module_precompile(cache_dir, arch, p->source.data(), p->source.size(), nullptr, 0, plan_hash,
                                          was_cached);
    if (st != GDB_OK) return setErr(st, gdb_last_error());
    return GDB_OK;
}

gdb_status gdb_plan_free(gdb_plan p) {
    delete p;
    return GDB_OK;
}

// ---- executor ----
gdb_status gdb_executor_create(gdb_ctx ctx, gdb_stream s, gdb_executor* out) {
    return guard([&]() -> gdb_status {
        auto* e = new gdb_executor_s();
        e->ctx = ctx;
        e->ex = std::make_unique<HipGenericExecutor>(ctx, s);
        *out = e;
        return GDB_OK;
    });
}

gdb_status gdb_executor_destroy(gdb_executor ex) {
    return guard([&]() -> gdb_status {
        delete ex;
        return GDB_OK;
    });
}

gdb_status gdb_executor_register_table(gdb_executor ex, const char* table, uint64_t rows) {
    return guard([&]() -> gdb_status {
        ex->ex->registerTable(table, rows);
        return GDB_OK;
    });
}

gdb_status gdb_executor_register_column(gdb_executor ex, const char* table, const char* column, const void* dptr) {
    return guard([&]() -> gdb_status {
        ex->ex->registerTableColumn(table, column, dptr);
        return GDB_OK;
    });
}

gdb_status gdb_executor_register_buffer(gdb_executor ex, const char* name, void* dptr, size_t bytes) {
    return guard([&]() -> gdb_status {
        ex->ex->registerBuffer(name, dptr, bytes);
        return GDB_OK;
    });
}

gdb_status gdb_executor_register_symbol(gdb_executor ex, const char* name, uint64_t value) {
    return guard([&]() -> gdb_status {
        ex->ex->registerSymbol(name, value);
        return GDB_OK;
    });
}

gdb_status gdb_executor_get_symbol(gdb_executor ex, const char* name, uint64_t* value, int* found) {
    return guard([&]() -> gdb_status {
        uint64_t v = 0;
        const bool f = ex->ex->tryGetSymbol(name, v);
        if (value) *value = v;
        if (found) *found = f ? 1 : 0;
        return GDB_OK;
    });
}

gdb_status gdb_executor_register_scalar(gdb_executor ex, const char* name, int64_t value) {
    return guard([&]() -> gdb_status {
        ex->ex->registerScalarInt(name, value);
        return GDB_OK;
    });
}

gdb_status gdb_executor_set_skip_init(gdb_executor ex, int skip) {
    ex->ex->setSkipInit(skip != 0);
    return GDB_OK;
}

gdb_status gdb_executor_set_phase_timing(gdb_executor ex, int on) {
    ex->ex->setPhaseTiming(on != 0);
    return GDB_OK;
}

static void registerPlanTables(gdb_executor ex, gdb_plan p) {
    for (const auto& bt : p->plan.bufferTables) ex->ex->registerBufferTable(bt.table, bt.column, bt.buffer, bt.rowsExpr);
}

// Install the plan's sizing pre-pass rules as phase hooks on the runner.
static void installSizingRules(gdb_executor ex, gdb_plan p) {
    ex->ex->clearPhaseHooks();
    ex->ex->clearLateSymbols();
    for (const auto& rule : p->plan.sizingRules) {
        ex->ex->addLateSymbol(rule.symbol);
        ex->ex->setPhaseHook(
            rule.afterPhase,
            [rule](HipGenericExecutor& e) {
            uint64_t count = 0;
            std::memcpy(&count, e.hookBytes(), 8);  // read behind the phase (one synchronisation)
            uint64_t v = std::max<uint64_t>(rule.minimum, count * rule.multiplier);
            if (rule.pow2) {
                uint64_t q = 1;
                while (q < v) q <<= 1;
                v = q;
            }
            e.registerSymbol(rule.symbol, v);
            e.registerSymbol(rule.symbol + "_count", count);
            },
            rule.buffer, static_cast<size_t>(rule.word) * 8, 8);
    }
}

static void fillTiming(const HipExecutionResult& r, gdb_exec_timing* out) {
    if (!out) return;
    std::memset(out, 0, sizeof(*out));
    out->total_ms = r.totalKernelTimeMs;
    out->alloc_ms = r.bufferAllocTimeMs;
    out->nphases = static_cast<int>(std::min<size_t>(r.phaseTimesMs.size(), GDB_MAX_PHASES));
    for (int i = 0; i < out->nphases; ++i) {
        out->phase_ms[i] = r.phaseTimesMs[static_cast<size_t>(i)];
        if (static_cast<size_t>(i) < r.phaseGrid.size()) {
            out->grid[i] = r.phaseGrid[static_cast<size_t>(i)];
            out->block[i] = r.phaseBlock[static_cast<size_t>(i)];
            out->items[i] = r.phaseItems[static_cast<size_t>(i)];
        }
        if (static_cast<size_t>(i) < r.phaseStatus.size())
            for (int w = 0; w < 4; ++w) out->status[i][w] = r.phaseStatus[static_cast<size_t>(i)][static_cast<size_t>(w)];
    }
}

gdb_status gdb_executor_execute_range(gdb_executor ex, gdb_plan p, int first, int last, int warmup, int measured,
                                      gdb_exec_timing* out) {
    return guard([&]() -> gdb_status {
        if (!p->compiled || p->ctx != ex->ctx) return setErr(GDB_ERR_INVALID, "gdb_executor_execute: plan not compiled for this context");
        registerPlanTables(ex, p);
        installSizingRules(ex, p);
        auto& rc = compilerFor(ex->ctx);
        HipExecutionResult r;
        try {
            r = ex->ex->execute(rc, p->cq, p->cg, warmup, measured, first, last);
        } catch (const GdbOverflowError& e) {
            if (out) {
                std::memset(out, 0, sizeof(*out));
                out->status[0][0] = 1;
                out->status[0][2] = e.inserted;
            }
            throw;
        }
        fillTiming(r, out);
        return GDB_OK;
    });
}

gdb_status gdb_executor_execute(gdb_executor ex, gdb_plan p, int warmup, int measured, gdb_exec_timing* out) {
    return gdb_executor_execute_range(ex, p, 0, -1, warmup, measured, out);
}

gdb_status gdb_executor_execute_streamed(gdb_executor ex, gdb_plan p, const char* table, uint64_t rows, int ncols,
                                         const char* const* names, const void* const* host_cols,
                                         const int* elem_bytes, uint64_t chunk_rows, gdb_chunk_stats* out) {
    return guard([&]() -> gdb_status {
        if (!p->compiled || p->ctx != ex->ctx)
            return setErr(GDB_ERR_INVALID, "gdb_executor_execute_streamed: plan not compiled for this context");
        if (!table || ncols < 1 || !names || !host_cols || !elem_bytes)
            return setErr(GDB_ERR_INVALID, "gdb_executor_execute_streamed: no stream columns");
        std::vector<runtime::StreamColumn> cols;
        for (int i = 0; i < ncols; ++i) cols.push_back({names[i], host_cols[i], elem_bytes[i]});
        ex->ex->registerTable(table, rows);  // row count known to the plan before the pre-stream phases
        registerPlanTables(ex, p);
        installSizingRules(ex, p);
        runtime::ChunkStreamer cs(ex->ctx, *ex->ex, p->plan, p->cg, compilerFor(ex->ctx), p->cq);
        const auto st = cs.run(table, rows, cols, chunk_rows);
        if (out) {
            out->total_ms = st.totalMs;
            out->stream_ms = st.streamMs;
            out->chunks = st.chunks;
            out->bytes_copied = st.bytesCopied;
            out->first_stream_phase = st.firstStreamPhase;
            out->last_stream_phase = st.lastStreamPhase;
        }
        return GDB_OK;
    });
}

namespace {

// Rows [r0, r0 + n) of one .colbin column in the engine's format: INT/DATE
// int32 and CHAR1/CHAR_FIXED bytes as stored, FLOAT (float32 in the file)
// as exact int64 hundredths rounded half to even (the loader's rule,
// to_engine_format in the Python ingest). Large chunks are split over threads.
void convertColbinRows(const infra::ColbinColumn& fc, void* dst, uint64_t r0, uint64_t n) {
    const size_t fileEb = infra::elemBytes(fc.type, fc.width);
    auto part = [&](uint64_t a, uint64_t b) {
        if (fc.type == infra::TpchType::FLOAT) {
            const float* src = reinterpret_cast<const float*>(fc.data) + r0;
            int64_t* out = static_cast<int64_t*>(dst);
            for (uint64_t i = a; i < b; ++i) out[i] = static_cast<int64_t>(std::nearbyint(static_cast<double>(src[i]) * 100.0));
        } else {
            std::memcpy(static_cast<uint8_t*>(dst) + a * fileEb, fc.data + (r0 + a) * fileEb, (b - a) * fileEb);
        }
    };
    const uint64_t kPar = 1u << 20;
    if (n < 2 * kPar) {
        part(0, n);
        return;
    }
    const unsigned t = std::min<unsigned>(8, std::max(1u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    const uint64_t step = (n + t - 1) / t;
    for (unsigned k = 0; k < t; ++k) {
        const uint64_t a = k * step, b = std::min<uint64_t>(n, a + step);
        if (a < b) th.emplace_back(part, a, b);
    }
    for (auto& x : th) x.join();
}

}  // namespace

gdb_status gdb_executor_execute_streamed_colbin(gdb_executor ex, gdb_plan p, const char* table, const char* colbin_path,
                                                uint64_t chunk_rows, gdb_chunk_stats* out) {
    return guard([&]() -> gdb_status {
        if (!p->compiled || p->ctx != ex->ctx)
            return setErr(GDB_ERR_INVALID, "gdb_executor_execute_streamed_colbin: plan not compiled for this context");
        if (!table || !colbin_path)
            return setErr(GDB_ERR_INVALID, "gdb_executor_execute_streamed_colbin: no table or file");
        const infra::TpchTable& tdef = infra::tpchTable(table);
        infra::ColbinFile file(colbin_path, "");
        // the columns the stream phases read, each converted chunk by chunk
        std::set<std::string> names;
        for (const auto& ph : p->cg.getPhases())
            if (ph.scannedTable == table)
                for (const auto& b : ph.bindings)
                    if (b.kind == codegen::HipParamKind::TableColumn && b.tableName == table) names.insert(b.columnName);
        std::vector<runtime::StreamColumn> cols;
        for (const auto& name : names) {
            const infra::TpchColumn& c = tdef.column(name);
            const infra::ColbinColumn& fc = file.column(c.index);
            if (fc.type != c.type || (c.type == infra::TpchType::CHAR_FIXED && fc.width != c.width))
                return setErr(GDB_ERR_INVALID, std::string("gdb_executor_execute_streamed_colbin: column ") + name +
                                                   " has another type or width in " + colbin_path);
            if (fc.bytes < file.rows() * infra::elemBytes(fc.type, fc.width))
                return setErr(GDB_ERR_INVALID, std::string("gdb_executor_execute_streamed_colbin: column ") + name +
                                                   " is shorter than the file's row count");
            runtime::StreamColumn sc;
            sc.name = name;
            sc.elemBytes = c.type == infra::TpchType::FLOAT ? 8 : static_cast<int>(infra::elemBytes(c.type, c.width));
            const infra::ColbinColumn* fcp = &fc;
            sc.produce = [fcp](void* dst, uint64_t r0, uint64_t n) { convertColbinRows(*fcp, dst, r0, n); };
            cols.push_back(std::move(sc));
        }
        if (cols.empty())
            return setErr(GDB_ERR_INVALID, std::string("gdb_executor_execute_streamed_colbin: no phase reads ") + table);
        ex->ex->registerTable(table, file.rows());
        registerPlanTables(ex, p);
        installSizingRules(ex, p);
        runtime::ChunkStreamer cs(ex->ctx, *ex->ex, p->plan, p->cg, compilerFor(ex->ctx), p->cq);
        const auto st = cs.run(table, file.rows(), cols, chunk_rows);
        if (out) {
            out->total_ms = st.totalMs;
            out->stream_ms = st.streamMs;
            out->chunks = st.chunks;
            out->bytes_copied = st.bytesCopied;
            out->first_stream_phase = st.firstStreamPhase;
            out->last_stream_phase = st.lastStreamPhase;
        }
        return GDB_OK;
    });
}

gdb_status gdb_plan_chunkable(gdb_plan p, int* chunkable) {
    return guard([&]() -> gdb_status {
        *chunkable = p->plan.chunkable ? 1 : 0;
        return GDB_OK;
    });
}

gdb_status gdb_executor_prepare(gdb_executor ex, gdb_plan p) {
    return guard([&]() -> gdb_status {
        registerPlanTables(ex, p);
        ex->ex->prepare(p->cg);
        return GDB_OK;
    });
}

gdb_status gdb_executor_enqueue(gdb_executor ex, gdb_plan p) {
    return guard([&]() -> gdb_status {
        if (!p->compiled || p->ctx != ex->ctx) return setErr(GDB_ERR_INVALID, "gdb_executor_enqueue: plan not compiled for this context");
        registerPlanTables(ex, p);
        ex->ex->enqueue(compilerFor(ex->ctx), p->cq, p->cg);
        return GDB_OK;
    });
}

gdb_status gdb_executor_check_status(gdb_executor ex, gdb_plan p) {
    return guard([&]() -> gdb_status {
        ex->ex->checkStatus(p->cg);
        return GDB_OK;
    });
}

gdb_status gdb_executor_read_buffer(gdb_executor ex, const char* name, size_t offset, void* host, size_t bytes) {
    return guard([&]() -> gdb_status {
        ex->ex->readBuffer(name, offset, host, bytes);
        return GDB_OK;
    });
}

gdb_status gdb_executor_buffer_ptr(gdb_executor ex, const char* name, void** dptr, size_t* bytes) {
    void* ptr = ex->ex->getAllocatedBuffer(name);
    if (!ptr) return setErr(GDB_ERR_NOT_FOUND, std::string("no buffer ") + name);
    if (dptr) *dptr = ptr;
    if (bytes) *bytes = ex->ex->getAllocatedBytes(name);
    return GDB_OK;
}

gdb_status gdb_executor_owned_bytes(gdb_executor ex, size_t* bytes) {
    *bytes = ex->ex->ownedBytes();
    return GDB_OK;
}

gdb_status gdb_executor_release(gdb_executor ex) {
    return guard([&]() -> gdb_status {
        ex->ex->releaseAllocatedBuffers();
        return GDB_OK;
    });
}

// ---- results ----
gdb_status gdb_executor_collect(gdb_executor ex, gdb_plan p, gdb_result* out) {
    return guard([&]() -> gdb_status {
        auto* r = new gdb_result_s();
        r->r = ex->ex->collectResult(p->cg);
        *out = r;
        return GDB_OK;
    });
}

gdb_status gdb_result_shape(gdb_result r, uint64_t* rows, int* cols) {
    if (rows) *rows = r->r.rows();
    if (cols) *cols = static_cast<int>(r->r.columns.size());
    return GDB_OK;
}

gdb_status gdb_result_column(gdb_result r, int idx, char* name, size_t cap, int* type) {
    if (idx < 0 || idx >= static_cast<int>(r->r.columns.size()))
        return setErr(GDB_ERR_INVALID, "gdb_result_column: index out of range");
    const auto& c = r->r.columns[static_cast<size_t>(idx)];
    copyOut(c.name, name, cap, nullptr);
    if (type) *type = static_cast<int>(c.type);
    return GDB_OK;
}

gdb_status gdb_result_column_format(gdb_result r, int idx, int* scale, int* is_date) {
    if (idx < 0 || idx >= static_cast<int>(r->r.columns.size()))
        return setErr(GDB_ERR_INVALID, "gdb_result_column_format: index out of range");
    const auto& c = r->r.columns[static_cast<size_t>(idx)];
    if (scale) *scale = c.scale;
    if (is_date) *is_date = c.date ? 1 : 0;
    return GDB_OK;
}

gdb_status gdb_result_i64(gdb_result r, int idx, int64_t* out) {
    const auto& c = r->r.columns.at(static_cast<size_t>(idx));
    if (c.type != ResultColumn::I64) return setErr(GDB_ERR_INVALID, "gdb_result_i64: column is not int64");
    std::memcpy(out, c.i64.data(), c.i64.size() * 8);
    return GDB_OK;
}

gdb_status gdb_result_i128(gdb_result r, int idx, int64_t* out) {
    const auto& c = r->r.columns.at(static_cast<size_t>(idx));
    if (c.type != ResultColumn::I128) return setErr(GDB_ERR_INVALID, "gdb_result_i128: column is not int128");
    for (size_t i = 0; i < c.i128.size(); ++i) {
        const unsigned __int128 u = static_cast<unsigned __int128>(c.i128[i]);
        out[2 * i] = static_cast<int64_t>(static_cast<uint64_t>(u));
        out[2 * i + 1] = static_cast<int64_t>(static_cast<uint64_t>(u >> 64));
    }
    return GDB_OK;
}

gdb_status gdb_result_str(gdb_result r, int idx, char* out, size_t width) {
    const auto& c = r->r.columns.at(static_cast<size_t>(idx));
    if (c.type != ResultColumn::STR) return setErr(GDB_ERR_INVALID, "gdb_result_str: column is not a string");
    for (size_t i = 0; i < c.str.size(); ++i) {
        std::memset(out + i * width, 0, width);
        std::memcpy(out + i * width, c.str[i].data(), std::min(width, c.str[i].size()));
    }
    return GDB_OK;
}

gdb_status gdb_result_canonical(gdb_result r, char* buf, size_t cap, size_t* need) {
    return guard([&]() -> gdb_status { return copyOut(r->r.toCanonical(), buf, cap, need); });
}

gdb_status gdb_result_free(gdb_result r) {
    delete r;
    return GDB_OK;
}

// ---- autotuner ----
gdb_status gdb_autotune_phase(gdb_executor ex, gdb_plan p, const char* phase, const int* blocks, int nblocks,
                              const int* items, int nitems, const char* scalar_name, const int64_t* scalar_values,
                              int nscalars, int trials, int* best_block, int* best_items, int64_t* best_scalar,
                              float* best_ms, double* table, int* ncand) {
    return guard([&]() -> gdb_status {
        if (!p->compiled || p->ctx != ex->ctx) return setErr(GDB_ERR_INVALID, "gdb_autotune_phase: plan not compiled");
        int idx = -1;
        const auto& phs = p->cg.getPhases();
        for (size_t i = 0; i < phs.size(); ++i)
            if (phs[i].name == phase) idx = static_cast<int>(i);
        if (idx < 0) return setErr(GDB_ERR_NOT_FOUND, std::string("gdb_autotune_phase: no phase ") + phase);
        registerPlanTables(ex, p);
        ex->ex->clearPhaseHooks();
        auto r = runtime::autotunePhase(*ex->ex, compilerFor(ex->ctx), p->cq, p->cg, idx,
                                        std::vector<int>(blocks, blocks + nblocks), std::vector<int>(items, items + nitems),
                                        scalar_name ? scalar_name : "",
                                        std::vector<int64_t>(scalar_values, scalar_values + (scalar_values ? nscalars : 0)),
                                        trials);
        if (r.block == 0) return setErr(GDB_ERR_INVALID, "gdb_autotune_phase: no candidate ran");
        for (auto& pp : p->plan.phases)
            if (pp.name == phase) {
                pp.threadgroupSize = r.block;
                pp.items = r.items;
            }
        if (best_block) *best_block = r.block;
        if (best_items) *best_items = r.items;
        if (best_scalar) *best_scalar = r.scalar;
        if (best_ms) *best_ms = r.ms;
        if (table)
            for (size_t i = 0; i < r.table.size(); ++i)
                for (int k = 0; k < 4; ++k) table[i * 4 + static_cast<size_t>(k)] = r.table[i][static_cast<size_t>(k)];
        if (ncand) *ncand = static_cast<int>(r.table.size());
        return GDB_OK;
    });
}

// ---- batch ----
gdb_status gdb_batch_create(gdb_ctx ctx, gdb_batch* out) {
    return guard([&]() -> gdb_status {
        *out = reinterpret_cast<gdb_batch>(new runtime::PipelineBatch(ctx));
        return GDB_OK;
    });
}

static runtime::PipelineBatch& bat(gdb_batch b) { return *reinterpret_cast<runtime::PipelineBatch*>(b); }

gdb_status gdb_batch_add(gdb_batch b, gdb_plan p, gdb_executor ex, int items, int* index) {
    return guard([&]() -> gdb_status {
        const int i = bat(b).add(&p->plan, &p->cg, ex->ex.get(), items);
        if (index) *index = i;
        return GDB_OK;
    });
}

gdb_status gdb_batch_build(gdb_batch b, int block, uint64_t chunk_rows, int grid, unsigned* grid_out,
                           uint64_t* work_items) {
    return guard([&]() -> gdb_status {
        bat(b).build(block, chunk_rows, grid);
        if (grid_out) *grid_out = bat(b).grid();
        if (work_items) *work_items = bat(b).workItems();
        return GDB_OK;
    });
}

gdb_status gdb_batch_set_level_chunk(gdb_batch b, int level, uint64_t rows) {
    return guard([&]() -> gdb_status {
        bat(b).setLevelChunk(level, rows);
        return GDB_OK;
    });
}

gdb_status gdb_batch_source(gdb_batch b, char* buf, size_t cap, size_t* need) {
    return guard([&]() -> gdb_status { return copyOut(bat(b).source(), buf, cap, need); });
}

gdb_status gdb_batch_kernel_source(const gdb_plan* plans, const int* items, int n, char* buf, size_t cap,
                                   size_t* need) {
    return guard([&]() -> gdb_status {
        if (n < 1 || !plans || !items) return setErr(GDB_ERR_INVALID, "gdb_batch_kernel_source: no pipelines");
        runtime::BatchKernelSource bs;
        for (int i = 0; i < n; ++i) bs.add(&plans[i]->plan, &plans[i]->cg, items[i]);
        return copyOut(bs.print(), buf, cap, need);
    });
}

gdb_status gdb_batch_capture(gdb_batch b) {
    return guard([&]() -> gdb_status {
        bat(b).capture();
        return GDB_OK;
    });
}

gdb_status gdb_batch_run(gdb_batch b, int use_graph, float* ms) {
    return guard([&]() -> gdb_status {
        const float t = bat(b).run(use_graph != 0);
        if (ms) *ms = t;
        return GDB_OK;
    });
}

gdb_status gdb_batch_destroy(gdb_batch b) {
    delete reinterpret_cast<runtime::PipelineBatch*>(b);
    return GDB_OK;
}

// ---- estimator ----
gdb_status gdb_est_features(uint64_t n_rows, int64_t key_min, int64_t key_max, int64_t val_cut,
                            const int32_t* sample_keys, const int64_t* sample_vals, uint64_t sample_n, float* out16,
                            double* upper_bound) {
    return guard([&]() -> gdb_status {
        runtime::FeatureInput fi;
        fi.nRows = n_rows;
        fi.keyMin = key_min;
        fi.keyMax = key_max;
        fi.valCut = val_cut;
        fi.sampleKeys = sample_keys;
        fi.sampleVals = sample_vals;
        fi.sampleSize = sample_n;
        auto f = runtime::computeFeatures(fi);
        std::memcpy(out16, f.x.data(), sizeof(float) * runtime::kFeatures);
        if (upper_bound) *upper_bound = f.upperBound;
        return GDB_OK;
    });
}

gdb_status gdb_est_sample(gdb_ctx ctx, const void* dkeys, const void* dvals, uint64_t n, uint64_t sample_n,
                          int32_t* hkeys, int64_t* hvals) {
    return guard([&]() -> gdb_status {
        runtime::sampleDeviceColumns(ctx, gdb_default_stream(ctx), static_cast<const int32_t*>(dkeys),
                                     static_cast<const int64_t*>(dvals), n, sample_n, hkeys, hvals);
        return GDB_OK;
    });
}

gdb_status gdb_est_create(gdb_ctx ctx, gdb_estimator* out) {
    return guard([&]() -> gdb_status {
        *out = reinterpret_cast<gdb_estimator>(new runtime::CardinalityEstimator(ctx));
        return GDB_OK;
    });
}

static runtime::CardinalityEstimator& est(gdb_estimator e) { return *reinterpret_cast<runtime::CardinalityEstimator*>(e); }

gdb_status gdb_est_destroy(gdb_estimator e) {
    delete reinterpret_cast<runtime::CardinalityEstimator*>(e);
    return GDB_OK;
}

gdb_status gdb_est_init(gdb_estimator e, uint64_t seed) {
    return guard([&]() -> gdb_status {
        est(e).initWeights(seed);
        return GDB_OK;
    });
}

gdb_status gdb_est_train(gdb_estimator e, const float* X, const float* y, int n, int epochs, int batch, float lr,
                         uint64_t seed, float* losses) {
    return guard([&]() -> gdb_status {
        runtime::TrainConfig cfg;
        cfg.epochs = epochs;
        cfg.batch = batch;
        cfg.lr = lr;
        cfg.seed = seed;
        std::vector<float> l;
        est(e).train(X, y, n, cfg, losses ? &l : nullptr);
        if (losses) std::memcpy(losses, l.data(), l.size() * sizeof(float));
        return GDB_OK;
    });
}

gdb_status gdb_est_predict(gdb_estimator e, const float* X, int n, float* out) {
    return guard([&]() -> gdb_status {
        est(e).predict(X, n, out);
        return GDB_OK;
    });
}

gdb_status gdb_est_save(gdb_estimator e, const char* path) {
    return guard([&]() -> gdb_status {
        est(e).save(path);
        return GDB_OK;
    });
}

gdb_status gdb_est_load(gdb_estimator e, const char* path) {
    return guard([&]() -> gdb_status {
        est(e).load(path);
        return GDB_OK;
    });
}

gdb_status gdb_est_get_state(gdb_estimator e, float* params, float* mean16, float* std16) {
    auto& x = est(e);
    if (params) std::memcpy(params, x.params().data(), x.params().size() * sizeof(float));
    if (mean16) std::memcpy(mean16, x.mean().data(), sizeof(float) * runtime::kFeatures);
    if (std16) std::memcpy(std16, x.stdev().data(), sizeof(float) * runtime::kFeatures);
    return GDB_OK;
}

gdb_status gdb_est_set_state(gdb_estimator e, const float* params, const float* mean16, const float* std16) {
    auto& x = est(e);
    if (params) std::memcpy(x.params().data(), params, x.params().size() * sizeof(float));
    if (mean16) std::memcpy(x.mean().data(), mean16, sizeof(float) * runtime::kFeatures);
    if (std16) std::memcpy(x.stdev().data(), std16, sizeof(float) * runtime::kFeatures);
    return GDB_OK;
}

gdb_status gdb_est_collect_runs(gdb_ctx ctx, int nruns, uint64_t seed, uint64_t nmin, uint64_t nmax, int64_t kmin,
                                int64_t kmax, double smax, double selmin, float* X, double* ub, double* groups,
                                double* spec) {
    return guard([&]() -> gdb_status {
        auto logs = runtime::collectRuns(ctx, nruns, seed, nmin, nmax, kmin, kmax, smax, selmin);
        for (size_t r = 0; r < logs.size(); ++r) {
            std::memcpy(X + r * runtime::kFeatures, logs[r].f.x.data(), sizeof(float) * runtime::kFeatures);
            ub[r] = logs[r].f.upperBound;
            groups[r] = logs[r].groups;
            spec[r * 4 + 0] = static_cast<double>(logs[r].spec.n);
            spec[r * 4 + 1] = logs[r].spec.skew;
            spec[r * 4 + 2] = static_cast<double>(logs[r].spec.nkeys);
            spec[r * 4 + 3] = logs[r].spec.sel;
        }
        return GDB_OK;
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// TPC-H ingest and the golden check
// ---------------------------------------------------------------------------

struct gdb_colbin_s {
    std::unique_ptr<infra::ColbinFile> f;
};

gdb_status gdb_tpch_column(const char* table, const char* column, int* index, int* dtype, int* width) {
    return guard([&]() -> gdb_status {
        if (!table || !column) return setErr(GDB_ERR_INVALID, "gdb_tpch_column: null name");
        const auto& c = infra::tpchTable(table).column(column);
        if (index) *index = c.index;
        if (dtype) *dtype = static_cast<int>(c.type);
        if (width) *width = c.width;
        return GDB_OK;
    });
}

gdb_status gdb_tpch_table_columns(const char* table, char* buf, size_t cap, size_t* need) {
    return guard([&]() -> gdb_status {
        if (!table) return setErr(GDB_ERR_INVALID, "gdb_tpch_table_columns: null table");
        std::string s;
        for (const auto& c : infra::tpchTable(table).columns) s += (s.empty() ? "" : ",") + c.name;
        return copyOut(s, buf, cap, need);
    });
}

gdb_status gdb_tbl_to_colbin(const char* tbl_path, const char* table, int force, int threads, int* built) {
    return guard([&]() -> gdb_status {
        if (!tbl_path || !table) return setErr(GDB_ERR_INVALID, "gdb_tbl_to_colbin: null argument");
        const int t = threads > 0 ? threads : static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
        const bool b = infra::tblToColbin(tbl_path, table, force != 0, t);
        if (built) *built = b ? 1 : 0;
        return GDB_OK;
    });
}

gdb_status gdb_colbin_open(const char* colbin_path, const char* tbl_path, gdb_colbin* out) {
    return guard([&]() -> gdb_status {
        if (!colbin_path || !out) return setErr(GDB_ERR_INVALID, "gdb_colbin_open: null argument");
        auto h = std::make_unique<gdb_colbin_s>();
        h->f = std::make_unique<infra::ColbinFile>(colbin_path, tbl_path ? tbl_path : "");
        *out = h.release();
        return GDB_OK;
    });
}

gdb_status gdb_colbin_shape(gdb_colbin f, uint64_t* rows, int* ncols) {
    if (!f) return setErr(GDB_ERR_INVALID, "gdb_colbin_shape: null handle");
    if (rows) *rows = f->f->rows();
    if (ncols) *ncols = static_cast<int>(f->f->columns().size());
    return GDB_OK;
}

gdb_status gdb_colbin_column(gdb_colbin f, int column_index, const void** data, uint64_t* bytes, int* dtype,
                             int* width) {
    return guard([&]() -> gdb_status {
        if (!f) return setErr(GDB_ERR_INVALID, "gdb_colbin_column: null handle");
        const auto& c = f->f->column(column_index);
        if (data) *data = c.data;
        if (bytes) *bytes = c.bytes;
        if (dtype) *dtype = static_cast<int>(c.type);
        if (width) *width = c.width;
        return GDB_OK;
    });
}

void gdb_colbin_close(gdb_colbin f) { delete f; }

gdb_status gdb_compare_canonical(const char* got, const char* expected, double abs_tol, double rel_tol, char* msg,
                                 size_t cap, size_t* need) {
    return guard([&]() -> gdb_status {
        if (!got || !expected) return setErr(GDB_ERR_INVALID, "gdb_compare_canonical: null text");
        return copyOut(infra::compareCanonical(got, expected, abs_tol, rel_tol), msg, cap, need);
    });
}

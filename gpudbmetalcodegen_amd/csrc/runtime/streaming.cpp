#include "streaming.h"

#include <algorithm>
#include <chrono>
#include <set>
#include <stdexcept>

namespace gdb {
namespace runtime {

using namespace codegen;

namespace {

void check(gdb_status st, const std::string& what) {
    if (st != GDB_OK) throw std::runtime_error("ChunkStreamer: " + what + ": " + gdb_last_error());
}

}  // namespace

ChunkStreamer::ChunkStreamer(gdb_ctx ctx, HipGenericExecutor& ex, const HipQueryPlan& plan, const HipCodegen& cg,
                             RuntimeCompiler& rc, const RuntimeCompiler::CompiledQuery& cq)
    : ctx_(ctx), ex_(ex), plan_(plan), cg_(cg), rc_(rc), cq_(cq) {
    check(gdb_stream_create(ctx_, &copy_), "gdb_stream_create");
    for (int s = 0; s < 2; ++s) {
        check(gdb_event_create(&copied_[s]), "gdb_event_create");
        check(gdb_event_create(&consumed_[s]), "gdb_event_create");
    }
    check(gdb_event_create(&t0_), "gdb_event_create");
    check(gdb_event_create(&t1_), "gdb_event_create");
}

ChunkStreamer::~ChunkStreamer() {
    if (copy_) gdb_stream_sync(copy_);
    freeSlots();
    for (int s = 0; s < 2; ++s) {
        gdb_event_destroy(copied_[s]);
        gdb_event_destroy(consumed_[s]);
    }
    gdb_event_destroy(t0_);
    gdb_event_destroy(t1_);
    if (copy_) gdb_stream_destroy(copy_);
}

void ChunkStreamer::freeSlots() {
    for (auto& v : slots_) {
        for (void* p : v)
            if (p) gdb_free(ctx_, p);
        v.clear();
    }
    for (auto& v : staging_) {
        for (void* p : v)
            if (p) gdb_host_free_pinned(p);
        v.clear();
    }
}

StreamStats ChunkStreamer::run(const std::string& table, uint64_t rows, const std::vector<StreamColumn>& cols,
                               uint64_t chunkRows) {
    if (!plan_.chunkable) throw std::runtime_error("ChunkStreamer: plan " + plan_.name + " is not chunkable");
    if (chunkRows == 0) throw std::runtime_error("ChunkStreamer: chunk_rows must be > 0");
    const auto& phases = cg_.getPhases();
    const int np = static_cast<int>(phases.size());
    if (plan_.phases.size() != phases.size()) throw std::runtime_error("ChunkStreamer: plan and codegen phases differ");
    // Stream passes: runs of consecutive phases that scan the table; a phase
    // marked newStreamPass starts a pass of its own (it reads what the
    // phases before it accumulated over every chunk). Phases between passes
    // run once.
    std::vector<std::pair<int, int>> passes;
    for (int i = 0; i < np; ++i) {
        if (phases[static_cast<size_t>(i)].scannedTable != table) continue;
        if (passes.empty() || passes.back().second != i || plan_.phases[static_cast<size_t>(i)].newStreamPass)
            passes.push_back({i, i + 1});
        else
            passes.back().second = i + 1;
    }
    if (passes.empty()) throw std::runtime_error("ChunkStreamer: no phase of " + plan_.name + " scans table " + table);
    const int first = passes.front().first, last = passes.back().second;
    for (const auto& r : plan_.sizingRules)
        for (const auto& ps : passes)
            for (int i = ps.first; i < ps.second; ++i)
                if (r.afterPhase == phases[static_cast<size_t>(i)].name)
                    throw std::runtime_error("ChunkStreamer: phase " + r.afterPhase +
                                             " sizes a later phase and needs the whole table");
    // Columns each pass reads; every one must be given.
    std::vector<std::vector<size_t>> passCols(passes.size());
    for (size_t p = 0; p < passes.size(); ++p) {
        std::set<std::string> needed;
        for (int i = passes[p].first; i < passes[p].second; ++i)
            for (const auto& b : phases[static_cast<size_t>(i)].bindings)
                if (b.kind == HipParamKind::TableColumn && b.tableName == table) needed.insert(b.columnName);
        for (size_t k = 0; k < cols.size(); ++k)
            if (needed.erase(cols[k].name)) passCols[p].push_back(k);
        if (!needed.empty())
            throw std::runtime_error("ChunkStreamer: column " + *needed.begin() + " of " + table + " not given");
    }
    for (const auto& c : cols)
        if ((!c.host && !c.produce) || c.elemBytes <= 0)
            throw std::runtime_error("ChunkStreamer: column " + c.name + " has no data");

    StreamStats st;
    st.firstStreamPhase = first;
    st.lastStreamPhase = last;
    st.passes = static_cast<int>(passes.size());
    const auto wall0 = std::chrono::steady_clock::now();
    gdb_stream comp = ex_.stream();

    // Pre-stream phases once (hooks and sizing rules run as usual).
    if (first > 0) ex_.execute(rc_, cq_, cg_, 0, 1, 0, first);

    // Device slots: two per column, chunkRows rows each.
    freeSlots();
    const uint64_t cr = std::min<uint64_t>(chunkRows, std::max<uint64_t>(rows, 1));
    for (int s = 0; s < 2; ++s)
        for (const auto& c : cols) {
            const size_t bytes = static_cast<size_t>(cr) * static_cast<size_t>(c.elemBytes);
            void* d = nullptr;
            check(gdb_alloc(ctx_, bytes, &d), "gdb_alloc");
            slots_[s].push_back(d);
            void* h = nullptr;
            if (c.produce) check(gdb_host_alloc_pinned(bytes, &h), "gdb_host_alloc_pinned");
            staging_[s].push_back(h);
        }

    check(gdb_event_record(t0_, comp), "event");
    check(gdb_stream_wait_event(copy_, t0_), "wait");  // copies start after the pre-stream phases
    const uint64_t nChunks = rows == 0 ? 1 : (rows + cr - 1) / cr;
    uint64_t issued = 0;  // chunks issued over all passes: slot = issued & 1
    int cur = first;
    try {
        for (size_t p = 0; p < passes.size(); ++p) {
            const int p0 = passes[p].first, p1 = passes[p].second;
            if (cur < p0) {  // phases between two passes, once (after every chunk of the pass before)
                ex_.setSkipInit(false);
                ex_.execute(rc_, cq_, cg_, 0, 1, cur, p0);
            }
            for (uint64_t c = 0; c < nChunks; ++c, ++issued) {
                const int slot = static_cast<int>(issued & 1u);
                const uint64_t r0 = c * cr;
                const uint64_t n = rows == 0 ? 0 : std::min<uint64_t>(cr, rows - r0);
                if (issued >= 2) check(gdb_stream_wait_event(copy_, consumed_[slot]), "wait consumed");
                bool staged = false;
                for (size_t k : passCols[p]) staged |= static_cast<bool>(cols[k].produce);
                // the staging slot is free once the copy that last read it is done
                if (staged && issued >= 2) check(gdb_event_sync(copied_[slot]), "wait staging");
                for (size_t k : passCols[p]) {
                    const size_t eb = static_cast<size_t>(cols[k].elemBytes);
                    if (n) {
                        const void* src = staging_[slot][k];
                        if (cols[k].produce)
                            cols[k].produce(staging_[slot][k], r0, n);
                        else
                            src = static_cast<const char*>(cols[k].host) + r0 * eb;
                        check(gdb_memcpy_h2d_async(slots_[slot][k], src, static_cast<size_t>(n) * eb, copy_),
                              "copy chunk");
                    }
                    st.bytesCopied += n * eb;
                }
                check(gdb_event_record(copied_[slot], copy_), "event");
                check(gdb_stream_wait_event(comp, copied_[slot]), "wait copied");
                ex_.registerTable(table, n);
                for (size_t k = 0; k < cols.size(); ++k) ex_.registerTableColumn(table, cols[k].name, slots_[slot][k]);
                ex_.setSkipInit(c > 0);
                ex_.enqueue(rc_, cq_, cg_, p0, p1);
                check(gdb_event_record(consumed_[slot], comp), "event");
            }
            st.chunks += nChunks;
            cur = p1;
        }
    } catch (...) {
        ex_.setSkipInit(false);
        gdb_stream_sync(comp);
        gdb_stream_sync(copy_);
        throw;
    }
    ex_.setSkipInit(false);
    check(gdb_event_record(t1_, comp), "event");
    check(gdb_stream_sync(comp), "sync");
    float ms = 0;
    check(gdb_event_elapsed_ms(t0_, t1_, &ms), "elapsed");
    st.streamMs = ms;
    ex_.checkStatus(cg_, first, last);
    // Post-stream phases once. The stream table keeps the last chunk's
    // registration; post phases read other tables or runner buffers.
    if (last < np) ex_.execute(rc_, cq_, cg_, 0, 1, last, np);
    st.totalMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    return st;
}

}  // namespace runtime
}  // namespace gdb

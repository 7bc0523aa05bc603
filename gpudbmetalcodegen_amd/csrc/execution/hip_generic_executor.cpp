#include "hip_generic_executor.h"

#include "gdb/gdb_kernels.h"

#include <cctype>

#include <algorithm>
#include <chrono>
#include <cstring>

namespace gdb {
namespace codegen {

namespace {

void check(gdb_status st, const std::string& what) {
    if (st != GDB_OK) throw std::runtime_error(what + ": " + gdb_last_error());
}

}  // namespace

HipGenericExecutor::HipGenericExecutor(gdb_ctx ctx, gdb_stream stream)
    : ctx_(ctx), stream_(stream ? stream : gdb_default_stream(ctx)) {
    gdb_device_info_t info;
    check(gdb_device_info(ctx_, &info), "gdb_device_info");
    cuCount_ = info.cu_count;
}

HipGenericExecutor::~HipGenericExecutor() {
    releaseAllocatedBuffers();
    for (auto e : events_) gdb_event_destroy(e);
    if (dStatusPtrs_) gdb_free(ctx_, dStatusPtrs_);
    if (dStatusOut_) gdb_free(ctx_, dStatusOut_);
    if (pinned_) gdb_host_free_pinned(pinned_);
}

uint64_t* HipGenericExecutor::ensurePinned(size_t bytes) {
    if (bytes > pinnedBytes_) {
        check(gdb_stream_sync(stream_), "gdb_stream_sync");  // no copy into the old block is in flight
        if (pinned_) gdb_host_free_pinned(pinned_);
        pinned_ = nullptr;
        pinnedBytes_ = 0;
        const size_t nb = std::max<size_t>(bytes, 8192);
        void* h = nullptr;
        check(gdb_host_alloc_pinned(nb, &h), "gdb_host_alloc_pinned");
        pinned_ = static_cast<uint64_t*>(h);
        pinnedBytes_ = nb;
    }
    return pinned_;
}

void HipGenericExecutor::registerTable(const std::string& table, uint64_t rows) {
    tables_[table].rows = rows;
    sizes_.registerSymbol(tableSizeName(table), rows);
}

void HipGenericExecutor::registerTableColumn(const std::string& table, const std::string& column, const void* devptr) {
    if (reinterpret_cast<uintptr_t>(devptr) % 16 != 0)
        throw std::runtime_error("registerTableColumn: column " + column + " must be 16-byte aligned");
    tables_[table].columns[column] = devptr;
}

void HipGenericExecutor::registerBuffer(const std::string& name, void* devptr, size_t bytes) {
    auto it = buffers_.find(name);
    if (it != buffers_.end() && it->second.owned) gdb_free(ctx_, it->second.ptr);
    buffers_[name] = Buffer{devptr, bytes, false};
}

void HipGenericExecutor::registerBufferTable(const std::string& table, const std::string& column,
                                             const std::string& buffer, const std::string& rowsExpr) {
    for (auto& bt : bufferTables_)
        if (bt.table == table && bt.column == column) {
            bt.buffer = buffer;
            bt.rowsExpr = rowsExpr;
            return;
        }
    bufferTables_.push_back({table, column, buffer, rowsExpr});
}

void HipGenericExecutor::registerSymbol(const std::string& name, uint64_t value) { sizes_.registerSymbol(name, value); }

bool HipGenericExecutor::tryGetSymbol(const std::string& name, uint64_t& out) const {
    if (!sizes_.hasSymbol(name)) return false;
    out = sizes_.getSymbol(name);
    return true;
}

void HipGenericExecutor::registerScalarInt(const std::string& name, int64_t value) { scalars_[name] = value; }

void HipGenericExecutor::setPhaseHook(const std::string& phase, std::function<void(HipGenericExecutor&)> hook,
                                      const std::string& readBuffer, size_t readOffset, size_t readBytes) {
    if (readBytes > 4096) throw std::runtime_error("setPhaseHook: read-ahead of at most 4096 bytes");
    hooks_[phase] = Hook{std::move(hook), readBuffer, readOffset, readBytes};
}

HipGenericExecutor::Buffer& HipGenericExecutor::ensureBuffer(const HipParamBinding& b) {
    auto it = buffers_.find(b.name);
    if (b.sizeExpr.empty()) {
        if (it == buffers_.end())
            throw std::runtime_error("executor: buffer '" + b.name + "' is neither registered nor sized by the plan");
        return it->second;
    }
    const uint64_t elems = sizes_.resolve(b.sizeExpr);
    const size_t bytes = std::max<size_t>(static_cast<size_t>(elems) * b.elemBytes, 16);
    if (it != buffers_.end()) {
        if (!it->second.owned || it->second.bytes >= bytes) return it->second;
        check(gdb_free(ctx_, it->second.ptr), "gdb_free");
        buffers_.erase(it);
    }
    Buffer nb;
    nb.bytes = ((bytes + 255) / 256) * 256;
    check(gdb_alloc(ctx_, nb.bytes, &nb.ptr), "gdb_alloc(" + b.name + ", " + std::to_string(nb.bytes) + " bytes)");
    nb.owned = true;
    return buffers_[b.name] = nb;
}

void HipGenericExecutor::resolveBufferTables() {
    for (const auto& bt : bufferTables_) {
        auto it = buffers_.find(bt.buffer);
        if (it == buffers_.end()) continue;
        tables_[bt.table].columns[bt.column] = it->second.ptr;
        const uint64_t rows = sizes_.resolve(bt.rowsExpr);
        tables_[bt.table].rows = rows;
        sizes_.registerSymbol(tableSizeName(bt.table), rows);
    }
}

std::unordered_map<std::string, const HipParamBinding*> HipGenericExecutor::initSpecs(const HipCodegen& cg) const {
    std::unordered_map<std::string, const HipParamBinding*> spec;
    for (const auto& p : cg.getPhases())
        for (const auto& b : p.bindings)
            if (b.kind == HipParamKind::DeviceBuffer && b.init != InitKind::None && !spec.count(b.name))
                spec[b.name] = &b;
    return spec;
}

namespace {

// A size and pattern the multi-buffer fill takes (gdb_fill_multi_async):
// whole 8-byte words, a pattern of 1, 2 or 4 bytes or 1..4 words dividing it.
bool multiFillable(size_t bytes, int pl) {
    if (bytes % 8 != 0 || pl < 1) return false;
    if (pl == 1 || pl == 2 || pl == 4) return true;
    return pl % 8 == 0 && pl <= 32 && bytes % static_cast<size_t>(pl) == 0;
}

// The shortest prefix of a fill pattern that repeats to the whole pattern,
// among the lengths the multi-buffer fill takes (1, 2, 4 bytes, 1..4 words):
// a 16-word all-zero bucket pattern fills as one zero word. The pattern
// itself when none does.
std::vector<uint8_t> shortestPeriod(const std::vector<uint8_t>& pat) {
    for (size_t p : {1, 2, 4, 8, 16, 24, 32}) {
        if (p >= pat.size()) break;
        if (pat.size() % p != 0) continue;
        bool periodic = true;
        for (size_t i = p; i < pat.size() && periodic; ++i) periodic = pat[i] == pat[i % p];
        if (periodic) return std::vector<uint8_t>(pat.begin(), pat.begin() + static_cast<long>(p));
    }
    return pat;
}

bool mentionsSymbol(const std::string& expr, const std::string& sym) {
    size_t pos = 0;
    while ((pos = expr.find(sym, pos)) != std::string::npos) {
        const bool startOk = pos == 0 || !(std::isalnum(static_cast<unsigned char>(expr[pos - 1])) || expr[pos - 1] == '_');
        const size_t end = pos + sym.size();
        const bool endOk = end >= expr.size() || !(std::isalnum(static_cast<unsigned char>(expr[end])) || expr[end] == '_');
        if (startOk && endOk) return true;
        pos = end;
    }
    return false;
}

}  // namespace

void HipGenericExecutor::hoistInits(const HipCodegen& cg, int first, int last,
                                    std::unordered_map<std::string, bool>& initDone,
                                    const std::unordered_map<std::string, const HipParamBinding*>& initSpec) {
    if (skipInit_) return;
    std::vector<void*> dst;
    std::vector<size_t> bytes;
    std::vector<const void*> pats;
    std::vector<int> plen;
    std::vector<std::vector<uint8_t>> periods;  // each buffer's pattern at its shortest period
    const auto& phases = cg.getPhases();
    for (int pi = first; pi < last; ++pi) {
        const auto& p = phases[static_cast<size_t>(pi)];
        for (const auto& b : p.bindings) {
            if (b.kind != HipParamKind::DeviceBuffer || initDone[b.name]) continue;
            auto sp = initSpec.find(b.name);
            if (sp == initSpec.end() || sp->second != &b) continue;  // not defined by this phase
            const HipParamBinding* ib = sp->second;
            bool late = false;
            for (const auto& sym : lateSymbols_) late |= mentionsSymbol(ib->sizeExpr, sym);
            if (late) continue;
            uint64_t elems = 0;
            try {
                elems = sizes_.resolve(ib->sizeExpr);
            } catch (const std::exception&) {
                continue;  // sized by something not known yet: initialised at its phase
            }
            size_t nb = static_cast<size_t>(elems) * ib->elemBytes;
            // zero fills may cover the allocation's padding (allocations are
            // whole 256-byte multiples), so byte arrays of any length join
            if (ib->init == InitKind::Zero) nb = (nb + 7) & ~static_cast<size_t>(7);
            periods.push_back(ib->init == InitKind::Zero ? std::vector<uint8_t>(8, 0) : shortestPeriod(ib->pattern));
            const int pl = static_cast<int>(periods.back().size());
            if (nb == 0 || !multiFillable(nb, pl)) {
                periods.pop_back();
                continue;
            }
            Buffer& buf = ensureBuffer(*ib);
            dst.push_back(buf.ptr);
            bytes.push_back(nb);
            plen.push_back(pl);
            initDone[b.name] = true;
        }
    }
    for (const auto& pp : periods) pats.push_back(pp.data());
    if (!dst.empty())
        check(gdb_fill_multi_async(static_cast<int>(dst.size()), dst.data(), bytes.data(), pats.data(), plen.data(),
                                   stream_),
              "gdb_fill_multi_async");
}

int HipGenericExecutor::occupancy(gdb_kernel k, int block) {
    auto key = std::make_pair(k, block);
    auto it = occupancy_.find(key);
    if (it != occupancy_.end()) return it->second;
    int nb = 1;
    if (gdb_kernel_occupancy(k, block, 0, &nb) != GDB_OK || nb < 1) nb = 1;
    occupancy_[key] = nb;
    return nb;
}

void HipGenericExecutor::prepare(const HipCodegen& cg, int firstPhase, int lastPhase) {
    const auto& phases = cg.getPhases();
    const int last = lastPhase < 0 ? static_cast<int>(phases.size()) : lastPhase;
    for (int pi = firstPhase; pi < last; ++pi) {
        for (const auto& b : phases[static_cast<size_t>(pi)].bindings) {
            if (b.kind != HipParamKind::DeviceBuffer) continue;
            try {
                ensureBuffer(b);
            } catch (const std::runtime_error&) {
                // Sized by a symbol a hook registers later; allocated at launch.
            }
        }
        resolveBufferTables();
    }
}

void HipGenericExecutor::launchPhase(RuntimeCompiler& rc, const RuntimeCompiler::CompiledQuery& cq,
                                     const HipCodegen::PhaseInfo& p, std::unordered_map<std::string, bool>& initDone,
                                     const std::unordered_map<std::string, const HipParamBinding*>& initSpec,
                                     unsigned* gridOut, gdb_event beforeKernel) {
    std::vector<uint64_t> args;
    args.reserve(p.bindings.size());
    resolveBufferTables();
    for (const auto& b : p.bindings) {
        switch (b.kind) {
            case HipParamKind::TableColumn: {
                auto t = tables_.find(b.tableName);
                if (t == tables_.end()) throw std::runtime_error("executor: table '" + b.tableName + "' not registered");
                auto c = t->second.columns.find(b.columnName);
                if (c == t->second.columns.end())
                    throw std::runtime_error("executor: column '" + b.columnName + "' of table '" + b.tableName +
                                             "' not registered");
                args.push_back(reinterpret_cast<uint64_t>(c->second));
                break;
            }
            case HipParamKind::TableSize: {
                auto t = tables_.find(b.tableName);
                if (t == tables_.end()) throw std::runtime_error("executor: table '" + b.tableName + "' not registered");
                args.push_back(t->second.rows);
                break;
            }
            case HipParamKind::DeviceBuffer: {
                Buffer& buf = ensureBuffer(b);
                if (!skipInit_ && !initDone[b.name]) {
                    // A buffer is initialised right before the phase that
                    // defines it (the first phase binding it with an init
                    // kind), so a run over a phase range never re-fills a
                    // buffer an earlier phase built.
                    auto sp = initSpec.find(b.name);
                    const bool owner = sp != initSpec.end() && sp->second >= p.bindings.data() &&
                                       sp->second < p.bindings.data() + p.bindings.size();
                    if (owner) {
                        const HipParamBinding* ib = sp->second;
                        const size_t bytes = static_cast<size_t>(sizes_.resolve(ib->sizeExpr)) * ib->elemBytes;
                        if (ib->init == InitKind::Zero) {
                            const uint8_t z = 0;
                            check(gdb_fill_async(buf.ptr, bytes, &z, 1, stream_), "fill " + b.name);
                        } else {
                            check(gdb_fill_async(buf.ptr, bytes, ib->pattern.data(), static_cast<int>(ib->pattern.size()),
                                                 stream_),
                                  "fill " + b.name);
                        }
                        initDone[b.name] = true;
                    }
                }
                args.push_back(reinterpret_cast<uint64_t>(buf.ptr));
                break;
            }
            case HipParamKind::ScalarInt: {
                auto s = scalars_.find(b.name);
                if (s == scalars_.end()) throw std::runtime_error("executor: scalar '" + b.name + "' not registered");
                args.push_back(static_cast<uint64_t>(s->second));
                break;
            }
            case HipParamKind::SymbolValue: args.push_back(sizes_.resolve(b.sizeExpr)); break;
        }
    }
    const int items = p.isSingleThread ? 1 : p.items;
    bool known = false;
    for (int v : (p.isSingleThread ? std::vector<int>{1} : p.itemsVariants)) known |= (v == items);
    if (!known) throw std::runtime_error("executor: phase " + p.name + " has no ITEMS=" + std::to_string(items) + " variant");
    const int block = p.isSingleThread ? 64 : p.threadgroupSize;
    if (block > p.launchBound && !p.isSingleThread)
        throw std::runtime_error("executor: block " + std::to_string(block) + " exceeds launch bound of " + p.name);
    gdb_kernel k = rc.getPipeline(cq, HipCodegen::kernelName(p.name, items));
    unsigned grid = 1;
    if (!p.isSingleThread) {
        const uint64_t rows = p.rowsExpr.empty() ? 1 : sizes_.resolve(p.rowsExpr);
        const uint64_t tile = static_cast<uint64_t>(block) * static_cast<uint64_t>(items);
        uint64_t need = std::max<uint64_t>(1, (rows + tile - 1) / tile);
        uint64_t cap = static_cast<uint64_t>(cuCount_) * static_cast<uint64_t>(occupancy(k, block));
        if (p.maxThreadgroups > 0) cap = std::min<uint64_t>(cap, static_cast<uint64_t>(p.maxThreadgroups));
        cap = std::max<uint64_t>(cap, 1);
        grid = static_cast<unsigned>(std::min(need, cap));
    }
    if (gridOut) *gridOut = grid;
    if (beforeKernel) check(gdb_event_record(beforeKernel, stream_), "gdb_event_record");
    check(gdb_launch(k, grid, static_cast<unsigned>(block), 0, args.data(), args.size() * 8, stream_),
          "launch " + p.name);
}

std::vector<uint64_t> HipGenericExecutor::packPhaseArgs(const HipCodegen& cg, int phase) {
    const auto& p = cg.getPhases().at(static_cast<size_t>(phase));
    resolveBufferTables();
    std::vector<uint64_t> args;
    for (const auto& b : p.bindings) {
        switch (b.kind) {
            case HipParamKind::TableColumn: {
                auto t = tables_.find(b.tableName);
                if (t == tables_.end()) throw std::runtime_error("executor: table '" + b.tableName + "' not registered");
                auto c = t->second.columns.find(b.columnName);
                if (c == t->second.columns.end())
                    throw std::runtime_error("executor: column '" + b.columnName + "' not registered");
                args.push_back(reinterpret_cast<uint64_t>(c->second));
                break;
            }
            case HipParamKind::TableSize: {
                auto t = tables_.find(b.tableName);
                if (t == tables_.end()) throw std::runtime_error("executor: table '" + b.tableName + "' not registered");
                args.push_back(t->second.rows);
                break;
            }
            case HipParamKind::DeviceBuffer: args.push_back(reinterpret_cast<uint64_t>(ensureBuffer(b).ptr)); break;
            case HipParamKind::ScalarInt: {
                auto s = scalars_.find(b.name);
                if (s == scalars_.end()) throw std::runtime_error("executor: scalar '" + b.name + "' not registered");
                args.push_back(static_cast<uint64_t>(s->second));
                break;
            }
            case HipParamKind::SymbolValue: args.push_back(sizes_.resolve(b.sizeExpr)); break;
        }
    }
    return args;
}

void HipGenericExecutor::collectInits(const HipCodegen& cg, FillList& out) {
    for (const auto& kv : initSpecs(cg)) {
        const HipParamBinding* ib = kv.second;
        Buffer& buf = ensureBuffer(*ib);
        size_t bytes = static_cast<size_t>(sizes_.resolve(ib->sizeExpr)) * ib->elemBytes;
        if (!bytes) continue;
        std::vector<uint8_t> pat = ib->init == InitKind::Zero ? std::vector<uint8_t>(8, 0) : shortestPeriod(ib->pattern);
        // zero fills may cover the allocation's padding (whole 256-byte multiples)
        if (ib->init == InitKind::Zero) bytes = (bytes + 7) & ~static_cast<size_t>(7);
        const int pl = static_cast<int>(pat.size());
        const bool multi = multiFillable(bytes, pl) && reinterpret_cast<uintptr_t>(buf.ptr) % 8 == 0 && bytes <= buf.bytes;
        if (!multi) {
            out.singles.push_back({buf.ptr, bytes, pat});
            continue;
        }
        std::array<uint8_t, 32> w{};
        for (int k = 0; k < 32; ++k) w[static_cast<size_t>(k)] = pat[static_cast<size_t>(k % pl)];
        out.dst.push_back(buf.ptr);
        out.bytes.push_back(bytes);
        out.pattern.push_back(w);
        out.plen.push_back(pl);
    }
}

std::vector<HipGenericExecutor::StatusRef> HipGenericExecutor::statusRefs(const HipCodegen& cg) const {
    std::vector<StatusRef> out;
    for (const auto& p : cg.getPhases()) {
        auto st = buffers_.find(p.statusBuffer);
        if (!p.statusBuffer.empty() && st != buffers_.end()) out.push_back({st->second.ptr, p.name, false, ""});
        auto ck = buffers_.find(p.checkBuffer);
        if (!p.checkBuffer.empty() && ck != buffers_.end()) out.push_back({ck->second.ptr, p.name, true, p.checkMessage});
    }
    return out;
}

void HipGenericExecutor::judgeStatus(const StatusRef& r, const uint64_t* w) {
    if (!r.check && w[0] != 0) throw GdbOverflowError(r.phase, w[2]);
    if (r.check && w[0] != 0)
        throw std::runtime_error("phase " + r.phase + ": " + r.message + " (" + std::to_string(w[2]) + " rows)");
}

void HipGenericExecutor::enqueueInits(const HipCodegen& cg) {
    for (const auto& kv : initSpecs(cg)) {
        const HipParamBinding* ib = kv.second;
        Buffer& buf = ensureBuffer(*ib);
        const size_t bytes = static_cast<size_t>(sizes_.resolve(ib->sizeExpr)) * ib->elemBytes;
        if (ib->init == InitKind::Zero) {
            const uint8_t z = 0;
            check(gdb_fill_async(buf.ptr, bytes, &z, 1, stream_), "fill " + ib->name);
        } else {
            check(gdb_fill_async(buf.ptr, bytes, ib->pattern.data(), static_cast<int>(ib->pattern.size()), stream_),
                  "fill " + ib->name);
        }
    }
}

void HipGenericExecutor::enqueue(RuntimeCompiler& rc, const RuntimeCompiler::CompiledQuery& cq, const HipCodegen& cg,
                                 int firstPhase, int lastPhase) {
    const auto& phases = cg.getPhases();
    const int last = lastPhase < 0 ? static_cast<int>(phases.size()) : lastPhase;
    auto spec = initSpecs(cg);
    std::unordered_map<std::string, bool> initDone;
    hoistInits(cg, firstPhase, last, initDone, spec);
    for (int pi = firstPhase; pi < last; ++pi)
        launchPhase(rc, cq, phases[static_cast<size_t>(pi)], initDone, spec, nullptr);
}

HipExecutionResult HipGenericExecutor::execute(RuntimeCompiler& rc, const RuntimeCompiler::CompiledQuery& cq,
                                               const HipCodegen& cg, int warmupRuns, int measuredRuns, int firstPhase,
                                               int lastPhase) {
    const auto& phases = cg.getPhases();
    const int last = lastPhase < 0 ? static_cast<int>(phases.size()) : lastPhase;
    const int np = last - firstPhase;
    HipExecutionResult res;
    auto t0 = std::chrono::steady_clock::now();
    prepare(cg, firstPhase, last);
    res.bufferAllocTimeMs =
        std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    while (static_cast<int>(events_.size()) < 2 * np + 1) {
        gdb_event e;
        check(gdb_event_create(&e), "gdb_event_create");
        events_.push_back(e);
    }
    auto spec = initSpecs(cg);
    const int runs = warmupRuns + std::max(1, measuredRuns);
    measuredPhase_.assign(static_cast<size_t>(np), 0.0f);
    measuredFill_.assign(static_cast<size_t>(np), 0.0f);
    PendingStatus pending;
    for (int r = 0; r < runs; ++r) {
        const bool measured = r >= warmupRuns;
        std::unordered_map<std::string, bool> initDone;
        res.phaseTimesMs.assign(static_cast<size_t>(np), 0.0f);
        res.phaseFillMs.assign(static_cast<size_t>(np), 0.0f);
        res.phaseNames.clear();
        res.phaseGrid.clear();
        res.phaseBlock.clear();
        res.phaseItems.clear();
        check(gdb_event_record(events_[0], stream_), "gdb_event_record");
        hoistInits(cg, firstPhase, last, initDone, spec);
        for (int pi = firstPhase; pi < last; ++pi) {
            const auto& p = phases[static_cast<size_t>(pi)];
            unsigned grid = 0;
            const size_t e = static_cast<size_t>(2 * (pi - firstPhase));
            launchPhase(rc, cq, p, initDone, spec, &grid, phaseTiming_ ? events_[e + 1] : nullptr);
            if (phaseTiming_ || pi + 1 == last) check(gdb_event_record(events_[e + 2], stream_), "gdb_event_record");
            res.phaseNames.push_back(p.name);
            res.phaseGrid.push_back(grid);
            res.phaseBlock.push_back(static_cast<unsigned>(p.isSingleThread ? 64 : p.threadgroupSize));
            res.phaseItems.push_back(static_cast<unsigned>(p.isSingleThread ? 1 : p.items));
            auto h = hooks_.find(p.name);
            if (h != hooks_.end()) {
                if (h->second.bytes) {  // the hook's read rides behind the phase: one synchronisation
                    auto bit = buffers_.find(h->second.buffer);
                    if (bit == buffers_.end()) throw std::runtime_error("executor: no buffer " + h->second.buffer);
                    if (h->second.offset + h->second.bytes > bit->second.bytes)
                        throw std::runtime_error("executor: hook read past the end of " + h->second.buffer);
                    uint64_t* dst = ensurePinned(h->second.bytes);
                    check(gdb_memcpy_d2h_async(dst, static_cast<const char*>(bit->second.ptr) + h->second.offset,
                                               h->second.bytes, stream_),
                          "hook read " + h->second.buffer);
                }
                check(gdb_stream_sync(stream_), "gdb_stream_sync");
                h->second.fn(*this);
            }
        }
        if (r + 1 == runs) {
            // the final run: status words gathered and read behind it, one wait for both
            pending = enqueueStatus(cg, firstPhase, last);
            check(gdb_stream_sync(stream_), "gdb_stream_sync");
        } else {
            check(gdb_event_sync(events_[static_cast<size_t>(2 * np)]), "gdb_event_sync");
        }
        if (measured) {
            const float w = 1.0f / static_cast<float>(std::max(1, measuredRuns));
            for (int i = 0; phaseTiming_ && i < np; ++i) {
                float ms = 0.0f, fill = 0.0f;
                const size_t e = static_cast<size_t>(2 * i);
                check(gdb_event_elapsed_ms(events_[e + 1], events_[e + 2], &ms), "gdb_event_elapsed_ms");
                check(gdb_event_elapsed_ms(events_[e], events_[e + 1], &fill), "gdb_event_elapsed_ms");
                measuredPhase_[static_cast<size_t>(i)] += ms * w;
                measuredFill_[static_cast<size_t>(i)] += fill * w;
            }
            float total = 0.0f;
            check(gdb_event_elapsed_ms(events_[0], events_[static_cast<size_t>(2 * np)], &total), "gdb_event_elapsed_ms");
            res.totalKernelTimeMs += total * w;
        }
    }
    res.phaseTimesMs = measuredPhase_;
    res.phaseFillMs = measuredFill_;
    res.phaseStatus = judgeStatuses(pending);
    return res;
}

HipGenericExecutor::PendingStatus HipGenericExecutor::enqueueStatus(const HipCodegen& cg, int firstPhase,
                                                                    int lastPhase) {
    // Every status and check word of the range is gathered on the device and
    // read back with one copy; the caller's next synchronisation covers it.
    const auto& phases = cg.getPhases();
    const int last = lastPhase < 0 ? static_cast<int>(phases.size()) : lastPhase;
    PendingStatus ps;
    ps.phases = static_cast<size_t>(std::max(0, last - firstPhase));
    std::vector<const void*> ptrs;
    for (int pi = firstPhase; pi < last; ++pi) {
        const auto& p = phases[static_cast<size_t>(pi)];
        auto st = buffers_.find(p.statusBuffer);
        if (!p.statusBuffer.empty() && st != buffers_.end()) {
            ps.refs.push_back({pi - firstPhase, {st->second.ptr, p.name, false, ""}});
            ptrs.push_back(st->second.ptr);
        }
        auto ck = buffers_.find(p.checkBuffer);
        if (!p.checkBuffer.empty() && ck != buffers_.end()) {
            ps.refs.push_back({pi - firstPhase, {ck->second.ptr, p.name, true, p.checkMessage}});
            ptrs.push_back(ck->second.ptr);
        }
    }
    if (ps.refs.empty()) return ps;
    const size_t n = ps.refs.size();
    uint64_t* host = ensurePinned(n * 32 + n * sizeof(void*));
    if (n > statusCap_) {
        check(gdb_stream_sync(stream_), "gdb_stream_sync");
        if (dStatusPtrs_) gdb_free(ctx_, dStatusPtrs_);
        if (dStatusOut_) gdb_free(ctx_, dStatusOut_);
        dStatusPtrs_ = dStatusOut_ = nullptr;
        statusCap_ = std::max<size_t>(n, 64);
        check(gdb_alloc(ctx_, statusCap_ * sizeof(void*), &dStatusPtrs_), "gdb_alloc");
        check(gdb_alloc(ctx_, statusCap_ * 32, &dStatusOut_), "gdb_alloc");
        statusPtrCache_.clear();
    }
    if (ptrs != statusPtrCache_) {
        std::memcpy(host + 4 * n, ptrs.data(), n * sizeof(void*));
        check(gdb_memcpy_h2d_async(dStatusPtrs_, host + 4 * n, n * sizeof(void*), stream_), "status pointers");
        statusPtrCache_ = ptrs;
    }
    check(gdb_gather_words(static_cast<const uint64_t* const*>(dStatusPtrs_), static_cast<int>(n), 4,
                           static_cast<uint64_t*>(dStatusOut_), stream_),
          "gather status words");
    check(gdb_memcpy_d2h_async(host, dStatusOut_, n * 32, stream_), "read status words");
    ps.host = host;
    return ps;
}

std::vector<std::array<uint64_t, 4>> HipGenericExecutor::judgeStatuses(const PendingStatus& ps) const {
    std::vector<std::array<uint64_t, 4>> out(ps.phases, {0, 0, 0, 0});
    for (size_t i = 0; i < ps.refs.size(); ++i) {
        const uint64_t* w = ps.host + 4 * i;
        judgeStatus(ps.refs[i].second, w);
        if (!ps.refs[i].second.check)
            for (int k = 0; k < 4; ++k) out[static_cast<size_t>(ps.refs[i].first)][static_cast<size_t>(k)] = w[k];
    }
    return out;
}

std::vector<std::array<uint64_t, 4>> HipGenericExecutor::checkStatus(const HipCodegen& cg, int firstPhase,
                                                                     int lastPhase) {
    const PendingStatus ps = enqueueStatus(cg, firstPhase, lastPhase);
    if (!ps.refs.empty()) check(gdb_stream_sync(stream_), "gdb_stream_sync");
    return judgeStatuses(ps);
}

GenericResult HipGenericExecutor::collectResult(const HipCodegen& cg) const {
    check(gdb_stream_sync(stream_), "gdb_stream_sync");
    return HipResultCollector::collect(
        cg.getResultSchema(),
        [this](const std::string& name, size_t off, void* dst, size_t bytes) { readBuffer(name, off, dst, bytes); },
        [this](const std::string& expr) { return sizes_.resolve(expr); });
}

void HipGenericExecutor::releaseAllocatedBuffers() {
    for (auto& kv : buffers_)
        if (kv.second.owned && kv.second.ptr) gdb_free(ctx_, kv.second.ptr);
    buffers_.clear();
}

void* HipGenericExecutor::getAllocatedBuffer(const std::string& name) const {
    auto it = buffers_.find(name);
    return it == buffers_.end() ? nullptr : it->second.ptr;
}

size_t HipGenericExecutor::getAllocatedBytes(const std::string& name) const {
    auto it = buffers_.find(name);
    return it == buffers_.end() ? 0 : it->second.bytes;
}

size_t HipGenericExecutor::ownedBytes() const {
    size_t n = 0;
    for (const auto& kv : buffers_)
        if (kv.second.owned) n += kv.second.bytes;
    return n;
}

void HipGenericExecutor::readBuffer(const std::string& name, size_t offset, void* host, size_t bytes) const {
    auto it = buffers_.find(name);
    if (it == buffers_.end()) throw std::runtime_error("executor: no buffer " + name);
    if (offset + bytes > it->second.bytes)
        throw std::runtime_error("executor: read past the end of buffer " + name);
    check(gdb_memcpy_d2h_async(host, static_cast<const char*>(it->second.ptr) + offset, bytes, stream_),
          "read " + name);
    check(gdb_stream_sync(stream_), "gdb_stream_sync");
}

}  // namespace codegen
}  // namespace gdb

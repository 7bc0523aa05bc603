#include "batcher.h"

#include "../core/gdb_device_header.h"
#include "gdb/gdb_kernels.h"

#include <algorithm>
#include <cstring>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <unordered_set>

namespace gdb {
namespace runtime {

using namespace codegen;

namespace {

void check(gdb_status st, const std::string& what) {
    if (st != GDB_OK) throw std::runtime_error(what + ": " + gdb_last_error());
}

std::string stripRestrict(std::string t) {
    const std::string r = " __restrict__";
    auto p = t.find(r);
    if (p != std::string::npos) t.erase(p, r.size());
    return t;
}

}  // namespace

PipelineBatch::PipelineBatch(gdb_ctx ctx) : ctx_(ctx), stream_(gdb_default_stream(ctx)), rc_(ctx) {
    check(gdb_event_create(&ev0_), "gdb_event_create");
    check(gdb_event_create(&ev1_), "gdb_event_create");
}

PipelineBatch::~PipelineBatch() {
    if (graph_) gdb_graph_destroy(graph_);
    freeDevice();
    gdb_event_destroy(ev0_);
    gdb_event_destroy(ev1_);
}

void PipelineBatch::freeDevice() {
    for (auto& L : levels_)
        for (void** p : {&L.dArgs, &L.dKindItems, &L.dItemPipe, &L.dItemBegin, &L.dItemEnd}) {
            if (*p) gdb_free(ctx_, *p);
            *p = nullptr;
        }
    for (void** p : {&dTicket_, &dFill_, &dStatusSrc_, &dStatusOut_}) {
        if (*p) gdb_free(ctx_, *p);
        *p = nullptr;
    }
    if (hStatus_) gdb_host_free_pinned(hStatus_);
    hStatus_ = nullptr;
}

std::vector<int> BatchKernelSource::add(const HipQueryPlan* plan, const HipCodegen* cg, int items) {
    const auto& phases = cg->getPhases();
    if (phases.empty() || plan->phases.size() != phases.size())
        throw std::runtime_error("PipelineBatch: plan " + plan->name + " has no phases");
    std::vector<int> out;
    std::unique_ptr<HipCodegen> rcg;  // range-mode codegen of the plan, made once
    for (size_t lv = 0; lv < phases.size(); ++lv) {
        const auto& ph = phases[lv];
        if (ph.isSingleThread) throw std::runtime_error("PipelineBatch: plan " + plan->name + " has a single-thread phase");
        const int it = lv == 0 ? items : ph.items;
        bool known = false;
        for (int v : ph.itemsVariants) known |= (v == it);
        if (!known)
            throw std::runtime_error("PipelineBatch: phase " + ph.name + " of " + plan->name + " has no ITEMS=" +
                                     std::to_string(it));
        const std::string key = plan->name + "\n" + std::to_string(lv) + "\n" + ph.code;
        int kind = -1;
        for (size_t k = 0; k < kinds_.size(); ++k)
            if (kinds_[k].key == key) kind = static_cast<int>(k);
        if (kind < 0) {
            if (!rcg) rcg = std::make_unique<HipCodegen>(generateFromPlan(*plan, true));
            Kind k;
            k.key = key;
            k.body = "k" + std::to_string(kinds_.size()) + "_item";
            k.level = static_cast<int>(lv);
            for (const auto& b : rcg->getPhases()[lv].bindings) k.argTypes.push_back(stripRestrict(b.declType));
            kindCg_.push_back(std::make_unique<HipCodegen>(*rcg));
            kinds_.push_back(k);
            kind = static_cast<int>(kinds_.size()) - 1;
        }
        auto& ki = kinds_[static_cast<size_t>(kind)].items;
        if (std::find(ki.begin(), ki.end(), it) == ki.end()) ki.push_back(it);
        out.push_back(kind);
    }
    return out;
}

int BatchKernelSource::levels() const {
    int n = 0;
    for (const auto& k : kinds_) n = std::max(n, k.level + 1);
    return n;
}

std::string BatchKernelSource::kernelName(int level) {
    return level == 0 ? "gdb_batch_kernel" : "gdb_batch_kernel_l" + std::to_string(level);
}

int BatchKernelSource::launchBound() const {
    int lb = 1024;
    for (size_t k = 0; k < kinds_.size(); ++k)
        lb = std::min(lb, kindCg_[k]->getPhases()[static_cast<size_t>(kinds_[k].level)].launchBound);
    return lb;
}

std::string BatchKernelSource::print() const {
    if (kinds_.empty()) throw std::runtime_error("PipelineBatch: no pipelines");
    std::ostringstream src;
    src << kGdbDeviceHeader << "\n// ===== helpers =====\n";
    std::unordered_set<std::string> seen;
    for (auto& kc : kindCg_) {
        std::string h = kc->printHelpers();
        if (seen.insert(h).second) src << h;
    }
    for (size_t k = 0; k < kinds_.size(); ++k)
        src << "\n// ===== pipeline kind " << k << " (phase " << kinds_[k].level << ") =====\n"
            << kindCg_[k]->printPhaseBody(kinds_[k].level, kinds_[k].body);
    src << R"(
template <typename T>
GDB_DEV T gdb_arg(u64 w) {
    if constexpr (__is_pointer(T)) return reinterpret_cast<T>(w);
    else return (T)w;
}
)";
    for (int lv = 0; lv < levels(); ++lv) {
        // One LDS buffer shared by the level's bodies: the largest need, not the sum.
        std::string ldsMax = "16u";
        for (size_t k = 0; k < kinds_.size(); ++k)
            if (kinds_[k].level == lv)
                for (int it : kinds_[k].items)
                    ldsMax = "gdb_cmax(" + ldsMax + ", " + kinds_[k].body + "_lds_bytes<" + std::to_string(it) + ">())";
        src << "extern \"C\" __global__ void __launch_bounds__(" << launchBound() << ") " << kernelName(lv)
            << "(const u64* __restrict__ args, "
            << "const u32 args_stride, const u32* __restrict__ kind_items, const u32* __restrict__ item_pipe, "
            << "const u64* __restrict__ item_begin, const u64* __restrict__ item_end, const u32 n_items, "
            << "u32* __restrict__ ticket) {\n"
            << "    __shared__ u32 _it;\n"
            << "    __shared__ __attribute__((aligned(16))) unsigned char _lds[" << ldsMax << "];\n"
            << "    for (;;) {\n"
            << "        __syncthreads();\n"
            << "        if (threadIdx.x == 0) _it = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, "
               "__HIP_MEMORY_SCOPE_AGENT);\n"
            << "        __syncthreads();\n"
            << "        const u32 it = _it;\n"
            << "        if (it >= n_items) break;\n"
            << "        const u32 p = item_pipe[it];\n"
            << "        const u64* a = args + (u64)p * args_stride;\n"
            << "        const u64 rb = item_begin[it], re = item_end[it];\n"
            << "        switch (kind_items[p]) {\n";
        for (size_t k = 0; k < kinds_.size(); ++k) {
            if (kinds_[k].level != lv) continue;
            for (int it : kinds_[k].items) {
                src << "        case " << (k * 32 + static_cast<size_t>(it)) << ": " << kinds_[k].body << "<" << it << ">(";
                for (size_t a = 0; a < kinds_[k].argTypes.size(); ++a)
                    src << (a ? ", " : "") << "gdb_arg<" << kinds_[k].argTypes[a] << ">(a[" << a << "])";
                src << (kinds_[k].argTypes.empty() ? "" : ", ") << "rb, re, _lds); break;\n";
            }
        }
        // the last block out resets the ticket for the next level or run
        src << "        default: break;\n        }\n    }\n"
            << "    if (threadIdx.x == 0) {\n"
            << "        const u32 done = __hip_atomic_fetch_add(ticket + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);\n"
            << "        if (done == gridDim.x - 1) { ticket[0] = 0; ticket[1] = 0; }\n"
            << "    }\n}\n";
    }
    return src.str();
}

int PipelineBatch::add(const HipQueryPlan* plan, const HipCodegen* cg, HipGenericExecutor* ex, int items) {
    Pipe p{plan, cg, ex, {}, src_.add(plan, cg, items), {}};
    const auto& phases = cg->getPhases();
    for (size_t lv = 0; lv < phases.size(); ++lv) {
        p.items.push_back(lv == 0 ? items : phases[lv].items);
        uint64_t rows = 0;
        if (lv == 0) {
            const std::string tbl = phases[0].scannedTable;
            if (!ex->tryGetSymbol(tableSizeName(tbl), rows))
                throw std::runtime_error("PipelineBatch: table " + tbl + " of pipeline not registered");
        } else {
            rows = phases[lv].rowsExpr.empty() ? 1 : ex->resolve(phases[lv].rowsExpr);
        }
        p.rows.push_back(rows);
    }
    pipes_.push_back(std::move(p));
    return static_cast<int>(pipes_.size()) - 1;
}

size_t PipelineBatch::workItems() const {
    size_t n = 0;
    for (const auto& L : levels_) n += L.itemPipe.size();
    return n;
}

void PipelineBatch::build(int block, uint64_t chunkRows, int grid) {
    if (pipes_.empty()) throw std::runtime_error("PipelineBatch: no pipelines");
    if (block < 64 || block > 1024 || block % 64) throw std::runtime_error("PipelineBatch: bad block size");
    if (block > src_.launchBound()) throw std::runtime_error("PipelineBatch: block size above a pipeline's launch bound");
    block_ = block;
    source_ = src_.print();
    cq_ = rc_.compile(source_);
    freeDevice();
    levels_.assign(static_cast<size_t>(src_.levels()), Level{});
    auto up = [&](void*& d, const void* h, size_t bytes) {
        check(gdb_alloc(ctx_, std::max<size_t>(bytes, 16), &d), "gdb_alloc");
        if (bytes) check(gdb_memcpy_h2d_async(d, h, bytes, stream_), "upload");
    };
    gdb_device_info_t info;
    check(gdb_device_info(ctx_, &info), "gdb_device_info");
    for (int lv = 0; lv < src_.levels(); ++lv) {
        Level& L = levels_[static_cast<size_t>(lv)];
        L.kernel = rc_.getPipeline(cq_, BatchKernelSource::kernelName(lv));
        // --- arguments of the level's phase, one row per pipeline -----------
        L.argStride = 1;
        for (size_t k = 0; k < src_.kinds(); ++k)
            if (src_.level(static_cast<int>(k)) == lv)
                L.argStride = std::max<uint32_t>(L.argStride,
                                                 static_cast<uint32_t>(src_.argTypes(static_cast<int>(k)).size()));
        std::vector<uint64_t> args(static_cast<size_t>(L.argStride) * pipes_.size(), 0);
        std::vector<uint32_t> kindItems(pipes_.size(), 0xFFFFFFFFu);
        for (size_t i = 0; i < pipes_.size(); ++i) {
            auto& p = pipes_[i];
            if (static_cast<size_t>(lv) >= p.kinds.size()) continue;
            const int kind = p.kinds[static_cast<size_t>(lv)];
            auto a = p.ex->packPhaseArgs(src_.kindCodegen(kind), lv);
            std::copy(a.begin(), a.end(), args.begin() + static_cast<long>(i * L.argStride));
            kindItems[i] = static_cast<uint32_t>(kind * 32 + p.items[static_cast<size_t>(lv)]);
        }
        // --- work items: whole tiles, about chunkRows rows each, largest first
        std::vector<size_t> order;
        for (size_t i = 0; i < pipes_.size(); ++i)
            if (static_cast<size_t>(lv) < pipes_[i].kinds.size()) order.push_back(i);
        std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
            return pipes_[a].rows[static_cast<size_t>(lv)] > pipes_[b].rows[static_cast<size_t>(lv)];
        });
        for (size_t i : order) {
            const auto& p = pipes_[i];
            const uint64_t rows = p.rows[static_cast<size_t>(lv)];
            const uint64_t tile = static_cast<uint64_t>(block) * static_cast<uint64_t>(p.items[static_cast<size_t>(lv)]);
            const uint64_t want =
                static_cast<size_t>(lv) < levelChunk_.size() && levelChunk_[static_cast<size_t>(lv)] > 0
                    ? levelChunk_[static_cast<size_t>(lv)]
                    : chunkRows;
            const uint64_t chunk = tile * std::max<uint64_t>(1, (want + tile - 1) / tile);
            for (uint64_t b = 0; b < rows; b += chunk) {
                L.itemPipe.push_back(static_cast<uint32_t>(i));
                L.itemBegin.push_back(b);
                L.itemEnd.push_back(std::min(rows, b + chunk));
            }
        }
        up(L.dArgs, args.data(), args.size() * 8);
        up(L.dKindItems, kindItems.data(), kindItems.size() * 4);
        up(L.dItemPipe, L.itemPipe.data(), L.itemPipe.size() * 4);
        up(L.dItemBegin, L.itemBegin.data(), L.itemBegin.size() * 8);
        up(L.dItemEnd, L.itemEnd.data(), L.itemEnd.size() * 8);
        int occ = 1;
        if (gdb_kernel_occupancy(L.kernel, block, 0, &occ) != GDB_OK || occ < 1) occ = 1;
        const unsigned full = static_cast<unsigned>(info.cu_count * occ);
        L.grid = grid > 0 ? static_cast<unsigned>(grid) : full;
        L.grid = std::max(1u, std::min<unsigned>(L.grid, static_cast<unsigned>(std::max<size_t>(1, L.itemPipe.size()))));
    }
    const uint64_t zero = 0;
    up(dTicket_, &zero, 8);
    HipGenericExecutor::FillList fl;
    for (auto& p : pipes_) p.ex->collectInits(*p.cg, fl);
    fl.dst.push_back(dTicket_);
    fl.bytes.push_back(8);
    fl.pattern.push_back({});
    fl.plen.push_back(8);
    std::vector<gdb_fill_desc> table(fl.dst.size());
    fillMaxWords_ = 0;
    for (size_t i = 0; i < fl.dst.size(); ++i) {
        gdb_fill_desc& d = table[i];
        const int pl = fl.plen[i];
        d.dst = fl.dst[i];
        d.words = fl.bytes[i] / 8;
        d.period = (pl == 1 || pl == 2 || pl == 4) ? 1 : static_cast<uint64_t>(pl / 8);
        std::memcpy(d.pattern, fl.pattern[i].data(), 32);
        fillMaxWords_ = std::max<uint64_t>(fillMaxWords_, d.words);
    }
    fillN_ = static_cast<int>(table.size());
    up(dFill_, table.data(), table.size() * sizeof(gdb_fill_desc));
    fillSingles_ = std::move(fl.singles);
    statusRefs_.clear();
    for (const auto& p : pipes_)
        for (const auto& r : p.ex->statusRefs(*p.cg)) statusRefs_.push_back(r);
    if (!statusRefs_.empty()) {
        std::vector<const void*> src;
        for (const auto& r : statusRefs_) src.push_back(r.ptr);
        up(dStatusSrc_, src.data(), src.size() * sizeof(void*));
        check(gdb_alloc(ctx_, statusRefs_.size() * 32, &dStatusOut_), "gdb_alloc");
        void* h = nullptr;
        check(gdb_host_alloc_pinned(statusRefs_.size() * 32, &h), "gdb_host_alloc_pinned");
        hStatus_ = static_cast<uint64_t*>(h);
    }
    check(gdb_stream_sync(stream_), "sync");
    built_ = true;
    if (graph_) {
        gdb_graph_destroy(graph_);
        graph_ = nullptr;
    }
}

void PipelineBatch::setLevelChunk(int level, uint64_t rows) {
    if (level < 0 || level > 64) throw std::runtime_error("PipelineBatch: bad level");
    if (levelChunk_.size() <= static_cast<size_t>(level)) levelChunk_.resize(static_cast<size_t>(level) + 1, 0);
    levelChunk_[static_cast<size_t>(level)] = rows;
}

void PipelineBatch::enqueue() {
    if (!built_) throw std::runtime_error("PipelineBatch: build() first");
    // every pipeline's buffer fills and the work ticket: one launch over the
    // fill table made at build (one kernel for 64 pipelines' ~130 buffers)
    check(gdb_fill_table_async(static_cast<const gdb_fill_desc*>(dFill_), fillN_, fillMaxWords_, stream_),
          "fill pipeline buffers");
    for (const auto& f : fillSingles_)
        check(gdb_fill_async(f.dst, f.bytes, f.pattern.data(), static_cast<int>(f.pattern.size()), stream_),
              "fill pipeline buffer");
    for (size_t lv = 0; lv < levels_.size(); ++lv) {
        const Level& L = levels_[lv];
        if (L.itemPipe.empty()) continue;
        // Kernel arguments: pointers 8 bytes; the two u32 arguments are
        // packed at their natural 4-byte alignment by the layout below.
        std::vector<uint8_t> buf;
        auto put = [&](const void* p, size_t n, size_t align) {
            while (buf.size() % align) buf.push_back(0);
            const uint8_t* b = static_cast<const uint8_t*>(p);
            buf.insert(buf.end(), b, b + n);
        };
        const uint64_t pa = reinterpret_cast<uint64_t>(L.dArgs);
        const uint32_t st = L.argStride;
        const uint64_t pk = reinterpret_cast<uint64_t>(L.dKindItems), pp = reinterpret_cast<uint64_t>(L.dItemPipe);
        const uint64_t pb = reinterpret_cast<uint64_t>(L.dItemBegin), pe = reinterpret_cast<uint64_t>(L.dItemEnd);
        const uint32_t ni = static_cast<uint32_t>(L.itemPipe.size());
        const uint64_t pt = reinterpret_cast<uint64_t>(dTicket_);
        put(&pa, 8, 8);
        put(&st, 4, 4);
        put(&pk, 8, 8);
        put(&pp, 8, 8);
        put(&pb, 8, 8);
        put(&pe, 8, 8);
        put(&ni, 4, 4);
        put(&pt, 8, 8);
        while (buf.size() % 8) buf.push_back(0);
        check(gdb_launch(L.kernel, L.grid, static_cast<unsigned>(block_), 0, buf.data(), buf.size(), stream_),
              "launch " + BatchKernelSource::kernelName(static_cast<int>(lv)));
    }
    if (!statusRefs_.empty()) {
        const int n = static_cast<int>(statusRefs_.size());
        check(gdb_gather_words(static_cast<const uint64_t* const*>(dStatusSrc_), n, 4, static_cast<uint64_t*>(dStatusOut_),
                               stream_),
              "gather status words");
        check(gdb_memcpy_d2h_async(hStatus_, dStatusOut_, static_cast<size_t>(n) * 32, stream_), "read status words");
    }
}

void PipelineBatch::capture() {
    if (graph_) {
        gdb_graph_destroy(graph_);
        graph_ = nullptr;
    }
    check(gdb_graph_begin_capture(stream_), "begin capture");
    try {
        enqueue();
    } catch (...) {
        gdb_graph g = nullptr;
        gdb_graph_end_capture(stream_, &g);
        if (g) gdb_graph_destroy(g);
        throw;
    }
    check(gdb_graph_end_capture(stream_, &graph_), "end capture");
}

float PipelineBatch::run(bool useGraph) {
    check(gdb_event_record(ev0_, stream_), "event");
    if (useGraph) {
        if (!graph_) capture();
        check(gdb_graph_launch(graph_, stream_), "graph launch");
    } else {
        enqueue();
    }
    check(gdb_event_record(ev1_, stream_), "event");
    check(gdb_event_sync(ev1_), "event sync");
    float ms = 0.0f;
    check(gdb_event_elapsed_ms(ev0_, ev1_, &ms), "elapsed");
    for (size_t i = 0; i < statusRefs_.size(); ++i) HipGenericExecutor::judgeStatus(statusRefs_[i], hStatus_ + 4 * i);
    return ms;
}

}  // namespace runtime
}  // namespace gdb

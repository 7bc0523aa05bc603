#pragma once
// HipGenericExecutor — binding-driven runner over the C-ABI.
//
// Counterpart of MetalGenericExecutor (codegen/execution/metal_generic_executor.h:46-164):
// resolves symbolic sizes, allocates and initialises runner-owned buffers,
// packs kernel arguments from the phase bindings, chooses the grid, launches
// every phase on one stream and times phases with events. Differences:
//   - buffer initialisation is a fill kernel on the stream (graph-safe), not
//     a host memset of shared memory;
//   - buffers are allocated lazily per phase, so a phase hook (a sizing
//     pre-pass that registers a symbol) can size later phases;
//   - enqueue() only launches (no sync, no events) so a whole plan can be
//     captured in a graph or overlapped with other queries;
//   - kernels report capacity overflow through status words that execute()
//     checks and raises as GdbOverflowError.

#include "../core/hip_codegen_base.h"
#include "gdb/gdb_hip.h"
#include "hip_result_collector.h"
#include "runtime_compiler.h"

#include <array>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace gdb {
namespace codegen {

struct GdbOverflowError : public std::runtime_error {
    std::string phase;
    uint64_t inserted = 0;
    GdbOverflowError(const std::string& ph, uint64_t ins)
        : std::runtime_error("hash table overflow in phase " + ph), phase(ph), inserted(ins) {}
};

struct HipExecutionResult {
    float totalKernelTimeMs = 0.0f;          // measured run: first fill -> last kernel end (incl. host hooks)
    std::vector<float> phaseTimesMs;         // measured run, per phase: the phase kernel alone
    std::vector<float> phaseFillMs;          // measured run, per phase: buffer-init fills before the kernel
    std::vector<std::string> phaseNames;
    std::vector<std::array<uint64_t, 4>> phaseStatus;  // status words of phases that have them
    float bufferAllocTimeMs = 0.0f;
    std::vector<unsigned> phaseGrid, phaseBlock, phaseItems;
};

class HipGenericExecutor {
public:
    HipGenericExecutor(gdb_ctx ctx, gdb_stream stream = nullptr);
    ~HipGenericExecutor();
    HipGenericExecutor(const HipGenericExecutor&) = delete;
    HipGenericExecutor& operator=(const HipGenericExecutor&) = delete;

    // Tables: row count (also symbol n_<table>) and device column pointers
    // (not owned). Replaces registerTableBuffer/registerTableRowCount.
    void registerTable(const std::string& table, uint64_t rows);
    void registerTableColumn(const std::string& table, const std::string& column, const void* devptr);
    // External device buffer referenced by name (not owned).
    void registerBuffer(const std::string& name, void* devptr, size_t bytes);
    // Let a phase scan a runner-owned buffer as a table column.
    void registerBufferTable(const std::string& table, const std::string& column, const std::string& buffer,
                             const std::string& rowsExpr);
    void registerSymbol(const std::string& name, uint64_t value);
    bool tryGetSymbol(const std::string& name, uint64_t& out) const;
    void registerScalarInt(const std::string& name, int64_t value);
    // Host callback after a phase completes (the runner synchronises first).
    // With readBytes > 0, bytes [readOffset, readOffset + readBytes) of
    // buffer readBuffer are copied behind the phase into pinned memory in
    // the same stream, so one synchronisation serves both; the hook finds
    // them at hookBytes().
    void setPhaseHook(const std::string& phase, std::function<void(HipGenericExecutor&)> hook,
                      const std::string& readBuffer = "", size_t readOffset = 0, size_t readBytes = 0);
    void clearPhaseHooks() { hooks_.clear(); }
    const void* hookBytes() const { return pinned_; }
    // When true, buffers are not re-initialised (accumulate across runs).
    void setSkipInit(bool skip) { skipInit_ = skip; }
    // Per-phase device times (an event pair around every phase) or only the
    // whole run's (one event at each end): a query that needs no breakdown
    // keeps the stream free of the per-phase event records.
    void setPhaseTiming(bool on) { phaseTiming_ = on; }
    // Symbols a phase hook sets during a run: buffers sized by them are
    // initialised at their phase, the others all at the start of a run in
    // one multi-buffer fill.
    void addLateSymbol(const std::string& name) { lateSymbols_.push_back(name); }
    void clearLateSymbols() { lateSymbols_.clear(); }

    // Allocate every buffer of phases [first, last) that can be sized now.
    void prepare(const HipCodegen& cg, int firstPhase = 0, int lastPhase = -1);
    // Launch phases [first, last) on the stream: init fills + kernels only.
    // Hooks are ignored; no synchronisation.
    void enqueue(RuntimeCompiler& rc, const RuntimeCompiler::CompiledQuery& cq, const HipCodegen& cg,
                 int firstPhase = 0, int lastPhase = -1);
    // Warm up, then time `measured` runs (results of the last run remain).
    HipExecutionResult execute(RuntimeCompiler& rc, const RuntimeCompiler::CompiledQuery& cq, const HipCodegen& cg,
                               int warmupRuns = 0, int measuredRuns = 1, int firstPhase = 0, int lastPhase = -1);
    // Pieces of a launch, for callers that fuse phases of several plans into
    // one kernel: allocate the phase's buffers and return its packed
    // arguments (8 bytes each, binding order) without launching; and enqueue
    // the initialisation fills of every buffer of the plan.
    std::vector<uint64_t> packPhaseArgs(const HipCodegen& cg, int phase);
    void enqueueInits(const HipCodegen& cg);
    // The same fills gathered by a caller that batches many runners.
    struct FillList {
        std::vector<void*> dst;
        std::vector<size_t> bytes;
        std::vector<std::array<uint8_t, 32>> pattern;  // the pattern repeated to 32 bytes
        std::vector<int> plen;                          // its length (1, 2, 4 bytes or 1..4 words)
        struct Single {                                 // fills the multi-buffer fill cannot take
            void* dst;
            size_t bytes;
            std::vector<uint8_t> pattern;
        };
        std::vector<Single> singles;
    };
    // Every initialisation of cg's buffers (allocating them); the caller
    // issues them (gdb_fill_multi_async / gdb_fill_table_async for the
    // lists, gdb_fill_async for each single).
    void collectInits(const HipCodegen& cg, FillList& out);
    // Device addresses of the phases' status and check words (4 x u64 each)
    // for a caller that reads them in bulk, and how to judge them.
    struct StatusRef {
        const void* ptr = nullptr;
        std::string phase;
        bool check = false;       // a check word (failure message) rather than an overflow status
        std::string message;
    };
    std::vector<StatusRef> statusRefs(const HipCodegen& cg) const;
    // Throws as checkStatus does for the words of one StatusRef.
    static void judgeStatus(const StatusRef& r, const uint64_t* words);
    // Read the status words of the plan's phases; throws GdbOverflowError.
    std::vector<std::array<uint64_t, 4>> checkStatus(const HipCodegen& cg, int firstPhase = 0, int lastPhase = -1);
    // checkStatus in two halves: the gather and the copy into pinned memory
    // enqueued on the stream, then (after the caller's synchronisation) the
    // words judged.
    struct PendingStatus {
        std::vector<std::pair<int, StatusRef>> refs;  // (phase index in the range, reference)
        const uint64_t* host = nullptr;
        size_t phases = 0;
    };
    PendingStatus enqueueStatus(const HipCodegen& cg, int firstPhase = 0, int lastPhase = -1);
    std::vector<std::array<uint64_t, 4>> judgeStatuses(const PendingStatus& ps) const;

    GenericResult collectResult(const HipCodegen& cg) const;
    void releaseAllocatedBuffers();
    void* getAllocatedBuffer(const std::string& name) const;
    size_t getAllocatedBytes(const std::string& name) const;
    void readBuffer(const std::string& name, size_t offset, void* host, size_t bytes) const;
    uint64_t resolve(const std::string& expr) const { return sizes_.resolve(expr); }
    size_t ownedBytes() const;

    gdb_stream stream() const { return stream_; }
    gdb_ctx context() const { return ctx_; }
    int cuCount() const { return cuCount_; }

private:
    struct Buffer {
        void* ptr = nullptr;
        size_t bytes = 0;
        bool owned = true;
    };
    struct TableInfo {
        uint64_t rows = 0;
        std::unordered_map<std::string, const void*> columns;
    };
    struct BufferTable {
        std::string table, column, buffer, rowsExpr;
    };

    gdb_ctx ctx_;
    gdb_stream stream_;
    int cuCount_ = 0;
    HipSizeResolver sizes_;
    std::unordered_map<std::string, TableInfo> tables_;
    std::unordered_map<std::string, Buffer> buffers_;
    std::vector<BufferTable> bufferTables_;
    std::unordered_map<std::string, int64_t> scalars_;
    struct Hook {
        std::function<void(HipGenericExecutor&)> fn;
        std::string buffer;
        size_t offset = 0, bytes = 0;
    };
    std::unordered_map<std::string, Hook> hooks_;
    // pinned host staging (hook reads, status words) and the device side of
    // the bulk status read: the status buffers' addresses and the gathered words
    uint64_t* pinned_ = nullptr;
    size_t pinnedBytes_ = 0;
    void* dStatusPtrs_ = nullptr;
    void* dStatusOut_ = nullptr;
    size_t statusCap_ = 0;
    std::vector<const void*> statusPtrCache_;
    uint64_t* ensurePinned(size_t bytes);
    std::map<std::pair<gdb_kernel, int>, int> occupancy_;
    bool skipInit_ = false;
    bool phaseTiming_ = true;
    std::vector<std::string> lateSymbols_;
    gdb_event evA_ = nullptr;
    std::vector<gdb_event> events_;
    std::vector<float> measuredPhase_, measuredFill_;

    Buffer& ensureBuffer(const HipParamBinding& b);
    void resolveBufferTables();
    void launchPhase(RuntimeCompiler& rc, const RuntimeCompiler::CompiledQuery& cq, const HipCodegen::PhaseInfo& p,
                     std::unordered_map<std::string, bool>& initDone, const std::unordered_map<std::string, const HipParamBinding*>& initSpec,
                     unsigned* gridOut, gdb_event beforeKernel = nullptr);
    std::unordered_map<std::string, const HipParamBinding*> initSpecs(const HipCodegen& cg) const;
    // Initialise, in one launch, every buffer defined by phases [first, last)
    // whose size is known now; marks them done.
    void hoistInits(const HipCodegen& cg, int first, int last, std::unordered_map<std::string, bool>& initDone,
                    const std::unordered_map<std::string, const HipParamBinding*>& initSpec);
    int occupancy(gdb_kernel k, int block);
};

}  // namespace codegen
}  // namespace gdb

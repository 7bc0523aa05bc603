#pragma once
// ChunkStreamer — out-of-core execution of a chunkable plan whose stream
// table lives in host memory.
//
// Counterpart of the reference's chunked execution (ChunkedColbinTable,
// codegen/execution/chunked_colbin_loader.h:12-60, and the chunk loop of
// codegen/codegen_main.cpp:686-808): phases before the first phase that scans
// the stream table run once, the stream phases run once per row chunk, the
// phases after them run once. A stream phase that needs what earlier stream
// phases accumulated over the whole table (plan phase flag newStreamPass)
// starts a second pass over the chunks; phases between passes run once.
// Differences that follow from the target:
//   - chunks are copied host->HBM with asynchronous copies on a copy stream
//     into two device slots per column (double buffering), so the copy of
//     chunk c+1 overlaps the kernels of chunk c; the compute stream waits on
//     a per-slot "copied" event and the copy stream on a per-slot "consumed"
//     event, and the host never blocks inside the loop;
//   - host columns should be pinned (gdb_host_alloc_pinned) for the copies to
//     run at link speed and overlap; pageable memory works, serialised;
//   - a column given by a producer (a .colbin file read chunk by chunk,
//     chunked_colbin_loader.cpp:167-194) is written into one of two pinned
//     staging slots while the device works on the previous chunk; the host
//     waits only for the copy that last used the slot;
//   - buffers of the stream phases are initialised before the first chunk
//     only (the runner's skip-init for later chunks), and overflow status is
//     checked once after the last chunk.

#include "../execution/hip_generic_executor.h"
#include "../execution/runtime_compiler.h"
#include "../planning/hip_plan_builder.h"
#include "gdb/gdb_hip.h"

#include <functional>
#include <string>
#include <vector>

namespace gdb {
namespace runtime {

struct StreamColumn {
    std::string name;
    const void* host = nullptr;  // the engine-format column in host memory, or
    int elemBytes = 0;
    // ... a producer that writes rows [r0, r0 + n) in the engine's format to
    // dst (pinned staging the streamer owns): file-backed columns converted
    // chunk by chunk (the table is never held whole in host memory)
    std::function<void(void* dst, uint64_t r0, uint64_t n)> produce;
};

struct StreamStats {
    double totalMs = 0;      // host wall time of the whole run (pre + chunks + post), synchronised
    double streamMs = 0;     // first copy issued -> last stream-phase kernel done (device events)
    uint64_t chunks = 0;
    uint64_t bytesCopied = 0;
    int firstStreamPhase = 0, lastStreamPhase = 0;  // [first, last)
    int passes = 0;                                   // passes over the stream table
};

class ChunkStreamer {
public:
    ChunkStreamer(gdb_ctx ctx, codegen::HipGenericExecutor& ex, const codegen::HipQueryPlan& plan,
                  const codegen::HipCodegen& cg, codegen::RuntimeCompiler& rc,
                  const codegen::RuntimeCompiler::CompiledQuery& cq);
    ~ChunkStreamer();
    ChunkStreamer(const ChunkStreamer&) = delete;
    ChunkStreamer& operator=(const ChunkStreamer&) = delete;

    // Run the plan with `table` (rows rows, the given host columns) streamed
    // in chunks of chunkRows rows. Every column the stream phases read must
    // be given. Throws on a non-chunkable plan or an invalid stream table.
    StreamStats run(const std::string& table, uint64_t rows, const std::vector<StreamColumn>& cols,
                    uint64_t chunkRows);

private:
    gdb_ctx ctx_;
    codegen::HipGenericExecutor& ex_;
    const codegen::HipQueryPlan& plan_;
    const codegen::HipCodegen& cg_;
    codegen::RuntimeCompiler& rc_;
    const codegen::RuntimeCompiler::CompiledQuery& cq_;
    gdb_stream copy_ = nullptr;
    gdb_event copied_[2] = {nullptr, nullptr}, consumed_[2] = {nullptr, nullptr};
    gdb_event t0_ = nullptr, t1_ = nullptr;
    std::vector<void*> slots_[2];
    std::vector<void*> staging_[2];  // pinned host slots of produced columns
    void freeSlots();
};

}  // namespace runtime
}  // namespace gdb

#pragma once
// PipelineBatch — many independent pipelines in one persistent fused launch
// per phase level.
//
// No counterpart in the reference, which dispatches one command buffer per
// phase and query (codegen/execution/metal_generic_executor.cpp:338-390).
// Each pipeline is a plan of one or more phases with its own runner (tables,
// scalars, output buffers). The batch emits every distinct plan phase's body
// in range mode (one block processes rows [_rb, _re) of one pipeline's
// phase), packs each pipeline's kernel arguments into a device table, splits
// every pipeline phase into row-range work items, and launches one persistent
// kernel per phase level (phase 0 of every pipeline, then phase 1 of those
// that have one, ...) whose blocks take work items from an atomic ticket until
// none are left; the kernel boundary orders a pipeline's phases. Launch shape
// per pipeline (ITEMS of its first phase) and the block size come from the
// autotuner; the fills and the kernels can be captured in a graph and
// replayed.
#include "../execution/hip_generic_executor.h"
#include "../execution/runtime_compiler.h"
#include "../planning/hip_plan_builder.h"
#include "gdb/gdb_hip.h"

#include <memory>
#include <string>
#include <vector>

namespace gdb {
namespace runtime {

// Source of the persistent batched kernels for a set of pipeline kinds (a
// kind = one plan phase's body, instantiated for each ITEMS it is used with;
// the phase index is the kind's level). Needs no device, so the kernels can
// be emitted and compiled offline.
class BatchKernelSource {
public:
    // Register a pipeline's plan: phase 0 with the given ITEMS, later phases
    // with their default ITEMS. Returns the kind index of every phase.
    std::vector<int> add(const codegen::HipQueryPlan* plan, const codegen::HipCodegen* cg, int items);
    // Every kind's body and one kernel per level: gdb_batch_kernel for level
    // 0, gdb_batch_kernel_l<L> for level L > 0.
    std::string print() const;
    static std::string kernelName(int level);
    size_t kinds() const { return kinds_.size(); }
    int levels() const;
    int level(int kind) const { return kinds_[static_cast<size_t>(kind)].level; }
    // Argument types of a kind's body, in binding order.
    const std::vector<std::string>& argTypes(int kind) const { return kinds_[static_cast<size_t>(kind)].argTypes; }
    const codegen::HipCodegen& kindCodegen(int kind) const { return *kindCg_[static_cast<size_t>(kind)]; }
    int launchBound() const;
    void clear() {
        kinds_.clear();
        kindCg_.clear();
    }

private:
    struct Kind {
        std::string key;   // plan phase source identity
        std::string body;  // range-mode body name
        int level = 0;     // phase index within its plan
        std::vector<std::string> argTypes;
        std::vector<int> items;
    };
    std::vector<Kind> kinds_;
    std::vector<std::unique_ptr<codegen::HipCodegen>> kindCg_;  // range-mode codegen per kind
};

class PipelineBatch {
public:
    explicit PipelineBatch(gdb_ctx ctx);
    ~PipelineBatch();
    PipelineBatch(const PipelineBatch&) = delete;
    PipelineBatch& operator=(const PipelineBatch&) = delete;

    // Register a pipeline: its plan, the generated codegen of that plan, and
    // the runner holding its registrations. `items` (phase 0) must be one of
    // the phase's ITEMS variants. Returns the pipeline index.
    int add(const codegen::HipQueryPlan* plan, const codegen::HipCodegen* cg, codegen::HipGenericExecutor* ex,
            int items);
    // Emit + compile the fused kernel, pack arguments, build work items.
    // chunkRows: target rows per work item (rounded to whole tiles); grid:
    // 0 = CUs x occupancy.
    void build(int block, uint64_t chunkRows, int grid);
    // Rows per work item of one phase level from the next build() on,
    // instead of chunkRows (0: back to chunkRows). A fold level over partial
    // tables has few rows per pipeline: smaller items keep more CUs busy.
    void setLevelChunk(int level, uint64_t rows);
    // Fills of every pipeline's buffers + the fused kernels, on the stream.
    void enqueue();
    void capture();                         // record enqueue() into a graph
    // Run (graph replay when captured and useGraph), synchronise, check the
    // pipelines' status words; returns the elapsed ms between events around
    // the launch.
    float run(bool useGraph);

    const std::string& source() const { return source_; }
    // Work items over all levels; grid of the level-0 kernel.
    size_t workItems() const;
    unsigned grid() const { return levels_.empty() ? 0u : levels_[0].grid; }

private:
    struct Pipe {
        const codegen::HipQueryPlan* plan;
        const codegen::HipCodegen* cg;
        codegen::HipGenericExecutor* ex;
        std::vector<int> items;    // per phase
        std::vector<int> kinds;    // per phase
        std::vector<uint64_t> rows;  // per phase
    };
    // One persistent kernel: the work items of one phase level.
    struct Level {
        gdb_kernel kernel = nullptr;
        unsigned grid = 0;
        uint32_t argStride = 0;
        std::vector<uint32_t> itemPipe;
        std::vector<uint64_t> itemBegin, itemEnd;
        void* dArgs = nullptr;
        void* dKindItems = nullptr;
        void* dItemPipe = nullptr;
        void* dItemBegin = nullptr;
        void* dItemEnd = nullptr;
    };
    gdb_ctx ctx_;
    gdb_stream stream_;
    codegen::RuntimeCompiler rc_;
    codegen::RuntimeCompiler::CompiledQuery cq_;
    bool built_ = false;
    std::vector<Pipe> pipes_;
    BatchKernelSource src_;
    std::string source_;
    int block_ = 256;
    std::vector<Level> levels_;
    std::vector<uint64_t> levelChunk_;      // per level; 0 = chunkRows
    void* dTicket_ = nullptr;
    // every pipeline's buffer initialisations and the work ticket, as a fill
    // table in device memory (the buffers are fixed at build): one launch per
    // replay; fills the table cannot express are issued one by one
    void* dFill_ = nullptr;
    int fillN_ = 0;
    uint64_t fillMaxWords_ = 0;
    std::vector<codegen::HipGenericExecutor::FillList::Single> fillSingles_;
    // every pipeline's status and check words: gathered on the device after
    // the fused kernel and read back with one copy into pinned memory
    std::vector<codegen::HipGenericExecutor::StatusRef> statusRefs_;
    void* dStatusSrc_ = nullptr;
    void* dStatusOut_ = nullptr;
    uint64_t* hStatus_ = nullptr;
    gdb_graph graph_ = nullptr;
    gdb_event ev0_ = nullptr, ev1_ = nullptr;
    void freeDevice();
};

}  // namespace runtime
}  // namespace gdb

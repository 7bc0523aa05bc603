#pragma once
// PipelineBatch — many independent pipelines in one persistent fused launch.
//
// No counterpart in the reference, which dispatches one command buffer per
// phase and query (codegen/execution/metal_generic_executor.cpp:338-390).
// Each pipeline is a single-phase plan with its own runner (tables, scalars,
// output buffers). The batch emits every distinct plan's phase body in range
// mode (one block processes rows [_rb, _re) of one pipeline), packs each
// pipeline's kernel arguments into a device table, splits every pipeline into
// row-range work items, and launches one persistent kernel whose blocks take
// work items from an atomic ticket until none are left. Launch shape per
// pipeline (ITEMS) and the block size come from the autotuner; the fills and
// the kernel can be captured in a graph and replayed.

#include "../execution/hip_generic_executor.h"
#include "../execution/runtime_compiler.h"
#include "../planning/hip_plan_builder.h"
#include "gdb/gdb_hip.h"

#include <memory>
#include <string>
#include <vector>

namespace gdb {
namespace runtime {

// Source of the persistent batched kernel for a set of pipeline kinds (a
// kind = one plan's phase body, instantiated for each ITEMS it is used with).
// Needs no device, so the kernel can be emitted and compiled offline.
class BatchKernelSource {
public:
    // Register a pipeline's plan and ITEMS; returns the kind index.
    int add(const codegen::HipQueryPlan* plan, const codegen::HipCodegen* cg, int items);
    std::string print() const;
    size_t kinds() const { return kinds_.size(); }
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
        std::string key;   // plan source identity
        std::string body;  // range-mode body name
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

    // Register a pipeline: its plan (one phase), the generated codegen of
    // that plan, and the runner holding its registrations. `items` must be
    // one of the phase's ITEMS variants. Returns the pipeline index.
    int add(const codegen::HipQueryPlan* plan, const codegen::HipCodegen* cg, codegen::HipGenericExecutor* ex,
            int items);
    // Emit + compile the fused kernel, pack arguments, build work items.
    // chunkRows: target rows per work item (rounded to whole tiles); grid:
    // 0 = CUs x occupancy.
    void build(int block, uint64_t chunkRows, int grid);
    // Fills of every pipeline's buffers + the fused kernel, on the stream.
    void enqueue();
    void capture();                         // record enqueue() into a graph
    // Run (graph replay when captured and useGraph), synchronise, check the
    // pipelines' status words; returns the elapsed ms between events around
    // the launch.
    float run(bool useGraph);

    const std::string& source() const { return source_; }
    size_t workItems() const { return itemPipe_.size(); }
    unsigned grid() const { return grid_; }

private:
    struct Pipe {
        const codegen::HipQueryPlan* plan;
        const codegen::HipCodegen* cg;
        codegen::HipGenericExecutor* ex;
        int items;
        int kind;
        uint64_t rows;
    };
    gdb_ctx ctx_;
    gdb_stream stream_;
    codegen::RuntimeCompiler rc_;
    codegen::RuntimeCompiler::CompiledQuery cq_;
    gdb_kernel kernel_ = nullptr;
    std::vector<Pipe> pipes_;
    BatchKernelSource src_;
    std::string source_;
    int block_ = 256;
    unsigned grid_ = 0;
    uint32_t argStride_ = 0;
    std::vector<uint32_t> itemPipe_;
    std::vector<uint64_t> itemBegin_, itemEnd_;
    void* dArgs_ = nullptr;
    void* dKindItems_ = nullptr;
    void* dItemPipe_ = nullptr;
    void* dItemBegin_ = nullptr;
    void* dItemEnd_ = nullptr;
    void* dTicket_ = nullptr;
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

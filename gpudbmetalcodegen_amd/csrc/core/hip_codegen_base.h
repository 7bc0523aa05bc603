#pragma once
// HipCodegen — composable HIP C++ emission for gfx950.
//
// Counterpart of MetalCodegen (codegen/core/metal_codegen_base.h:27-184):
// code emission with indentation, helper functions emitted once, one kernel
// per phase, parameter bindings and the result schema. Differences that
// follow from the target:
//   - each phase body is a template over ITEMS (rows per thread per tile);
//     print() instantiates one extern "C" kernel per requested ITEMS value,
//     named <phase>_i<ITEMS>, so the runner and the autotuner pick the variant
//     at launch time without recompiling;
//   - bindings become positional kernel arguments (natural alignment) rather
//     than [[buffer(N)]] slots;
//   - the device runtime (gdb_device_header.h) replaces the Metal common header.

#include "hip_param_binding.h"

#include <functional>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace gdb {
namespace codegen {

using ConsumerFn = std::function<void()>;

// Column element types of the catalog. CHAR columns are fixed-width
// NUL-padded byte strings (width in ColumnType::width).
enum class ColType { I32, I64, U8, CHAR };

struct ColumnType {
    ColType type = ColType::I32;
    int width = 0;  // CHAR only
    size_t bytes() const {
        switch (type) {
            case ColType::I32: return 4;
            case ColType::I64: return 8;
            case ColType::U8: return 1;
            case ColType::CHAR: return static_cast<size_t>(width);
        }
        return 4;
    }
    std::string cxx() const {
        switch (type) {
            case ColType::I32: return "i32";
            case ColType::I64: return "i64";
            case ColType::U8: return "u8";
            case ColType::CHAR: return "char";
        }
        return "i32";
    }
    std::string tag() const {
        switch (type) {
            case ColType::I32: return "i32";
            case ColType::I64: return "i64";
            case ColType::U8: return "u8";
            case ColType::CHAR: return "char" + std::to_string(width);
        }
        return "i32";
    }
};

class HipCodegen {
public:
    // --- code emission ---
    void addLine(const std::string& line);
    void addBlock(const std::string& header, const std::function<void()>& body, const std::string& trailing = "");
    void addIf(const std::string& condition, const std::function<void()>& body);
    void addComment(const std::string& comment);
    void increaseIndent() { ++indentLevel_; }
    void decreaseIndent() {
        if (indentLevel_ > 0) --indentLevel_;
    }
    void addRawCode(const std::string& code);

    // Helper device functions / constant tables emitted once before all kernels.
    void addHelper(const std::string& code);

    // --- phases: one kernel (family) per phase ---
    void beginPhase(const std::string& phaseName);
    void endPhase();
    int phaseCount() const { return static_cast<int>(phases_.size()); }

    void setPhaseScannedTable(const std::string& tableName);
    void setPhaseRowsExpr(const std::string& rowsExpr);
    void setPhaseThreadgroupSize(int size);
    void setPhaseSingleThread(bool single);
    void setPhaseMaxThreadgroups(int max);
    void setPhaseItems(int items);
    void setPhaseItemsVariants(const std::vector<int>& items);
    void setPhaseLaunchBound(int bound);
    void setPhaseStatusBuffer(const std::string& name);
    // A row check's status words (HipRowCheck): the run fails with `message`
    // when any row of the phase failed it.
    void setPhaseCheck(const std::string& checkBuffer, const std::string& message);

    // --- parameters ---
    // Column of a registered table; returns the pointer parameter name.
    std::string addColumnParam(const std::string& table, const std::string& column, const ColumnType& t);
    // Row count of a table; returns the parameter name n_<table>.
    std::string addTableSizeParam(const std::string& table);
    void addBufferParam(const std::string& name, const std::string& elemType, size_t elemBytes,
                        const std::string& sizeExpr, InitKind init = InitKind::Zero,
                        const std::vector<uint8_t>& pattern = {}, bool readOnly = false, bool persistent = false);
    void addScalarParam(const std::string& name);
    // Parameter carrying the resolved value of a size expression.
    void addSymbolParam(const std::string& name, const std::string& sizeExpr);
    void addBitmapWriteParam(const std::string& name, const std::string& bitsExpr);
    void addBitmapReadParam(const std::string& name, const std::string& bitsExpr);

    // Block-shared (LDS) array `name` of countExpr elements of `type`
    // (countExpr is a constant expression and may use ITEMS). Grid mode: a
    // __shared__ array. Range mode: a 16-byte aligned slice of the batched
    // kernel's single LDS buffer (_lds), so the LDS of the pipeline bodies of
    // one kernel overlap instead of adding up.
    void declareLds(const std::string& type, const std::string& name, const std::string& countExpr);

    // --- staged (item-batched) emission; see hip_operators.h ---
    void setPhaseStaged(bool staged);
    void setPhasePrefetch(bool on) { cur("setPhasePrefetch").prefetch = on; }
    bool phaseStaged() const;
    // True while emitting for a batch of ITEMS rows with active mask _act.
    bool staged() const { return staged_; }
    void setStaged(bool on) { staged_ = on; }
    // Lines that bind a row's locals for item _u (columns, computed values,
    // probe results); emitted at the top of every loop over the items.
    void addRowBinding(const std::string& line) { rowBindings_.push_back(line); }
    std::vector<std::string>& rowBindings() { return rowBindings_; }
    // One unrolled loop over the active items with the row bindings in
    // scope; body() runs in per-row context.
    void rowLoop(const std::function<void()>& body, const std::string& mask = "_act");
    // Whether the operator about to be produced consumes staged batches (set
    // by the parent operator before it produces its child).
    bool nextConsumerStageable() const { return nextConsumerStageable_; }
    void setNextConsumerStageable(bool v) { nextConsumerStageable_ = v; }

    // Fresh identifier with the given prefix (unique within this codegen).
    std::string uniqueName(const std::string& prefix);

    // Range mode: scans cover rows [_rb, _re) with one block (a work item of
    // a persistent batched kernel) instead of the whole table with the grid.
    // The phase body then takes (..., const u64 _rb, const u64 _re).
    void setRangeMode(bool on) { rangeMode_ = on; }
    bool rangeMode() const { return rangeMode_; }
    // Segment scans: like range mode for the scan alone (rows [_rb, _re)
    // walked by this block), with _rb/_re defined by the consuming operator's
    // own loop (logical blocks of a multi-kernel operator).
    void setScanSegment(bool on) { scanSegment_ = on; }
    bool scanSegment() const { return scanSegment_; }
    // The phase body alone (template <int ITEMS> __device__ function named
    // <bodyName>), without helpers, runtime header or kernel wrappers. In
    // range mode the body also takes (unsigned char* _lds) last, and a
    // constexpr <bodyName>_lds_bytes<ITEMS>() gives the LDS bytes it uses.
    std::string printPhaseBody(int phase, const std::string& bodyName) const;
    std::string printHelpers() const;

    // Size expressions of runner-owned buffers (first registration wins).
    const std::unordered_map<std::string, std::string>& getGlobalBufferSizes() const { return bufferSizes_; }

    // --- result schema ---
    HipResultSchema& resultSchema() { return resultSchema_; }
    const HipResultSchema& getResultSchema() const { return resultSchema_; }

    // Full HIP source: device runtime, helpers, then every phase.
    std::string print() const;

    struct PhaseInfo {
        std::string name;
        std::string code;
        std::string scannedTable;
        std::string rowsExpr;      // rows the grid covers (resolved by the runner)
        int threadgroupSize = 256;
        int maxThreadgroups = 0;   // 0 = occupancy-derived
        bool isSingleThread = false;
        int items = 4;             // default ITEMS variant to launch
        std::vector<int> itemsVariants{4};
        int launchBound = 1024;
        std::string statusBuffer;  // optional overflow/status words checked after the phase
        std::string checkBuffer;   // optional row check's words (failed flag, -, failed rows) and its message
        std::string checkMessage;
        std::string ldsBytesExpr = "0u";  // range mode: LDS bytes of the body (constant expression)
        bool staged = false;              // item-batched emission (hip_operators.h)
        bool prefetch = false;            // scans load the next tile before processing this one
        std::vector<HipParamBinding> bindings;
    };
    const std::vector<PhaseInfo>& getPhases() const { return phases_; }
    std::vector<PhaseInfo>& getPhasesMutable() { return phases_; }
    std::vector<HipParamBinding> getAllBindings() const;

    // Kernel name of a phase's ITEMS variant.
    static std::string kernelName(const std::string& phase, int items) {
        return phase + "_i" + std::to_string(items);
    }

private:
    std::vector<PhaseInfo> phases_;
    int current_ = -1;
    std::vector<std::string> helpers_;
    std::unordered_set<std::string> helperSet_;
    std::unordered_map<std::string, std::string> bufferSizes_;
    std::unordered_map<std::string, int> nameCounters_;
    HipResultSchema resultSchema_;
    unsigned indentLevel_ = 1;
    bool rangeMode_ = false;
    bool scanSegment_ = false;
    bool staged_ = false;
    bool nextConsumerStageable_ = false;
    std::vector<std::string> rowBindings_;

    PhaseInfo& cur(const char* op);
    void pushBinding(const char* op, HipParamBinding b);
    std::string indent() const { return std::string(indentLevel_ * 4, ' '); }
    static std::string paramList(const PhaseInfo& p, bool withTypes);
};

// RAII indentation scope (same role as the reference's IndentGuard).
class IndentGuard {
public:
    explicit IndentGuard(HipCodegen& cg) : cg_(cg) { cg_.increaseIndent(); }
    ~IndentGuard() { cg_.decreaseIndent(); }
    IndentGuard(const IndentGuard&) = delete;
    IndentGuard& operator=(const IndentGuard&) = delete;

private:
    HipCodegen& cg_;
};

}  // namespace codegen
}  // namespace gdb

#pragma once
// HipQueryPlan and its builders.
//
// Counterpart of codegen/planning/metal_plan_builder.h:22-78: a plan is a list
// of phases, each an operator tree that becomes one kernel; generateFromPlan()
// runs the emitters. There is no SQL front end (the reference's needs
// libpg_query); builders are selected by plan name and take shape options
// (variant, LDS slots, ITEMS, block size) as key=value pairs. Runtime values
// (cutoff dates, segment ids, LIKE words) are kernel scalars or registered
// buffers, never baked into the source, so one compiled plan serves every
// query of its shape and the code-object cache stays small.

#include "../operators/hip_operators.h"

#include <map>
#include <memory>
#include <optional>
#include <string>
#include <vector>

namespace gdb {
namespace codegen {

struct HipQueryPlan {
    std::string name;

    struct Phase {
        std::string name;
        std::unique_ptr<HipOperator> root;
        int threadgroupSize = 256;
        int items = 4;
        std::vector<int> itemsVariants{4};
        int launchBound = 1024;
        int maxThreadgroups = 0;
        bool singleThread = false;
        // Item-batched emission: the random loads (bitmap words, first hash
        // slots) of a thread's ITEMS rows are issued together.
        bool staged = false;
        // Software-pipelined scan: the next tile's columns load while this
        // tile is processed.
        bool prefetch = false;
        // Out-of-core runs: this phase scans the stream table after earlier
        // stream phases have accumulated over every chunk (a second pass).
        bool newStreamPass = false;
        std::vector<std::string> scalarParams;  // registered before the operators run
        // Buffers the caller registers (e.g. LIKE words, prebuilt bitmaps):
        // (name, element type) read-only parameters.
        std::vector<std::pair<std::string, std::string>> externalBuffers;
    };
    std::vector<Phase> phases;
    std::vector<std::string> helpers;
    // Out-of-core eligible (metal_plan_builder.cpp:2078-2107's allowlist):
    // every phase that scans the stream table updates its outputs only by
    // associative atomics, so running those phases once per row chunk (with
    // buffers initialised before the first chunk only) gives the result of
    // one run over the whole table.
    bool chunkable = false;

    // Runner-owned buffers that a later phase scans as a table column:
    // (table, column, buffer, rows expression).
    struct BufferTable {
        std::string table, column, buffer, rowsExpr;
    };
    std::vector<BufferTable> bufferTables;

    // Exact-count pre-pass sizing: after `afterPhase` completes, the runner
    // reads u64 word `word` of `buffer` (a count), computes
    // max(minimum, count * multiplier), rounds up to a power of two when
    // pow2 is set, and registers it as `symbol` for the later phases.
    // Counterpart of registerMaxKeySymbols (max_key_symbols.cpp:188-235).
    struct SizingRule {
        std::string afterPhase, symbol, buffer;
        int word = 0;
        uint64_t multiplier = 2;
        uint64_t minimum = 64;
        bool pow2 = true;
    };
    std::vector<SizingRule> sizingRules;

    // Scalars the caller must register before execute (documentation + checks).
    std::vector<std::string> requiredScalars;
    // Size symbols the caller (or a sizing pre-pass) must register.
    std::vector<std::string> requiredSymbols;

    // Plans built from SQL text: how the caller finishes the result rows
    // (JSON; sql_frontend.h). Empty for named plans.
    std::string sqlLayout;

    std::string describe() const;  // the --print-plan view
};

using PlanOptions = std::map<std::string, std::string>;
PlanOptions parsePlanOptions(const std::string& text);  // "k=v,k=v"

// Catalog of the synthetic schema: table -> ordered (column, type).
struct CatalogTable {
    std::string name;
    std::vector<std::pair<std::string, ColumnType>> columns;
    const ColumnType& type(const std::string& col) const;
};
const CatalogTable& catalogTable(const std::string& name);
std::vector<std::string> catalogTableNames();

// Build a plan by name. Known names:
//   Q1      filter + 6-group exact aggregate over F        (cfg 0; reference Q1)
//   Q6      filtered scalar revenue over F                 (reference Q6 / MB)
//   Q3      A-bitmap -> B hash build -> F probe + per-key revenue -> top-10
//           (cfg 1; reference Q3 with a real hash join)
//   Q3_SIZE sizing pre-pass for Q3: A bitmap + exact count of qualifying B rows
//   GROUPBY hash group-by sum/count/min/max over G         (cfg 2; Q18 analog)
//   GB_COUNT exact distinct-count pre-pass for GROUPBY (key bitmap + popcount)
//   MAXKEY  min/max of a key column                        (max_key_symbols.cpp)
//   Q13L    LIKE '%w1%w2%' per-key count over S, then histogram (cfg 4; Q13)
//   Q22A    anti-join: A rows with no qualifying B row, per segment (Q22)
// Returns nullopt for an unknown name.
std::optional<HipQueryPlan> buildHipPlan(const std::string& name, const PlanOptions& opts = {});
std::vector<std::string> knownPlanNames();

// Run the emitters: helpers, then one phase per plan phase. rangeMode emits
// scans over a work item [_rb, _re) (for the persistent batched kernel).
HipCodegen generateFromPlan(const HipQueryPlan& plan, bool rangeMode = false);

}  // namespace codegen
}  // namespace gdb

#include "autotuner.h"

#include <algorithm>
#include <limits>

namespace gdb {
namespace runtime {

using namespace codegen;

TuneResult autotunePhase(HipGenericExecutor& ex, RuntimeCompiler& rc, const RuntimeCompiler::CompiledQuery& cq,
                         HipCodegen& cg, int phase, const std::vector<int>& blocks, const std::vector<int>& items,
                         const std::string& scalarName, const std::vector<int64_t>& scalarValues, int trials) {
    auto& ph = cg.getPhasesMutable().at(static_cast<size_t>(phase));
    const int saveBlock = ph.threadgroupSize, saveItems = ph.items;
    std::vector<int64_t> svals = scalarName.empty() ? std::vector<int64_t>{0} : scalarValues;
    TuneResult best;
    best.ms = std::numeric_limits<float>::infinity();
    for (int b : blocks) {
        if (b < 64 || b > ph.launchBound || b % 64) continue;
        for (int it : items) {
            if (std::find(ph.itemsVariants.begin(), ph.itemsVariants.end(), it) == ph.itemsVariants.end()) continue;
            for (int64_t sv : svals) {
                ph.threadgroupSize = b;
                ph.items = it;
                if (!scalarName.empty()) ex.registerScalarInt(scalarName, sv);
                double p50 = std::numeric_limits<double>::infinity();
                try {
                    ex.execute(rc, cq, cg, 1, 1, phase, phase + 1);  // untimed warmup
                    std::vector<double> t;
                    for (int k = 0; k < std::max(2, trials); ++k) {
                        auto r = ex.execute(rc, cq, cg, 0, 1, phase, phase + 1);
                        t.push_back(r.phaseTimesMs.empty() ? 0.0 : r.phaseTimesMs[0]);
                    }
                    std::sort(t.begin(), t.end());
                    t.pop_back();  // drop the slowest
                    p50 = t[t.size() / 2];
                } catch (const std::exception&) {
                    // a shape the kernel cannot run (e.g. too much LDS) is skipped
                }
                best.table.push_back({static_cast<double>(b), static_cast<double>(it), static_cast<double>(sv), p50});
                if (p50 < best.ms) {
                    best.ms = static_cast<float>(p50);
                    best.block = b;
                    best.items = it;
                    best.scalar = sv;
                }
            }
        }
    }
    ph.threadgroupSize = best.block ? best.block : saveBlock;
    ph.items = best.items ? best.items : saveItems;
    if (!scalarName.empty() && best.block) ex.registerScalarInt(scalarName, best.scalar);
    return best;
}

}  // namespace runtime
}  // namespace gdb

#pragma once
// Small helpers shared by the plan builders (option lookup, catalog scans,
// phase append).

#include "hip_plan_builder.h"

#include <string>
#include <vector>

namespace gdb {
namespace codegen {

inline std::string opt(const PlanOptions& o, const std::string& k, const std::string& def) {
    auto it = o.find(k);
    return it == o.end() ? def : it->second;
}

inline int optInt(const PlanOptions& o, const std::string& k, int def) {
    auto it = o.find(k);
    return it == o.end() ? def : std::stoi(it->second);
}

// Scan of catalog table `table` reading the named columns.
inline std::unique_ptr<HipGridStrideScan> makeScan(const std::string& table, std::initializer_list<const char*> cols) {
    auto scan = std::make_unique<HipGridStrideScan>(table, "i");
    const auto& t = catalogTable(table);
    for (const char* c : cols) scan->addColumn(c, t.type(c));
    return scan;
}

// Same, with the named char columns read in place (HipGridStrideScan::setLazy).
inline std::unique_ptr<HipGridStrideScan> makeScanLazy(const std::string& table,
                                                       std::initializer_list<const char*> cols,
                                                       std::initializer_list<const char*> lazy) {
    auto scan = makeScan(table, cols);
    for (const char* c : lazy) scan->setLazy(c);
    return scan;
}

inline HipQueryPlan::Phase& appendPhase(HipQueryPlan& plan, const std::string& name,
                                        std::unique_ptr<HipOperator> root, int block, int items,
                                        std::vector<int> variants = {}) {
    HipQueryPlan::Phase p;
    p.name = name;
    p.root = std::move(root);
    p.threadgroupSize = block;
    p.items = items;
    p.itemsVariants = variants.empty() ? std::vector<int>{items} : variants;
    bool has = false;
    for (int v : p.itemsVariants) has |= (v == items);
    if (!has) p.itemsVariants.push_back(items);
    plan.phases.push_back(std::move(p));
    return plan.phases.back();
}

}  // namespace codegen
}  // namespace gdb

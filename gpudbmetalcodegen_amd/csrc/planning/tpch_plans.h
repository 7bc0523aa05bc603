#pragma once
// Named TPC-H plans over the TPC-H tables as loaded from .colbin files, built
// from the same operators as the synthetic workloads.
//
// Q1, Q3 and Q6 share their names with the synthetic workload plans and are
// selected with the option schema=tpch; Q4, Q12, Q13, Q14, Q18 and Q22 exist
// only over TPC-H. Constants of the SQL text (dates, segment, ship modes,
// LIKE words, country codes, quantity bound) are kernel scalars or a
// registered byte buffer; the caller computes them from the query text.

#include "hip_plan_builder.h"

#include <optional>
#include <string>
#include <vector>

namespace gdb {
namespace codegen {

// nullopt when name/options do not select a TPC-H plan.
std::optional<HipQueryPlan> buildTpchPlan(const std::string& name, const PlanOptions& opts);
std::vector<std::string> tpchPlanNames();

}  // namespace codegen
}  // namespace gdb

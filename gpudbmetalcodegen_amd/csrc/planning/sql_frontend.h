#pragma once
// SQL front end: SELECT text -> analyzed query -> operator plan.
//
// Counterpart of the reference's analyzeSQL + the analysis-based builders
// (codegen/planning/query_analyzer.cpp:640-727, metal_plan_builder.cpp:253-360),
// without libpg_query: a hand-written lexer and recursive-descent parser for
// the SELECT subset those builders consume. Queries over TPC-H tables joined
// in a tree of key equalities are planned from the SQL itself: WHERE predicates and aggregate
// arguments are compiled to HIP expressions in exact fixed-point integers
// (money columns are hundredths, literals keep their decimal digits, scales
// are aligned at + - and added at *). Other queries are planned by name
// (buildTpchPlan), as the reference's name-based dispatch does for them.

#include "hip_plan_builder.h"

#include <string>

namespace gdb {
namespace codegen {

// Plan a SELECT over tables joined in a tree of key equalities, with [NOT]
// EXISTS / [NOT] IN subqueries as semi- / anti-joins. Options: groups=N (dense group capacity of a GROUP BY; default what
// fits register accumulators, at most 8; an overflow asks for a rebuild with
// the exact count).
// The plan's sqlLayout (JSON) tells the caller how to finish the rows: the
// select list as expressions over the aggregate columns and group keys, the
// ORDER BY and the LIMIT. Throws std::runtime_error naming the unsupported
// construct.
HipQueryPlan buildSqlPlan(const std::string& sql, const PlanOptions& opts);

}  // namespace codegen
}  // namespace gdb

#pragma once
// Launch-shape autotuner.
//
// Generalises the reference's threadgroup-size sweep
// (codegen/codegen_main.cpp:837-929: each candidate gets 1 untimed + 5 timed
// runs, the slowest is dropped, the p50 of the rest wins) to three axes:
// block size, ITEMS per thread (a compiled variant of the phase) and one
// optional runtime scalar (e.g. the LDS hash slots in use, as log2). The
// candidate set is the cross product; the phase's kernel time comes from
// events around the kernel alone.

#include "../execution/hip_generic_executor.h"
#include "../execution/runtime_compiler.h"

#include <string>
#include <vector>

namespace gdb {
namespace runtime {

struct TuneResult {
    int block = 0, items = 0;
    int64_t scalar = 0;
    float ms = 0.0f;
    // one row per candidate: block, items, scalar, p50 ms (infinity if the launch failed)
    std::vector<std::array<double, 4>> table;
};

TuneResult autotunePhase(codegen::HipGenericExecutor& ex, codegen::RuntimeCompiler& rc,
                         const codegen::RuntimeCompiler::CompiledQuery& cq, codegen::HipCodegen& cg, int phase,
                         const std::vector<int>& blocks, const std::vector<int>& items, const std::string& scalarName,
                         const std::vector<int64_t>& scalarValues, int trials);

}  // namespace runtime
}  // namespace gdb

#pragma once
// RuntimeCompiler — hipRTC through the C-ABI, with a code-object cache keyed
// by plan hash.
//
// Counterpart of codegen/execution/runtime_compiler.h:9-41. The reference
// compiles each query's MSL into a new MTL::Library and caches pipeline states
// by kernel name for that query only; here gdb_module_load() hashes
// (source, options, architecture, hipRTC version) and reuses code objects
// across queries and processes (memory + on-disk cache), and this class keeps
// one loaded module per plan hash so repeated queries of one shape share it.

#include "gdb/gdb_hip.h"

#include <cstdint>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

namespace gdb {
namespace codegen {

class RuntimeCompiler {
public:
    explicit RuntimeCompiler(gdb_ctx ctx) : ctx_(ctx) {}
    ~RuntimeCompiler();
    RuntimeCompiler(const RuntimeCompiler&) = delete;
    RuntimeCompiler& operator=(const RuntimeCompiler&) = delete;

    struct CompiledQuery {
        gdb_module module = nullptr;
        uint64_t planHash = 0;
        double compileMs = 0.0;   // wall time of this compile() call
        bool cacheHit = false;    // module already loaded, or code object cached
    };

    // Throws std::runtime_error carrying the hipRTC log on failure.
    CompiledQuery compile(const std::string& source);
    gdb_kernel getPipeline(const CompiledQuery& cq, const std::string& kernelName);
    size_t cacheSize() const { return modules_.size(); }
    void clearCache();

private:
    gdb_ctx ctx_;
    std::mutex mu_;
    std::unordered_map<std::string, CompiledQuery> modules_;  // source -> module
};

}  // namespace codegen
}  // namespace gdb

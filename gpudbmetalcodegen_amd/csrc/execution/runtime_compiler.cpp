#include "runtime_compiler.h"

#include <chrono>
#include <stdexcept>

namespace gdb {
namespace codegen {

RuntimeCompiler::~RuntimeCompiler() { clearCache(); }

void RuntimeCompiler::clearCache() {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& kv : modules_) gdb_module_unload(kv.second.module);
    modules_.clear();
}

RuntimeCompiler::CompiledQuery RuntimeCompiler::compile(const std::string& source) {
    auto t0 = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> lk(mu_);
    auto it = modules_.find(source);
    if (it != modules_.end()) {
        CompiledQuery cq = it->second;
        cq.cacheHit = true;
        cq.compileMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return cq;
    }
    uint64_t hitsBefore = 0, hitsAfter = 0;
    gdb_cache_stats(ctx_, &hitsBefore, nullptr, nullptr);
    CompiledQuery cq;
    gdb_status st = gdb_module_load(ctx_, source.data(), source.size(), nullptr, 0, &cq.planHash, &cq.module);
    if (st != GDB_OK) throw std::runtime_error(std::string("RuntimeCompiler: ") + gdb_last_error());
    gdb_cache_stats(ctx_, &hitsAfter, nullptr, nullptr);
    cq.cacheHit = hitsAfter > hitsBefore;
    cq.compileMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    modules_[source] = cq;
    return cq;
}

gdb_kernel RuntimeCompiler::getPipeline(const CompiledQuery& cq, const std::string& kernelName) {
    gdb_kernel k = nullptr;
    if (gdb_kernel_get(cq.module, kernelName.c_str(), &k) != GDB_OK)
        throw std::runtime_error(std::string("RuntimeCompiler: ") + gdb_last_error());
    return k;
}

}  // namespace codegen
}  // namespace gdb

#include "hip_codegen_base.h"

#include "gdb_device_header.h"

#include <sstream>
#include <stdexcept>

namespace gdb {
namespace codegen {

// --- emission ---------------------------------------------------------------

void HipCodegen::addLine(const std::string& line) { cur("addLine").code += indent() + line + "\n"; }

void HipCodegen::addBlock(const std::string& header, const std::function<void()>& body, const std::string& trailing) {
    addLine(header + " {");
    {
        IndentGuard g(*this);
        body();
    }
    addLine("}" + trailing);
}

void HipCodegen::addIf(const std::string& condition, const std::function<void()>& body) {
    addBlock("if (" + condition + ")", body);
}

void HipCodegen::addComment(const std::string& comment) { addLine("// " + comment); }

void HipCodegen::addRawCode(const std::string& code) { cur("addRawCode").code += code; }

void HipCodegen::addHelper(const std::string& code) {
    if (helperSet_.insert(code).second) helpers_.push_back(code);
}

// --- phases -----------------------------------------------------------------

HipCodegen::PhaseInfo& HipCodegen::cur(const char* op) {
    if (current_ < 0) throw std::runtime_error(std::string("HipCodegen::") + op + ": no active phase");
    return phases_[static_cast<size_t>(current_)];
}

void HipCodegen::beginPhase(const std::string& phaseName) {
    if (current_ >= 0) throw std::runtime_error("HipCodegen::beginPhase: phase '" + phaseName + "' nested");
    for (const auto& p : phases_)
        if (p.name == phaseName) throw std::runtime_error("HipCodegen::beginPhase: duplicate phase " + phaseName);
    PhaseInfo p;
    p.name = phaseName;
    phases_.push_back(p);
    current_ = static_cast<int>(phases_.size()) - 1;
    indentLevel_ = 1;
    staged_ = false;
    nextConsumerStageable_ = false;
    rowBindings_.clear();
}

void HipCodegen::endPhase() {
    if (current_ < 0) throw std::runtime_error("HipCodegen::endPhase: no active phase");
    current_ = -1;
}

void HipCodegen::setPhaseScannedTable(const std::string& tableName) {
    auto& p = cur("setPhaseScannedTable");
    p.scannedTable = tableName;
    p.rowsExpr = tableSizeName(tableName);
}
void HipCodegen::setPhaseRowsExpr(const std::string& rowsExpr) { cur("setPhaseRowsExpr").rowsExpr = rowsExpr; }
void HipCodegen::setPhaseThreadgroupSize(int size) {
    if (size <= 0 || size > 1024 || size % 64 != 0)
        throw std::runtime_error("HipCodegen: block size must be a multiple of 64 in [64, 1024]");
    cur("setPhaseThreadgroupSize").threadgroupSize = size;
}
void HipCodegen::setPhaseSingleThread(bool single) { cur("setPhaseSingleThread").isSingleThread = single; }
void HipCodegen::setPhaseMaxThreadgroups(int max) { cur("setPhaseMaxThreadgroups").maxThreadgroups = max; }
void HipCodegen::setPhaseItems(int items) { cur("setPhaseItems").items = items; }
void HipCodegen::setPhaseItemsVariants(const std::vector<int>& items) {
    if (items.empty()) throw std::runtime_error("HipCodegen: empty ITEMS variant list");
    cur("setPhaseItemsVariants").itemsVariants = items;
}
void HipCodegen::setPhaseLaunchBound(int bound) { cur("setPhaseLaunchBound").launchBound = bound; }
void HipCodegen::setPhaseStatusBuffer(const std::string& name) { cur("setPhaseStatusBuffer").statusBuffer = name; }
void HipCodegen::setPhaseCheck(const std::string& checkBuffer, const std::string& message) {
    auto& p = cur("setPhaseCheck");
    if (!p.checkBuffer.empty()) throw std::runtime_error("phase " + p.name + ": one row check per phase");
    p.checkBuffer = checkBuffer;
    p.checkMessage = message;
}

// --- parameters -------------------------------------------------------------

void HipCodegen::pushBinding(const char* op, HipParamBinding b) {
    auto& p = cur(op);
    for (auto& existing : p.bindings) {
        if (existing.name == b.name) {
            // A buffer first registered read-only and later written in the same
            // phase becomes writable; any other re-registration is a no-op.
            if (existing.readOnly && !b.readOnly) {
                existing.readOnly = false;
                existing.declType = b.declType;
            }
            if (existing.sizeExpr.empty()) existing.sizeExpr = b.sizeExpr;
            return;
        }
    }
    if (b.kind == HipParamKind::DeviceBuffer && !b.sizeExpr.empty() && !bufferSizes_.count(b.name))
        bufferSizes_[b.name] = b.sizeExpr;
    b.argIndex = static_cast<int>(p.bindings.size());
    p.bindings.push_back(std::move(b));
}

std::string HipCodegen::addColumnParam(const std::string& table, const std::string& column, const ColumnType& t) {
    HipParamBinding b;
    b.name = tableDataName(table, column);
    b.declType = "const " + t.cxx() + "* __restrict__";
    b.kind = HipParamKind::TableColumn;
    b.tableName = table;
    b.columnName = column;
    b.elementType = t.tag();
    b.elemBytes = t.bytes();
    b.readOnly = true;
    pushBinding("addColumnParam", b);
    return b.name;
}

std::string HipCodegen::addTableSizeParam(const std::string& table) {
    HipParamBinding b;
    b.name = tableSizeName(table);
    b.declType = "const u64";
    b.kind = HipParamKind::TableSize;
    b.tableName = table;
    b.elementType = "u64";
    pushBinding("addTableSizeParam", b);
    return b.name;
}

void HipCodegen::addBufferParam(const std::string& name, const std::string& elemType, size_t elemBytes,
                                const std::string& sizeExpr, InitKind init, const std::vector<uint8_t>& pattern,
                                bool readOnly, bool persistent) {
    HipParamBinding b;
    b.name = name;
    b.declType = (readOnly ? "const " : "") + elemType + "* __restrict__";
    b.kind = HipParamKind::DeviceBuffer;
    b.elementType = elemType;
    b.elemBytes = elemBytes;
    b.sizeExpr = sizeExpr;
    b.init = init;
    b.pattern = pattern;
    b.readOnly = readOnly;
    b.persistent = persistent;
    pushBinding("addBufferParam", b);
}

void HipCodegen::addScalarParam(const std::string& name) {
    HipParamBinding b;
    b.name = name;
    b.declType = "const i64";
    b.kind = HipParamKind::ScalarInt;
    b.elementType = "i64";
    pushBinding("addScalarParam", b);
}

void HipCodegen::addSymbolParam(const std::string& name, const std::string& sizeExpr) {
    HipParamBinding b;
    b.name = name;
    b.declType = "const u64";
    b.kind = HipParamKind::SymbolValue;
    b.elementType = "u64";
    b.sizeExpr = sizeExpr;
    pushBinding("addSymbolParam", b);
}

void HipCodegen::addBitmapWriteParam(const std::string& name, const std::string& bitsExpr) {
    addBufferParam(name, "u32", 4, "(" + bitsExpr + " + 31) / 32 + 1", InitKind::Zero, {}, false, true);
    addSymbolParam("nbits_" + name, bitsExpr);
}

void HipCodegen::addBitmapReadParam(const std::string& name, const std::string& bitsExpr) {
    addBufferParam(name, "u32", 4, "(" + bitsExpr + " + 31) / 32 + 1", InitKind::None, {}, true, true);
    addSymbolParam("nbits_" + name, bitsExpr);
}

std::string HipCodegen::uniqueName(const std::string& prefix) {
    int n = nameCounters_[prefix]++;
    return n == 0 ? prefix : prefix + std::to_string(n);
}

std::vector<HipParamBinding> HipCodegen::getAllBindings() const {
    std::vector<HipParamBinding> all;
    for (const auto& p : phases_)
        for (const auto& b : p.bindings) all.push_back(b);
    return all;
}

// --- printing ---------------------------------------------------------------

std::string HipCodegen::paramList(const PhaseInfo& p, bool withTypes) {
    std::ostringstream os;
    for (size_t i = 0; i < p.bindings.size(); ++i) {
        if (i) os << ", ";
        if (withTypes) os << p.bindings[i].declType << " ";
        os << p.bindings[i].name;
    }
    return os.str();
}

std::string HipCodegen::printHelpers() const {
    std::ostringstream out;
    for (const auto& h : helpers_) out << h << "\n";
    return out.str();
}

void HipCodegen::setPhaseStaged(bool staged) { cur("setPhaseStaged").staged = staged; }

bool HipCodegen::phaseStaged() const { return !phases_.empty() && current_ >= 0 && phases_.back().staged; }

void HipCodegen::rowLoop(const std::function<void()>& body, const std::string& mask) {
    const bool was = staged_;
    addLine("#pragma unroll");
    addBlock("for (int _u = 0; _u < ITEMS; ++_u)", [&]() {
        addBlock("if (" + mask + "[_u])", [&]() {
            const std::vector<std::string> binds = rowBindings_;
            for (const auto& b : binds) addLine(b);
            staged_ = false;
            body();
            staged_ = was;
        });
    });
}

void HipCodegen::declareLds(const std::string& type, const std::string& name, const std::string& countExpr) {
    auto& p = cur("declareLds");
    if (!rangeMode_) {
        addLine("__shared__ __attribute__((aligned(16))) " + type + " " + name + "[" + countExpr + "];");
        return;
    }
    const std::string off = "gdb_lds_align16(" + p.ldsBytesExpr + ")";
    addLine(type + "* const " + name + " = reinterpret_cast<" + type + "*>(_lds + " + off + ");");
    p.ldsBytesExpr = "(" + off + " + (u32)sizeof(" + type + ") * (u32)(" + countExpr + "))";
}

std::string HipCodegen::printPhaseBody(int phase, const std::string& bodyName) const {
    const auto& p = phases_.at(static_cast<size_t>(phase));
    std::ostringstream out;
    if (rangeMode_)
        out << "template <int ITEMS>\n__host__ __device__ constexpr u32 " << bodyName << "_lds_bytes() { return "
            << p.ldsBytesExpr << "; }\n";
    out << "template <int ITEMS>\n__device__ __forceinline__ void " << bodyName << "(" << paramList(p, true);
    if (rangeMode_)
        out << (p.bindings.empty() ? "" : ", ") << "const u64 _rb, const u64 _re, unsigned char* __restrict__ _lds";
    out << ") {\n" << p.code << "}\n";
    return out.str();
}

std::string HipCodegen::print() const {
    std::ostringstream out;
    out << kGdbDeviceHeader;
    out << "\n// ===== helpers =====\n";
    for (const auto& h : helpers_) out << h << "\n";
    for (const auto& p : phases_) {
        out << "\n// ===== phase " << p.name << " (rows: " << (p.rowsExpr.empty() ? "1" : p.rowsExpr) << ") =====\n";
        out << "template <int ITEMS>\n__device__ __forceinline__ void " << p.name << "_body(" << paramList(p, true)
            << ") {\n";
        out << p.code;
        out << "}\n";
        std::vector<int> variants = p.isSingleThread ? std::vector<int>{1} : p.itemsVariants;
        const int lb = p.isSingleThread ? 64 : p.launchBound;
        for (int it : variants) {
            out << "extern \"C\" __global__ void __launch_bounds__(" << lb << ") " << kernelName(p.name, it) << "("
                << paramList(p, true) << ") {\n    " << p.name << "_body<" << it << ">(" << paramList(p, false)
                << ");\n}\n";
        }
    }
    return out.str();
}

}  // namespace codegen
}  // namespace gdb

#pragma once
// GenericResult and HipResultCollector — device buffers -> exact rows.
//
// Counterpart of codegen/execution/metal_result_collector.h:19-60. Results
// are columnar and exact: 64-bit and 128-bit integers (fixed-point values
// keep their digits; no scale-down to double), strings for group labels.
// toCanonical() renders CSV with a header, like GenericResult::toCanonical
// (metal_result_collector.cpp:85-122) but with exact integers.

#include "../core/hip_param_binding.h"

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace gdb {
namespace codegen {

struct ResultColumn {
    enum Type { I64, I128, STR };
    std::string name;
    Type type = I64;
    std::vector<int64_t> i64;
    std::vector<__int128> i128;
    std::vector<std::string> str;
    int scale = 0;      // integer columns: fixed-point digits, shown as an exact decimal
    bool date = false;  // integer columns: YYYYMMDD shown as YYYY-MM-DD
    size_t size() const { return type == I64 ? i64.size() : (type == I128 ? i128.size() : str.size()); }
};

struct GenericResult {
    std::vector<ResultColumn> columns;
    size_t rows() const { return columns.empty() ? 0 : columns[0].size(); }
    const ResultColumn& column(const std::string& name) const;
    std::string toCanonical() const;
};

std::string int128ToString(__int128 v);
// Exact decimal of v / 10^scale ("-0.05" for v = -5, scale = 2).
std::string fixedToString(__int128 v, int scale);

// Reads a runner buffer: (name, offset bytes, dst, bytes). Synchronous.
using BufferReader = std::function<void(const std::string&, size_t, void*, size_t)>;
using SymbolReader = std::function<uint64_t(const std::string&)>;

class HipResultCollector {
public:
    static GenericResult collect(const HipResultSchema& schema, const BufferReader& read, const SymbolReader& sym);

private:
    static GenericResult collectKeyedAgg(const HipResultSchema& schema, const BufferReader& read);
    static GenericResult collectHashAgg(const HipResultSchema& schema, const BufferReader& read,
                                        const SymbolReader& sym);
    static GenericResult collectTopK(const HipResultSchema& schema, const BufferReader& read);
    static GenericResult collectCounts(const HipResultSchema& schema, const BufferReader& read,
                                       const SymbolReader& sym);
    static GenericResult collectRows(const HipResultSchema& schema, const BufferReader& read,
                                     const SymbolReader& sym);
};

}  // namespace codegen
}  // namespace gdb

#pragma once
// HipParamBinding, HipSizeResolver, HipResultSchema — the metadata that ties
// generated kernels to the runner.
//
// Counterpart of codegen/core/metal_param_binding.h:29-144. A binding
// describes one kernel parameter: its C++ declaration, where its value comes
// from (table column, row count, runner-owned device buffer, scalar), its
// symbolic size and its initialisation. Instead of Metal's [[buffer(N)]]
// slots, bindings become positional kernel arguments packed into one kernarg
// buffer with natural alignment.

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace gdb {
namespace codegen {

// Multiplicative hash constant shared by every hash table the engine builds
// and probes (same constant as codegen/core/metal_param_binding.h:26). The
// engine takes the HIGH bits of key*C (Fibonacci hashing) rather than the low
// bits, so power-of-two strided keys do not collide into a few slots.
inline constexpr uint32_t kKnuthHashMul = 2654435769u;

// Sentinel for an empty 32-bit hash-table key (query_preprocessing.cpp:362).
inline constexpr uint32_t kEmptyKey32 = 0xFFFFFFFFu;

enum class HipParamKind {
    TableColumn,   // const T* __restrict__ — one column of a registered table
    TableSize,     // unsigned long long     — row count of a registered table
    DeviceBuffer,  // T* __restrict__        — runner-owned scratch / output buffer
    ScalarInt,     // long long              — value registered by name
    SymbolValue,   // unsigned long long     — value of a size symbol (e.g. slot count)
};

enum class InitKind { None, Zero, Pattern };

inline std::string tableDataName(const std::string& table, const std::string& col) {
    return "p_" + col;
}
inline std::string tableSizeName(const std::string& table) { return "n_" + table; }

struct HipParamBinding {
    std::string name;          // parameter name inside the kernel
    std::string declType;      // e.g. "const int* __restrict__", "unsigned long long"
    HipParamKind kind = HipParamKind::ScalarInt;
    std::string tableName;     // TableColumn / TableSize
    std::string columnName;    // TableColumn: catalog column name
    std::string elementType;   // "i32","i64","u8","u32","u64","char","slot"
    size_t elemBytes = 0;      // bytes per element (DeviceBuffer sizing)
    std::string sizeExpr;      // DeviceBuffer: element count expression
    InitKind init = InitKind::None;
    std::vector<uint8_t> pattern;  // InitKind::Pattern: repeating bytes (divides 256)
    bool readOnly = false;
    bool persistent = false;   // DeviceBuffer shared across phases (bitmap, table)
    int argIndex = -1;
};

// Symbolic size resolution: literals, symbols, + - * / and parentheses.
// Counterpart of MetalSizeResolver (codegen/core/metal_param_binding.cpp:32-101);
// an unknown symbol raises, as there.
class HipSizeResolver {
public:
    void registerSymbol(const std::string& name, uint64_t value) { symbols_[name] = value; }
    bool hasSymbol(const std::string& name) const { return symbols_.count(name) > 0; }
    uint64_t getSymbol(const std::string& name) const {
        auto it = symbols_.find(name);
        if (it == symbols_.end()) throw std::runtime_error("HipSizeResolver: unknown symbol '" + name + "'");
        return it->second;
    }
    uint64_t resolve(const std::string& expr) const;
    const std::unordered_map<std::string, uint64_t>& symbols() const { return symbols_; }

private:
    std::unordered_map<std::string, uint64_t> symbols_;
};

// What a plan produces and where the runner finds it.
// Counterpart of MetalResultSchema (codegen/core/metal_param_binding.h:103-144),
// with exact integer kinds instead of lo/hi 32-bit pairs and double scale-down.
enum class AggKind { SumI64, SumI128, Count, MinI64, MaxI64 };

struct HipResultSchema {
    enum Kind { NONE, KEYED_AGG, HASH_AGG, TOPK, HISTOGRAM, COUNTS, ROWS };
    Kind kind = NONE;

    struct AggSlot {
        std::string name;
        AggKind agg = AggKind::SumI64;
        int wordOffset = 0;    // u64 word offset inside the bucket / slot
        int scaleDigits = 0;   // fixed-point digits (display only; values stay integers)
    };

    // KEYED_AGG: numBuckets buckets of wordsPerBucket u64 words in bufferName.
    struct KeyedAggInfo {
        std::string bufferName;
        int numBuckets = 0;
        int wordsPerBucket = 0;
        std::vector<AggSlot> slots;
        std::vector<std::string> bucketLabels;  // optional display labels per bucket
        std::string keyColumn = "bucket";
        std::vector<std::pair<std::string, std::pair<std::string, std::string>>> averages;
        // (display name, (numerator slot, denominator slot)); exact derived averages
    };
    KeyedAggInfo keyed;

    // HASH_AGG: open-addressing table of slotWords-u64 slots; word 0 holds the
    // 32-bit key in its low half (empty = 0xFFFFFFFF) and the 32-bit count in
    // its high half when countInKeyWord is set.
    struct HashAggInfo {
        std::string tableName;
        std::string slotsSymbol;
        int slotWords = 4;
        std::string keyName = "key";
        std::vector<AggSlot> slots;
    };
    HashAggInfo hash;

    // TOPK: K entries of two u64 words (rank hi, rank lo) in bufferName, best
    // first; the number of valid entries is word 3 of ctlBuffer. Output
    // columns are bit fields of the rank words.
    struct TopKColumn {
        std::string name;
        int word = 0;       // 0 = rank hi, 1 = rank lo
        int shift = 0;
        int bits = 64;
        long long complementOf = 0;  // value = complementOf - field when != 0
        int scale = 0;               // fixed-point digits (display only)
        bool date = false;           // YYYYMMDD integer shown as YYYY-MM-DD
    };
    struct TopKInfo {
        std::string bufferName;
        std::string ctlBuffer;
        int k = 10;
        std::vector<TopKColumn> columns;
    };
    TopKInfo topk;

    // HISTOGRAM / COUNTS: u32 counters in bufferName of size sizeExpr.
    struct CountsInfo {
        std::string bufferName;
        std::string sizeExpr;
        std::string keyName = "key";
        std::string valueName = "count";
        bool sortByCountDesc = false;  // Q13-style (custdist desc, c_count desc)
    };
    CountsInfo counts;

    // ROWS: records of `words` u64 fields in bufferName, their number in word
    // 0 of ctlBuffer; ordered by (field index, descending) keys, then cut to
    // limit rows (0 = all).
    struct RowsField {
        std::string name;
        int scale = 0;
        bool date = false;
    };
    struct RowsInfo {
        std::string bufferName;
        std::string ctlBuffer;
        std::string capExpr;
        std::vector<RowsField> fields;
        std::vector<std::pair<int, bool>> orderBy;
        int limit = 0;
    };
    RowsInfo rows;
};

}  // namespace codegen
}  // namespace gdb

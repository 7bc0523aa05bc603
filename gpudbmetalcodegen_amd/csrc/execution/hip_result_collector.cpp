#include "hip_result_collector.h"

#include <algorithm>
#include <cstdio>
#include <numeric>
#include <sstream>
#include <stdexcept>

namespace gdb {
namespace codegen {

std::string int128ToString(__int128 v) {
    if (v == 0) return "0";
    const bool neg = v < 0;
    unsigned __int128 u = neg ? (unsigned __int128)(-(v + 1)) + 1u : (unsigned __int128)v;
    std::string s;
    while (u) {
        s.push_back(static_cast<char>('0' + static_cast<int>(u % 10u)));
        u /= 10u;
    }
    if (neg) s.push_back('-');
    std::reverse(s.begin(), s.end());
    return s;
}

std::string fixedToString(__int128 v, int scale) {
    if (scale <= 0) return int128ToString(v);
    const bool neg = v < 0;
    std::string d = int128ToString(neg ? -v : v);  // |v| of any sum here is far below 2^127
    if (d.size() <= static_cast<size_t>(scale)) d.insert(0, static_cast<size_t>(scale) + 1 - d.size(), '0');
    d.insert(d.size() - static_cast<size_t>(scale), 1, '.');
    return neg ? "-" + d : d;
}

namespace {

std::string dateString(int64_t ymd) {
    char buf[32];
    std::snprintf(buf, sizeof(buf), "%04lld-%02lld-%02lld", static_cast<long long>(ymd / 10000),
                  static_cast<long long>(ymd / 100 % 100), static_cast<long long>(ymd % 100));
    return buf;
}

// Quoted when it holds a comma, a quote or a line break; quotes doubled
// (metal_result_collector.cpp:102-115).
std::string csvField(const std::string& s) {
    if (s.find_first_of(",\"\n\r") == std::string::npos) return s;
    std::string q = "\"";
    for (char ch : s) {
        if (ch == '"') q += '"';
        q += ch;
    }
    return q + "\"";
}

}  // namespace

const ResultColumn& GenericResult::column(const std::string& name) const {
    for (const auto& c : columns)
        if (c.name == name) return c;
    throw std::runtime_error("GenericResult: no column " + name);
}

std::string GenericResult::toCanonical() const {
    std::ostringstream os;
    for (size_t c = 0; c < columns.size(); ++c) os << (c ? "," : "") << columns[c].name;
    os << "\n";
    const size_t n = rows();
    for (size_t r = 0; r < n; ++r) {
        for (size_t c = 0; c < columns.size(); ++c) {
            if (c) os << ",";
            const auto& col = columns[c];
            if (col.type == ResultColumn::STR)
                os << csvField(col.str[r]);
            else if (col.date)
                os << dateString(col.type == ResultColumn::I64 ? col.i64[r] : static_cast<int64_t>(col.i128[r]));
            else
                os << fixedToString(col.type == ResultColumn::I64 ? col.i64[r] : col.i128[r], col.scale);
        }
        os << "\n";
    }
    return os.str();
}

namespace {

ResultColumn mk(const std::string& name, ResultColumn::Type t) {
    ResultColumn c;
    c.name = name;
    c.type = t;
    return c;
}

std::vector<std::string> split(const std::string& s, char d) {
    std::vector<std::string> out;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, d)) out.push_back(item);
    return out;
}

__int128 slotValue(const HipResultSchema::AggSlot& slot, const uint64_t* words) {
    switch (slot.agg) {
        case AggKind::SumI128: {
            const uint64_t lo = words[slot.wordOffset];
            const int64_t hi = static_cast<int64_t>(words[slot.wordOffset + 1]);
            return (static_cast<__int128>(hi) << 32) + static_cast<__int128>(lo);
        }
        default:
            return static_cast<int64_t>(words[slot.wordOffset]);
    }
}

}  // namespace

GenericResult HipResultCollector::collect(const HipResultSchema& schema, const BufferReader& read,
                                          const SymbolReader& sym) {
    switch (schema.kind) {
        case HipResultSchema::KEYED_AGG: return collectKeyedAgg(schema, read);
        case HipResultSchema::HASH_AGG: return collectHashAgg(schema, read, sym);
        case HipResultSchema::TOPK: return collectTopK(schema, read);
        case HipResultSchema::HISTOGRAM:
        case HipResultSchema::COUNTS: return collectCounts(schema, read, sym);
        case HipResultSchema::ROWS: return collectRows(schema, read, sym);
        default: return GenericResult{};
    }
}

// Buckets with no qualifying row are skipped, as in
// metal_result_collector.cpp:228-235 (there: all slots zero; here: the count
// slot, when present, is zero).
GenericResult HipResultCollector::collectKeyedAgg(const HipResultSchema& schema, const BufferReader& read) {
    const auto& k = schema.keyed;
    std::vector<uint64_t> words(static_cast<size_t>(k.numBuckets) * k.wordsPerBucket);
    read(k.bufferName, 0, words.data(), words.size() * 8);

    GenericResult res;
    std::vector<std::string> keyNames = split(k.keyColumn, ',');
    const bool labelled = !k.bucketLabels.empty();
    if (labelled) {
        for (const auto& kn : keyNames) res.columns.push_back(mk(kn, ResultColumn::STR));
    } else {
        res.columns.push_back(mk(k.keyColumn, ResultColumn::I64));
    }
    for (const auto& s : k.slots) {
        res.columns.push_back(mk(s.name, s.agg == AggKind::SumI128 ? ResultColumn::I128 : ResultColumn::I64));
        res.columns.back().scale = s.scaleDigits;
    }

    int countSlot = -1;
    for (size_t i = 0; i < k.slots.size(); ++i)
        if (k.slots[i].agg == AggKind::Count) countSlot = static_cast<int>(i);

    auto slotIndex = [&](const std::string& name) -> int {
        for (size_t i = 0; i < k.slots.size(); ++i)
            if (k.slots[i].name == name) return static_cast<int>(i);
        throw std::runtime_error("collector: unknown slot " + name);
    };
    // An average carries 6 more digits than numerator / denominator.
    for (const auto& a : k.averages) {
        res.columns.push_back(mk(a.first, ResultColumn::I128));
        res.columns.back().scale = std::max(0, 6 + k.slots[static_cast<size_t>(slotIndex(a.second.first))].scaleDigits -
                                                   k.slots[static_cast<size_t>(slotIndex(a.second.second))].scaleDigits);
    }

    for (int b = 0; b < k.numBuckets; ++b) {
        const uint64_t* w = words.data() + static_cast<size_t>(b) * k.wordsPerBucket;
        bool has;
        if (countSlot >= 0) {
            has = w[k.slots[countSlot].wordOffset] != 0;
        } else {
            has = false;
            for (int i = 0; i < k.wordsPerBucket; ++i) has |= (w[i] != 0);
        }
        if (!has) continue;
        size_t c = 0;
        if (labelled) {
            std::vector<std::string> parts = split(k.bucketLabels[static_cast<size_t>(b)], '|');
            for (size_t j = 0; j < keyNames.size(); ++j) res.columns[c++].str.push_back(j < parts.size() ? parts[j] : "");
        } else {
            res.columns[c++].i64.push_back(b);
        }
        for (const auto& s : k.slots) {
            __int128 v = slotValue(s, w);
            auto& col = res.columns[c++];
            if (col.type == ResultColumn::I128)
                col.i128.push_back(v);
            else
                col.i64.push_back(static_cast<int64_t>(v));
        }
        // Exact derived averages: trunc(numerator * 10^6 / denominator).
        for (const auto& a : k.averages) {
            const __int128 num = slotValue(k.slots[slotIndex(a.second.first)], w);
            const __int128 den = slotValue(k.slots[slotIndex(a.second.second)], w);
            res.columns[c++].i128.push_back(den == 0 ? 0 : (num * 1000000) / den);
        }
    }
    return res;
}

GenericResult HipResultCollector::collectHashAgg(const HipResultSchema& schema, const BufferReader& read,
                                                 const SymbolReader& sym) {
    const auto& h = schema.hash;
    const uint64_t slots = sym(h.slotsSymbol);
    std::vector<uint64_t> tab(slots * static_cast<uint64_t>(h.slotWords));
    read(h.tableName, 0, tab.data(), tab.size() * 8);
    std::vector<uint64_t> occ;
    for (uint64_t s = 0; s < slots; ++s)
        if (static_cast<uint32_t>(tab[s * h.slotWords]) != kEmptyKey32) occ.push_back(s);
    std::sort(occ.begin(), occ.end(), [&](uint64_t a, uint64_t b) {
        return static_cast<uint32_t>(tab[a * h.slotWords]) < static_cast<uint32_t>(tab[b * h.slotWords]);
    });
    GenericResult res;
    res.columns = {mk(h.keyName, ResultColumn::I64), mk("count", ResultColumn::I64), mk("sum", ResultColumn::I64),
                   mk("min", ResultColumn::I64), mk("max", ResultColumn::I64)};
    for (auto& c : res.columns) c.i64.reserve(occ.size());
    for (uint64_t s : occ) {
        const uint64_t* w = tab.data() + s * h.slotWords;
        res.columns[0].i64.push_back(static_cast<uint32_t>(w[0]));
        res.columns[1].i64.push_back(static_cast<int64_t>(w[0] >> 32));
        res.columns[2].i64.push_back(static_cast<int64_t>(w[1]));
        res.columns[3].i64.push_back(static_cast<int64_t>(w[2]));
        res.columns[4].i64.push_back(static_cast<int64_t>(w[3]));
    }
    return res;
}

GenericResult HipResultCollector::collectTopK(const HipResultSchema& schema, const BufferReader& read) {
    const auto& t = schema.topk;
    std::vector<uint64_t> ctl(4), out(2 * static_cast<size_t>(t.k));
    read(t.ctlBuffer, 0, ctl.data(), 32);
    read(t.bufferName, 0, out.data(), out.size() * 8);
    GenericResult res;
    for (const auto& c : t.columns) {
        res.columns.push_back(mk(c.name, ResultColumn::I64));
        res.columns.back().scale = c.scale;
        res.columns.back().date = c.date;
    }
    const uint64_t valid = std::min<uint64_t>(ctl[3], static_cast<uint64_t>(t.k));
    for (uint64_t j = 0; j < valid; ++j) {
        for (size_t ci = 0; ci < t.columns.size(); ++ci) {
            const auto& c = t.columns[ci];
            uint64_t field = out[2 * j + static_cast<uint64_t>(c.word)] >> c.shift;
            if (c.bits < 64) field &= (1ull << c.bits) - 1ull;
            int64_t v = c.complementOf ? c.complementOf - static_cast<int64_t>(field) : static_cast<int64_t>(field);
            res.columns[ci].i64.push_back(v);
        }
    }
    return res;
}

GenericResult HipResultCollector::collectCounts(const HipResultSchema& schema, const BufferReader& read,
                                                const SymbolReader& sym) {
    const auto& c = schema.counts;
    HipSizeResolver r;
    (void)r;
    const uint64_t n = sym(c.sizeExpr);
    std::vector<uint32_t> v(n);
    read(c.bufferName, 0, v.data(), n * 4);
    std::vector<std::pair<int64_t, int64_t>> rows;  // (key, count)
    for (uint64_t i = 0; i < n; ++i)
        if (v[i]) rows.push_back({static_cast<int64_t>(i), static_cast<int64_t>(v[i])});
    if (c.sortByCountDesc) {
        // Q13 ordering: custdist desc, then c_count desc (codegen_main.cpp:1250-1277).
        std::sort(rows.begin(), rows.end(), [](const auto& a, const auto& b) {
            if (a.second != b.second) return a.second > b.second;
            return a.first > b.first;
        });
    }
    GenericResult res;
    res.columns = {mk(c.keyName, ResultColumn::I64), mk(c.valueName, ResultColumn::I64)};
    for (const auto& rw : rows) {
        res.columns[0].i64.push_back(rw.first);
        res.columns[1].i64.push_back(rw.second);
    }
    return res;
}

// Appended records, ordered by the schema's keys with every field ascending
// as the final tie-break (so the order is total), then cut to the limit.
GenericResult HipResultCollector::collectRows(const HipResultSchema& schema, const BufferReader& read,
                                              const SymbolReader& sym) {
    const auto& r = schema.rows;
    const size_t W = r.fields.size();
    std::vector<uint64_t> ctl(4);
    read(r.ctlBuffer, 0, ctl.data(), 32);
    const uint64_t n = std::min<uint64_t>(ctl[2], sym(r.capExpr));
    std::vector<int64_t> rec(static_cast<size_t>(n) * W);
    if (n) read(r.bufferName, 0, rec.data(), rec.size() * 8);
    std::vector<size_t> idx(static_cast<size_t>(n));
    std::iota(idx.begin(), idx.end(), size_t{0});
    std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
        const int64_t* x = rec.data() + a * W;
        const int64_t* y = rec.data() + b * W;
        for (const auto& o : r.orderBy) {
            const int64_t u = x[o.first], v = y[o.first];
            if (u != v) return o.second ? u > v : u < v;
        }
        for (size_t f = 0; f < W; ++f)
            if (x[f] != y[f]) return x[f] < y[f];
        return false;
    });
    if (r.limit > 0 && idx.size() > static_cast<size_t>(r.limit)) idx.resize(static_cast<size_t>(r.limit));
    GenericResult res;
    for (const auto& f : r.fields) {
        res.columns.push_back(mk(f.name, ResultColumn::I64));
        res.columns.back().scale = f.scale;
        res.columns.back().date = f.date;
    }
    for (size_t i : idx)
        for (size_t f = 0; f < W; ++f) res.columns[f].i64.push_back(rec[i * W + f]);
    return res;
}

}  // namespace codegen
}  // namespace gdb

#include "colbin.h"

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <fcntl.h>
#include <stdexcept>
#include <sys/mman.h>
#include <sys/stat.h>
#include <thread>
#include <unistd.h>

namespace gdb {
namespace infra {

namespace {

constexpr char kMagic[8] = {'T', 'P', 'C', 'H', 'C', 'B', '0', '1'};
constexpr uint32_t kVersion = 2;
constexpr uint64_t kAlign = 16384;

#pragma pack(push, 1)
struct FileHeader {
    char magic[8];
    uint32_t version;
    uint32_t nCols;
    uint64_t nRows;
    uint64_t sourceSize;
    int64_t sourceMtimeNs;
    uint64_t pad;
};
struct ColDesc {
    int32_t columnIndex;
    uint8_t dtype;
    uint8_t pad0;
    uint16_t fixedWidth;
    uint64_t offset;
    uint64_t sizeBytes;
    uint64_t pad1;
};
#pragma pack(pop)
static_assert(sizeof(FileHeader) == 48, "colbin header is 48 bytes");
static_assert(sizeof(ColDesc) == 32, "colbin column descriptor is 32 bytes");

uint64_t alignUp(uint64_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

bool statFile(const std::string& p, uint64_t& size, int64_t& mtimeNs) {
    struct stat st{};
    if (::stat(p.c_str(), &st) != 0) return false;
    size = static_cast<uint64_t>(st.st_size);
    mtimeNs = static_cast<int64_t>(st.st_mtim.tv_sec) * 1000000000LL + static_cast<int64_t>(st.st_mtim.tv_nsec);
    return true;
}

TpchType decodeType(uint8_t b) {
    switch (b) {
        case 0: return TpchType::INT;
        case 1: return TpchType::FLOAT;
        case 2: return TpchType::DATE;
        case 3: return TpchType::CHAR1;
        default: return TpchType::CHAR_FIXED;  // as src/infra.h:559-564
    }
}

TpchTable mkTable(const std::string& name, std::vector<TpchColumn> cols) { return TpchTable{name, std::move(cols)}; }

constexpr TpchType I = TpchType::INT, F = TpchType::FLOAT, D = TpchType::DATE, C1 = TpchType::CHAR1,
                   CF = TpchType::CHAR_FIXED;

// --- .tbl field parsers (the rules of src/infra.h:229-258) -----------------

int32_t parseInt(const char* p, const char* e) {
    while (p < e && (*p == ' ' || *p == '\t')) ++p;
    bool neg = false;
    if (p < e && (*p == '-' || *p == '+')) neg = (*p++ == '-');
    int64_t v = 0;
    while (p < e && *p >= '0' && *p <= '9') v = v * 10 + (*p++ - '0');
    return static_cast<int32_t>(neg ? -v : v);
}

float parseFloat(const char* p, const char* e) {
    char buf[64];
    const size_t n = std::min<size_t>(static_cast<size_t>(e - p), sizeof(buf) - 1);
    std::memcpy(buf, p, n);
    buf[n] = '\0';
    return std::strtof(buf, nullptr);
}

int32_t parseDate(const char* p, const char* e) {
    int y = 0, m = 0, d = 0;
    while (p < e && *p >= '0' && *p <= '9') y = y * 10 + (*p++ - '0');
    if (p < e) ++p;
    while (p < e && *p >= '0' && *p <= '9') m = m * 10 + (*p++ - '0');
    if (p < e) ++p;
    while (p < e && *p >= '0' && *p <= '9') d = d * 10 + (*p++ - '0');
    return y * 10000 + m * 100 + d;
}

struct Mapped {
    void* p = nullptr;
    size_t n = 0;
    explicit Mapped(const std::string& path) {
        const int fd = ::open(path.c_str(), O_RDONLY);
        if (fd < 0) throw std::runtime_error("cannot open " + path + ": " + std::strerror(errno));
        struct stat st{};
        if (::fstat(fd, &st) != 0) {
            ::close(fd);
            throw std::runtime_error("cannot stat " + path);
        }
        n = static_cast<size_t>(st.st_size);
        if (n > 0) {
            p = ::mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
            if (p == MAP_FAILED) {
                p = nullptr;
                ::close(fd);
                throw std::runtime_error("cannot map " + path);
            }
        }
        ::close(fd);
    }
    ~Mapped() {
        if (p) ::munmap(p, n);
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// Catalog (codegen/planning/tpch_schema.h:66-162: same names, indices, types
// and widths, since .colbin descriptors are keyed on them)
// ---------------------------------------------------------------------------

const std::vector<TpchTable>& tpchTables() {
    static const std::vector<TpchTable> tables = {
        mkTable("lineitem", {{"l_orderkey", 0, I},       {"l_partkey", 1, I},         {"l_suppkey", 2, I},
                             {"l_linenumber", 3, I},     {"l_quantity", 4, F},        {"l_extendedprice", 5, F},
                             {"l_discount", 6, F},       {"l_tax", 7, F},             {"l_returnflag", 8, C1},
                             {"l_linestatus", 9, C1},    {"l_shipdate", 10, D},       {"l_commitdate", 11, D},
                             {"l_receiptdate", 12, D},   {"l_shipinstruct", 13, CF, 25},
                             {"l_shipmode", 14, CF, 2},  {"l_comment", 15, CF, 44}}),
        mkTable("orders", {{"o_orderkey", 0, I},      {"o_custkey", 1, I},        {"o_orderstatus", 2, C1},
                           {"o_totalprice", 3, F},    {"o_orderdate", 4, D},      {"o_orderpriority", 5, C1},
                           {"o_clerk", 6, CF, 15},    {"o_shippriority", 7, I},   {"o_comment", 8, CF, 79}}),
        mkTable("customer", {{"c_custkey", 0, I},      {"c_name", 1, CF, 25},   {"c_address", 2, CF, 40},
                             {"c_nationkey", 3, I},    {"c_phone", 4, CF, 15},  {"c_acctbal", 5, F},
                             {"c_mktsegment", 6, C1},  {"c_comment", 7, CF, 117}}),
        mkTable("supplier", {{"s_suppkey", 0, I}, {"s_name", 1, CF, 25}, {"s_address", 2, CF, 40},
                             {"s_nationkey", 3, I}, {"s_phone", 4, CF, 15}, {"s_acctbal", 5, F},
                             {"s_comment", 6, CF, 101}}),
        mkTable("part", {{"p_partkey", 0, I},        {"p_name", 1, CF, 55},     {"p_mfgr", 2, CF, 25},
                         {"p_brand", 3, CF, 10},     {"p_type", 4, CF, 25},     {"p_size", 5, I},
                         {"p_container", 6, CF, 10}, {"p_retailprice", 7, F},   {"p_comment", 8, CF, 23}}),
        mkTable("partsupp", {{"ps_partkey", 0, I}, {"ps_suppkey", 1, I}, {"ps_availqty", 2, I},
                             {"ps_supplycost", 3, F}, {"ps_comment", 4, CF, 199}}),
        mkTable("nation", {{"n_nationkey", 0, I}, {"n_name", 1, CF, 25}, {"n_regionkey", 2, I},
                           {"n_comment", 3, CF, 152}}),
        mkTable("region", {{"r_regionkey", 0, I}, {"r_name", 1, CF, 25}, {"r_comment", 2, CF, 152}}),
    };
    return tables;
}

const TpchColumn& TpchTable::column(const std::string& n) const {
    for (const auto& c : columns)
        if (c.name == n) return c;
    throw std::runtime_error("unknown column " + name + "." + n);
}

const TpchTable& tpchTable(const std::string& name) {
    for (const auto& t : tpchTables())
        if (t.name == name) return t;
    throw std::runtime_error("unknown TPC-H table " + name);
}

std::vector<ColSpec> specsOf(const TpchTable& t) {
    std::vector<ColSpec> v;
    for (const auto& c : t.columns) v.push_back({c.index, c.type, c.width});
    return v;
}

size_t elemBytes(TpchType t, int width) {
    switch (t) {
        case TpchType::INT:
        case TpchType::DATE:
        case TpchType::FLOAT: return 4;
        case TpchType::CHAR1: return 1;
        case TpchType::CHAR_FIXED: return static_cast<size_t>(width);
    }
    return 4;
}

// ---------------------------------------------------------------------------
// .tbl parser
// ---------------------------------------------------------------------------

ParsedTable parseTbl(const std::string& tblPath, const std::vector<ColSpec>& specs, int threads) {
    if (specs.empty()) throw std::runtime_error("parseTbl: no columns requested");
    int maxIdx = 0;
    for (const auto& s : specs) {
        if (s.index < 0) throw std::runtime_error("parseTbl: negative column index");
        if (s.type == TpchType::CHAR_FIXED && s.width <= 0) throw std::runtime_error("parseTbl: CHAR_FIXED needs a width");
        maxIdx = std::max(maxIdx, s.index);
    }
    std::vector<int> specOf(static_cast<size_t>(maxIdx) + 1, -1);
    for (size_t i = 0; i < specs.size(); ++i) specOf[static_cast<size_t>(specs[i].index)] = static_cast<int>(i);

    Mapped m(tblPath);
    const char* base = static_cast<const char*>(m.p);
    const size_t size = m.n;
    const int T = std::max(1, std::min(threads, 64));
    // Range t covers the lines that start in [cut[t], cut[t+1]).
    std::vector<size_t> cut(static_cast<size_t>(T) + 1, size);
    cut[0] = 0;
    for (int t = 1; t < T; ++t) {
        size_t pos = size * static_cast<size_t>(t) / static_cast<size_t>(T);
        if (pos < cut[static_cast<size_t>(t) - 1]) pos = cut[static_cast<size_t>(t) - 1];
        while (pos < size && pos > 0 && base[pos - 1] != '\n') ++pos;
        cut[static_cast<size_t>(t)] = pos;
    }
    auto forLines = [&](size_t lo, size_t hi, auto&& fn) {
        size_t p = lo;
        while (p < hi) {
            const char* nl = static_cast<const char*>(std::memchr(base + p, '\n', size - p));
            const size_t e = nl ? static_cast<size_t>(nl - base) : size;
            fn(base + p, base + e);
            p = e + 1;
        }
    };
    auto isData = [](const char* b, const char* e) { return std::memchr(b, '|', static_cast<size_t>(e - b)) != nullptr; };

    std::vector<uint64_t> lines(static_cast<size_t>(T), 0);
    {
        std::vector<std::thread> ts;
        for (int t = 0; t < T; ++t)
            ts.emplace_back([&, t]() {
                uint64_t c = 0;
                forLines(cut[static_cast<size_t>(t)], cut[static_cast<size_t>(t) + 1],
                         [&](const char* b, const char* e) { c += isData(b, e) ? 1 : 0; });
                lines[static_cast<size_t>(t)] = c;
            });
        for (auto& th : ts) th.join();
    }
    std::vector<uint64_t> first(static_cast<size_t>(T) + 1, 0);
    for (int t = 0; t < T; ++t) first[static_cast<size_t>(t) + 1] = first[static_cast<size_t>(t)] + lines[static_cast<size_t>(t)];

    ParsedTable out;
    out.rows = first[static_cast<size_t>(T)];
    out.specs = specs;
    out.data.resize(specs.size());
    for (size_t i = 0; i < specs.size(); ++i) out.data[i].assign(out.rows * elemBytes(specs[i].type, specs[i].width), 0);

    std::atomic<bool> bad{false};
    std::string badMsg;
    std::vector<std::thread> ts;
    for (int t = 0; t < T; ++t)
        ts.emplace_back([&, t]() {
            uint64_t row = first[static_cast<size_t>(t)];
            forLines(cut[static_cast<size_t>(t)], cut[static_cast<size_t>(t) + 1], [&](const char* b, const char* e) {
                if (bad.load(std::memory_order_relaxed) || !isData(b, e)) return;
                const char* tok = b;
                int col = 0, seen = 0;
                while (col <= maxIdx) {
                    const char* bar = static_cast<const char*>(std::memchr(tok, '|', static_cast<size_t>(e - tok)));
                    if (!bar) break;
                    const int si = specOf[static_cast<size_t>(col)];
                    if (si >= 0) {
                        const ColSpec& s = specs[static_cast<size_t>(si)];
                        uint8_t* dst = out.data[static_cast<size_t>(si)].data();
                        switch (s.type) {
                            case TpchType::INT: {
                                const int32_t v = parseInt(tok, bar);
                                std::memcpy(dst + row * 4, &v, 4);
                                break;
                            }
                            case TpchType::DATE: {
                                const int32_t v = parseDate(tok, bar);
                                std::memcpy(dst + row * 4, &v, 4);
                                break;
                            }
                            case TpchType::FLOAT: {
                                const float v = parseFloat(tok, bar);
                                std::memcpy(dst + row * 4, &v, 4);
                                break;
                            }
                            case TpchType::CHAR1: dst[row] = bar > tok ? static_cast<uint8_t>(*tok) : 0; break;
                            case TpchType::CHAR_FIXED: {
                                const size_t w = static_cast<size_t>(s.width);
                                const size_t n = std::min(w, static_cast<size_t>(bar - tok));
                                std::memcpy(dst + row * w, tok, n);  // the rest stays NUL
                                break;
                            }
                        }
                        ++seen;
                    }
                    tok = bar + 1;
                    ++col;
                }
                if (seen != static_cast<int>(specs.size())) {
                    if (!bad.exchange(true)) badMsg = "line " + std::to_string(row + 1) + " of " + tblPath +
                                                      " has fewer fields than column " + std::to_string(maxIdx) + " needs";
                    return;
                }
                ++row;
            });
        });
    for (auto& th : ts) th.join();
    if (bad) throw std::runtime_error("parseTbl: " + badMsg);
    return out;
}

// ---------------------------------------------------------------------------
// .colbin v2
// ---------------------------------------------------------------------------

std::string colbinPath(const std::string& tblPath) {
    const auto dot = tblPath.rfind('.');
    const auto slash = tblPath.rfind('/');
    const bool hasExt = dot != std::string::npos && (slash == std::string::npos || dot > slash);
    return (hasExt ? tblPath.substr(0, dot) : tblPath) + ".colbin";
}

void writeColbin(const std::string& tblPath, const ParsedTable& t) {
    uint64_t srcSize = 0;
    int64_t srcMtime = 0;
    if (!statFile(tblPath, srcSize, srcMtime)) throw std::runtime_error("writeColbin: cannot stat " + tblPath);
    if (t.specs.empty() || t.specs.size() != t.data.size()) throw std::runtime_error("writeColbin: no columns");
    std::vector<ColDesc> descs(t.specs.size());
    uint64_t cursor = alignUp(sizeof(FileHeader) + t.specs.size() * sizeof(ColDesc));
    for (size_t i = 0; i < t.specs.size(); ++i) {
        const auto& s = t.specs[i];
        ColDesc& d = descs[i];
        std::memset(&d, 0, sizeof(d));
        d.columnIndex = s.index;
        d.dtype = static_cast<uint8_t>(s.type);
        d.fixedWidth = static_cast<uint16_t>(s.type == TpchType::CHAR_FIXED ? s.width : 0);
        d.sizeBytes = t.rows * elemBytes(s.type, s.width);
        if (t.data[i].size() != d.sizeBytes) throw std::runtime_error("writeColbin: column payload size mismatch");
        d.offset = cursor;
        cursor = alignUp(cursor + d.sizeBytes);
    }
    FileHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.nCols = static_cast<uint32_t>(t.specs.size());
    h.nRows = t.rows;
    h.sourceSize = srcSize;
    h.sourceMtimeNs = srcMtime;

    const std::string dst = colbinPath(tblPath), tmp = dst + ".tmp";
    const int fd = ::open(tmp.c_str(), O_CREAT | O_RDWR | O_TRUNC, 0644);
    if (fd < 0) throw std::runtime_error("writeColbin: cannot create " + tmp);
    auto fail = [&](const std::string& what) {
        ::close(fd);
        ::unlink(tmp.c_str());
        throw std::runtime_error("writeColbin: " + what + " (" + tmp + ")");
    };
    auto writeAll = [&](const void* p, uint64_t n, uint64_t off) {
        const char* c = static_cast<const char*>(p);
        while (n > 0) {
            const size_t chunk = static_cast<size_t>(std::min<uint64_t>(n, 1ull << 30));
            const ssize_t w = ::pwrite(fd, c, chunk, static_cast<off_t>(off));
            if (w <= 0) fail("write failed");
            c += w;
            off += static_cast<uint64_t>(w);
            n -= static_cast<uint64_t>(w);
        }
    };
    if (::ftruncate(fd, static_cast<off_t>(cursor)) != 0) fail("ftruncate failed");
    writeAll(&h, sizeof(h), 0);
    writeAll(descs.data(), descs.size() * sizeof(ColDesc), sizeof(FileHeader));
    for (size_t i = 0; i < descs.size(); ++i) writeAll(t.data[i].data(), descs[i].sizeBytes, descs[i].offset);
    ::fsync(fd);
    ::close(fd);
    if (::rename(tmp.c_str(), dst.c_str()) != 0) {
        ::unlink(tmp.c_str());
        throw std::runtime_error("writeColbin: rename to " + dst + " failed");
    }
}

bool colbinUpToDate(const std::string& tblPath, const std::vector<ColSpec>& specs) {
    uint64_t srcSize = 0;
    int64_t srcMtime = 0;
    if (!statFile(tblPath, srcSize, srcMtime)) return false;
    try {
        ColbinFile f(colbinPath(tblPath), tblPath);
        if (f.columns().size() != specs.size()) return false;
        for (const auto& s : specs) {
            const auto& c = f.column(s.index);
            if (c.type != s.type) return false;
            if (s.type == TpchType::CHAR_FIXED && c.width != s.width) return false;
            if (c.bytes != f.rows() * elemBytes(s.type, s.width)) return false;
        }
        return true;
    } catch (const std::exception&) {
        return false;
    }
}

bool tblToColbin(const std::string& tblPath, const std::string& table, bool force, int threads) {
    const auto specs = specsOf(tpchTable(table));
    if (!force && colbinUpToDate(tblPath, specs)) return false;
    writeColbin(tblPath, parseTbl(tblPath, specs, threads));
    return true;
}

ColbinFile::ColbinFile(const std::string& path, const std::string& tblPath) {
    const int fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) throw std::runtime_error("colbin: cannot open " + path);
    struct stat st{};
    if (::fstat(fd, &st) != 0) {
        ::close(fd);
        throw std::runtime_error("colbin: cannot stat " + path);
    }
    size_ = static_cast<size_t>(st.st_size);
    if (size_ < sizeof(FileHeader)) {
        ::close(fd);
        throw std::runtime_error("colbin: " + path + " is shorter than its header");
    }
    map_ = ::mmap(nullptr, size_, PROT_READ, MAP_SHARED, fd, 0);
    ::close(fd);
    if (map_ == MAP_FAILED) {
        map_ = nullptr;
        throw std::runtime_error("colbin: cannot map " + path);
    }
    auto fail = [&](const std::string& why) {
        ::munmap(map_, size_);
        map_ = nullptr;
        throw std::runtime_error("colbin: " + path + ": " + why);
    };
    const uint8_t* base = static_cast<const uint8_t*>(map_);
    FileHeader h;
    std::memcpy(&h, base, sizeof(h));
    if (std::memcmp(h.magic, kMagic, 8) != 0) fail("bad magic");
    if (h.version != kVersion) fail("version " + std::to_string(h.version) + ", expected 2");
    uint64_t srcSize = 0;
    int64_t srcMtime = 0;
    if (!tblPath.empty() && statFile(tblPath, srcSize, srcMtime)) {
        if (h.sourceSize != srcSize || h.sourceMtimeNs != srcMtime) fail("stale: " + tblPath + " changed since it was built");
    }
    if (sizeof(FileHeader) + static_cast<uint64_t>(h.nCols) * sizeof(ColDesc) > size_) fail("descriptors out of range");
    rows_ = h.nRows;
    for (uint32_t i = 0; i < h.nCols; ++i) {
        ColDesc d;
        std::memcpy(&d, base + sizeof(FileHeader) + i * sizeof(ColDesc), sizeof(d));
        if (d.offset > size_ || d.sizeBytes > size_ - d.offset) fail("payload out of range");
        ColbinColumn c;
        c.index = d.columnIndex;
        c.type = decodeType(d.dtype);
        c.width = d.fixedWidth;
        c.data = base + d.offset;
        c.bytes = d.sizeBytes;
        if (c.bytes != rows_ * elemBytes(c.type, c.width)) fail("column " + std::to_string(c.index) + " size mismatch");
        cols_.push_back(c);
    }
}

ColbinFile::~ColbinFile() {
    if (map_) ::munmap(map_, size_);
}

const ColbinColumn& ColbinFile::column(int index) const {
    for (const auto& c : cols_)
        if (c.index == index) return c;
    throw std::runtime_error("colbin: no column with index " + std::to_string(index));
}

}  // namespace infra
}  // namespace gdb

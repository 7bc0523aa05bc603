#pragma once
// TPC-H ingest: the .tbl text parser, the .colbin v2 column file (writer and
// memory-mapped reader) and the TPC-H catalog whose column indices, types and
// fixed widths both formats are keyed on.
//
// Counterparts in the reference:
//   - catalog: TPCHSchema (codegen/planning/tpch_schema.h:47-166);
//   - .tbl parsing rules: loadColumnsMulti (src/infra.h:151-262) — INT by
//     atoi, FLOAT by strtof, DATE "YYYY-MM-DD" -> YYYYMMDD, CHAR1 the first
//     byte (NUL when empty), CHAR_FIXED truncated / NUL-padded to its width;
//   - .colbin v2: src/infra.h:511-771 (48-byte header, 32-byte column
//     descriptors, payloads at 16 KiB-aligned offsets, the source .tbl's size
//     and mtime recorded so a stale file is refused);
//   - the per-table build: tools/tbl_to_colbin.cpp:106-172.
// The byte layout is kept identical on purpose so either side reads the
// other's files. The parser is this build's own: the .tbl is memory-mapped,
// cut at line boundaries into one range per thread, lines are counted, then
// every thread parses its range straight into its slice of pre-sized columns.
// Lines without any '|' are skipped; a line that lacks a field a requested
// column needs is an error (the reference would misalign the columns).

#include <cstdint>
#include <string>
#include <vector>

namespace gdb {
namespace infra {

enum class TpchType : uint8_t { INT = 0, FLOAT = 1, DATE = 2, CHAR1 = 3, CHAR_FIXED = 4 };

struct TpchColumn {
    std::string name;
    int index = 0;
    TpchType type = TpchType::INT;
    int width = 0;  // CHAR_FIXED only
};

struct TpchTable {
    std::string name;
    std::vector<TpchColumn> columns;
    const TpchColumn& column(const std::string& name) const;
};

const std::vector<TpchTable>& tpchTables();
const TpchTable& tpchTable(const std::string& name);  // throws on an unknown table

struct ColSpec {
    int index = 0;
    TpchType type = TpchType::INT;
    int width = 0;
};

std::vector<ColSpec> specsOf(const TpchTable& t);

// Bytes per row of a column of this type (INT/DATE/FLOAT 4, CHAR1 1, CHAR_FIXED width).
size_t elemBytes(TpchType t, int width);

struct ParsedTable {
    uint64_t rows = 0;
    std::vector<ColSpec> specs;
    std::vector<std::vector<uint8_t>> data;  // one payload per spec, rows * elemBytes
};

ParsedTable parseTbl(const std::string& tblPath, const std::vector<ColSpec>& specs, int threads);

// <base>.colbin next to <base>.tbl (binaryPath, src/infra.h:565-569).
std::string colbinPath(const std::string& tblPath);

// Writes colbinPath(tblPath) through a temporary file and a rename; the
// header records the .tbl's size and mtime.
void writeColbin(const std::string& tblPath, const ParsedTable& t);

// True when colbinPath(tblPath) exists, matches the .tbl's size and mtime and
// holds every spec with the expected type, width, size and alignment.
bool colbinUpToDate(const std::string& tblPath, const std::vector<ColSpec>& specs);

// Parse the .tbl and write the .colbin unless it is up to date (or force).
// Returns true when a file was written.
bool tblToColbin(const std::string& tblPath, const std::string& table, bool force, int threads);

struct ColbinColumn {
    int index = 0;
    TpchType type = TpchType::INT;
    int width = 0;
    const uint8_t* data = nullptr;
    uint64_t bytes = 0;
};

// Read-only memory map of a .colbin file. With a non-empty tblPath whose file
// exists, the recorded source size and mtime must match it.
class ColbinFile {
public:
    ColbinFile(const std::string& colbinPath, const std::string& tblPath);
    ~ColbinFile();
    ColbinFile(const ColbinFile&) = delete;
    ColbinFile& operator=(const ColbinFile&) = delete;

    uint64_t rows() const { return rows_; }
    const std::vector<ColbinColumn>& columns() const { return cols_; }
    const ColbinColumn& column(int index) const;  // throws when absent

private:
    void* map_ = nullptr;
    size_t size_ = 0;
    uint64_t rows_ = 0;
    std::vector<ColbinColumn> cols_;
};

}  // namespace infra
}  // namespace gdb

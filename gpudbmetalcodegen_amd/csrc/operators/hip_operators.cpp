#include "hip_operators.h"

#include <cstdlib>
#include <sstream>
#include <stdexcept>
#include <tuple>
#include <utility>

namespace gdb {
namespace codegen {

namespace {

std::vector<uint8_t> wordsPattern(const std::vector<uint64_t>& words) {
    std::vector<uint8_t> p(words.size() * 8);
    for (size_t i = 0; i < words.size(); ++i)
        for (int b = 0; b < 8; ++b) p[i * 8 + b] = static_cast<uint8_t>(words[i] >> (8 * b));
    return p;
}

constexpr uint64_t kEmptyWord = 0x00000000FFFFFFFFull;
constexpr uint64_t kI64Max = 0x7FFFFFFFFFFFFFFFull;
constexpr uint64_t kI64Min = 0x8000000000000000ull;

std::string s(long long v) { return std::to_string(v); }

// One term of a status-word update: per-thread counter `var` added into word
// `word`, or (flag) word set to 1 when any thread's counter is nonzero.
struct StatusTerm {
    std::string var;
    int word;
    bool flag;
};

// Status words updated once per block: wave sums into LDS, then thread 0
// issues one global atomic per nonzero word. (Per-wave global atomics on the
// same few words serialise at the memory-side atomic unit.) Every thread of
// the block must reach this code.
void emitBlockStatus(HipCodegen& cg, const std::string& st, const std::vector<StatusTerm>& terms) {
    cg.addBlock("", [&]() {
        cg.declareLds("u64", "_stl", "4");
        cg.addLine("if (threadIdx.x < 4) _stl[threadIdx.x] = 0;");
        cg.addLine("__syncthreads();");
        for (const auto& t : terms) {
            cg.addLine("{ const u64 _v = gdb_wave_sum_u64(" + t.var + "); if (gdb_lane() == 0 && _v) gdb_lds_add_u64(&_stl[" +
                       s(t.word) + "], _v); }");
        }
        cg.addLine("__syncthreads();");
        cg.addBlock("if (threadIdx.x == 0)", [&]() {
            for (const auto& t : terms) {
                const std::string w = s(t.word);
                if (t.flag)
                    cg.addLine("if (_stl[" + w + "]) gdb_atomic_max_u64(" + st + " + " + w + ", 1ull);");
                else
                    cg.addLine("if (_stl[" + w + "]) gdb_atomic_add_u64(" + st + " + " + w + ", _stl[" + w + "]);");
            }
        });
    });
}

}  // namespace

void HipOperator::produce(HipCodegen& cg, ConsumerFn consume) {
    const bool consumerStageable = cg.nextConsumerStageable();
    cg.setNextConsumerStageable(stageable());
    produceImpl(cg, [&cg, consume, consumerStageable]() {
        if (cg.staged() && !consumerStageable)
            cg.rowLoop(consume);
        else
            consume();
    });
}

void HipOperator::describeTree(std::string& out, int depth) const {
    out += std::string(depth * 2, ' ') + describe() + "\n";
}

void HipUnaryOperator::describeTree(std::string& out, int depth) const {
    out += std::string(depth * 2, ' ') + describe() + "\n";
    if (child_) child_->describeTree(out, depth + 1);
}

// ===========================================================================
// HipGridStrideScan
// ===========================================================================

HipGridStrideScan::HipGridStrideScan(const std::string& table, const std::string& idxVar)
    : table_(table), idxVar_(idxVar) {}

void HipGridStrideScan::addColumn(const std::string& column, ColumnType type) {
    for (const auto& c : columns_)
        if (c.name == column) return;
    columns_.push_back({column, type, false});
}

void HipGridStrideScan::setLazy(const std::string& column) {
    for (auto& c : columns_)
        if (c.name == column) {
            if (c.type.type != ColType::CHAR)
                throw std::runtime_error("HipGridStrideScan: only char columns are read lazily (" + column + ")");
            c.lazy = true;
            return;
        }
    throw std::runtime_error("HipGridStrideScan(" + table_ + "): no column " + column);
}

void HipGridStrideScan::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    if (columns_.empty())
        throw std::runtime_error("HipGridStrideScan(" + table_ + "): no columns registered; call addColumn() first");
    cg.setPhaseScannedTable(table_);
    std::vector<std::string> ptr;
    for (const auto& c : columns_) ptr.push_back(cg.addColumnParam(table_, c.name, c.type));
    const std::string nParam = cg.addTableSizeParam(table_);
    const bool range = cg.rangeMode() || cg.scanSegment();
    // Grid mode: the grid strides over the whole table. Range mode: this
    // block walks the work item [_rb, _re) (_rb is a multiple of the tile).
    const std::string n = range ? "_re" : nParam;
    const int lb = cg.getPhases().back().launchBound;

    auto staged = [](const ColumnDesc& c) { return c.type.type == ColType::CHAR && !c.lazy; };
    bool anyCharCol = false;
    for (const auto& c : columns_) anyCharCol |= staged(c);
    // Software pipelining across tiles (grid mode, fixed-width columns): the
    // next tile's column loads are in flight while this tile is processed.
    const bool prefetch = cg.getPhases().back().prefetch && !range && !anyCharCol;
    // Column loads of the tile starting at `tileExpr` (empty: the current
    // tile, _r0/_full in scope) into arrays <prefix><column>.
    auto emitLoads = [&](const std::string& tileExpr, const std::string& prefix) {
        cg.addBlock("", [&]() {
            std::string r0 = "_r0", full = "_full";
            if (!tileExpr.empty()) {
                cg.addLine("const u64 _tn = " + tileExpr + ";");
                cg.addLine("const u64 _rn = _tn + (u64)threadIdx.x * ITEMS;");
                cg.addLine("const bool _fn = _tn + _tile <= " + n + ";");
                r0 = "_rn";
                full = "_fn";
            }
            cg.addBlock(tileExpr.empty() ? "if (" + full + ")" : "if (_tn < " + n + " && " + full + ")", [&]() {
                for (size_t k = 0; k < columns_.size(); ++k) {
                    const auto& c = columns_[k];
                    if (c.type.type == ColType::CHAR) continue;
                    cg.addLine("gdb_load_items<" + c.type.cxx() + ", ITEMS>(" + ptr[k] + " + " + r0 + ", " + prefix +
                               c.name + ");");
                }
            }, tileExpr.empty() ? " else {" : " else if (_tn < " + n + ") {");
            {
                IndentGuard g(cg);
                for (size_t k = 0; k < columns_.size(); ++k) {
                    const auto& c = columns_[k];
                    if (c.type.type == ColType::CHAR) continue;
                    cg.addLine("gdb_load_tail<" + c.type.cxx() + ", ITEMS>(" + ptr[k] + ", " + r0 + ", " + n + ", " +
                               prefix + c.name + ");");
                }
            }
            cg.addLine("}");
        });
    };

    cg.addComment("scan " + table_ + ": tiles of blockDim.x*ITEMS rows, ITEMS consecutive rows per thread");
    cg.addBlock("", [&]() {
        cg.addLine("const u64 _tile = (u64)blockDim.x * ITEMS;");
        cg.addLine(range ? "const u64 _step = _tile;" : "const u64 _step = (u64)gridDim.x * _tile;");
        if (prefetch) {
            for (const auto& c : columns_)
                if (c.type.type != ColType::CHAR) cg.addLine(c.type.cxx() + " _p_" + c.name + "[ITEMS];");
            emitLoads("(u64)blockIdx.x * _tile", "_p_");
        }
        for (size_t k = 0; k < columns_.size(); ++k) {
            const auto& c = columns_[k];
            if (staged(c)) {
                // + 16: row readers may load the 4-byte word past a row's end
                cg.declareLds("char", "_s_" + c.name, s(lb) + " * ITEMS * " + s(c.type.width) + " + 16");
            }
        }
        const std::string start = range ? "_rb" : "(u64)blockIdx.x * _tile";
        cg.addBlock("for (u64 _t0 = " + start + "; _t0 < " + n + "; _t0 += _step)", [&]() {
            cg.addLine("const u64 _r0 = _t0 + (u64)threadIdx.x * ITEMS;");
            cg.addLine("const bool _full = _t0 + _tile <= " + n + ";");
            bool anyChar = false;
            for (size_t k = 0; k < columns_.size(); ++k) {
                const auto& c = columns_[k];
                if (c.type.type == ColType::CHAR && c.lazy) continue;
                if (c.type.type == ColType::CHAR) {
                    anyChar = true;
                    const std::string W = s(c.type.width);
                    cg.addComment("stage the tile's " + c.name + " bytes in LDS with 16-byte loads");
                    cg.addBlock("", [&]() {
                        cg.addLine("const u64 _nrows = (_full ? _tile : " + n + " - _t0);");
                        cg.addLine("const u64 _nb = _nrows * " + W + ";");
                        cg.addLine("const char* _src = (const char*)" + ptr[k] + " + _t0 * " + W + ";");
                        cg.addBlock("if (_full && (_nb % 16) == 0)", [&]() {
                            // four 16-byte loads in flight per thread before their LDS stores
                            const std::string dst = "reinterpret_cast<u32x4*>(_s_" + c.name + ")";
                            const std::string src = "reinterpret_cast<const u32x4*>(_src)";
                            cg.addLine("const u64 _nv = _nb / 16, _bd = blockDim.x;");
                            cg.addLine("u64 _k = threadIdx.x;");
                            cg.addBlock("for (; _k + 3 * _bd < _nv; _k += 4 * _bd)", [&]() {
                                cg.addLine("const u32x4 _v0 = gdb_ld_stream(" + src + " + _k), _v1 = gdb_ld_stream(" +
                                           src + " + _k + _bd);");
                                cg.addLine("const u32x4 _v2 = gdb_ld_stream(" + src + " + _k + 2 * _bd), _v3 = gdb_ld_stream(" +
                                           src + " + _k + 3 * _bd);");
                                cg.addLine(dst + "[_k] = _v0; " + dst + "[_k + _bd] = _v1; " + dst + "[_k + 2 * _bd] = _v2; " +
                                           dst + "[_k + 3 * _bd] = _v3;");
                            });
                            cg.addLine("for (; _k < _nv; _k += _bd) " + dst + "[_k] = gdb_ld_stream(" + src + " + _k);");
                        }, " else {");
                        {
                            IndentGuard g(cg);
                            cg.addLine("for (u64 _k = threadIdx.x; _k < _nb; _k += blockDim.x) _s_" + c.name +
                                       "[_k] = _src[_k];");
                        }
                        cg.addLine("}");
                    });
                    continue;
                }
                const std::string T = c.type.cxx();
                cg.addLine(T + " _r_" + c.name + "[ITEMS];");
            }
            if (anyChar) cg.addLine("__syncthreads();");
            if (prefetch) {
                cg.addComment("this tile's columns were loaded one iteration ahead; issue the next tile's now");
                for (const auto& c : columns_) {
                    if (c.type.type == ColType::CHAR) continue;
                    cg.addLine("#pragma unroll");
                    cg.addLine("for (int _u = 0; _u < ITEMS; ++_u) _r_" + c.name + "[_u] = _p_" + c.name + "[_u];");
                }
                emitLoads("_t0 + _step", "_p_");
            } else {
                emitLoads("", "_r_");
            }
            if (cg.phaseStaged()) {
                cg.addLine("bool _act[ITEMS];");
                cg.addLine("#pragma unroll");
                cg.addLine("for (int _u = 0; _u < ITEMS; ++_u) _act[_u] = _full || _r0 + (u64)_u < " + n + ";");
                const size_t mark = cg.rowBindings().size();
                cg.addRowBinding("const u64 " + idxVar_ + " = _r0 + (u64)_u;");
                for (size_t k = 0; k < columns_.size(); ++k) {
                    const auto& c = columns_[k];
                    if (c.type.type == ColType::CHAR && c.lazy)
                        cg.addRowBinding("const char* " + c.name + " = (const char*)" + ptr[k] + " + (_r0 + (u64)_u) * " +
                                         s(c.type.width) + ";");
                    else if (c.type.type == ColType::CHAR)
                        cg.addRowBinding("const char* " + c.name + " = _s_" + c.name +
                                         " + ((u64)threadIdx.x * ITEMS + _u) * " + s(c.type.width) + ";");
                    else
                        cg.addRowBinding("const " + c.type.cxx() + " " + c.name + " = _r_" + c.name + "[_u];");
                }
                cg.setStaged(true);
                consume();
                cg.setStaged(false);
                cg.rowBindings().resize(mark);
                if (anyChar) cg.addLine("__syncthreads();");
                return;
            }
            cg.addLine("#pragma unroll");
            cg.addBlock("for (int _u = 0; _u < ITEMS; ++_u)", [&]() {
                cg.addLine("const u64 " + idxVar_ + " = _r0 + (u64)_u;");
                cg.addBlock("if (_full || " + idxVar_ + " < " + n + ")", [&]() {
                    for (size_t k = 0; k < columns_.size(); ++k) {
                        const auto& c = columns_[k];
                        if (c.type.type == ColType::CHAR && c.lazy) {
                            cg.addLine("const char* " + c.name + " = (const char*)" + ptr[k] + " + " + idxVar_ + " * " +
                                       s(c.type.width) + ";");
                        } else if (c.type.type == ColType::CHAR) {
                            cg.addLine("const char* " + c.name + " = _s_" + c.name + " + ((u64)threadIdx.x * ITEMS + _u) * " +
                                       s(c.type.width) + ";");
                        } else {
                            cg.addLine("const " + c.type.cxx() + " " + c.name + " = _r_" + c.name + "[_u];");
                        }
                    }
                    consume();
                });
            });
            if (anyChar) cg.addLine("__syncthreads();");
        });
    });
}

std::string HipGridStrideScan::describe() const {
    std::string d = "GridStrideScan(" + table_ + ":";
    for (const auto& c : columns_) d += " " + c.name + "/" + c.type.tag();
    return d + ")";
}

// ===========================================================================
// HipHashTableScan
// ===========================================================================

HipHashTableScan::HipHashTableScan(const std::string& table, const std::string& slotsExpr, int slotWords,
                                   const std::string& prefix)
    : table_(table), slotsExpr_(slotsExpr), prefix_(prefix), slotWords_(slotWords) {}

void HipHashTableScan::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    cg.setPhaseRowsExpr(slotsExpr_);
    cg.addBufferParam(table_, "u64", 8, slotsExpr_ + " * " + s(slotWords_), InitKind::None, {}, true, true);
    const std::string n = "nslots_" + table_;
    cg.addSymbolParam(n, slotsExpr_);
    const std::string W = s(slotWords_);
    const int vecs = slotWords_ / 2;  // 16-byte loads per slot (even slot widths)
    cg.addComment("scan the occupied slots of hash table " + table_);
    cg.addBlock("", [&]() {
        cg.addLine("const u64 _tile = (u64)blockDim.x * ITEMS;");
        cg.addBlock("for (u64 _t0 = (u64)blockIdx.x * _tile; _t0 < " + n + "; _t0 += (u64)gridDim.x * _tile)", [&]() {
            // Lane-contiguous slots (consecutive lanes read consecutive
            // slots); all ITEMS slots of a thread are loaded before any is
            // examined, so ITEMS 16-byte loads per lane are in flight at once.
            if (slotWords_ % 2 == 0) {
                cg.addLine("u64x2 _sv[ITEMS][" + s(vecs) + "];");
                cg.addLine("#pragma unroll");
                cg.addBlock("for (int _u = 0; _u < ITEMS; ++_u)", [&]() {
                    cg.addLine("const u64 _slot = _t0 + (u64)_u * blockDim.x + (u64)threadIdx.x;");
                    cg.addLine("#pragma unroll");
                    cg.addLine("for (int _v = 0; _v < " + s(vecs) + "; ++_v) _sv[_u][_v] = _slot < " + n +
                               " ? __builtin_nontemporal_load(reinterpret_cast<const u64x2*>(" + table_ + " + _slot * " + W +
                               ") + _v) : u64x2{(u64)GDB_EMPTY_KEY, 0ull};");
                });
            }
            cg.addLine("#pragma unroll");
            cg.addBlock("for (int _u = 0; _u < ITEMS; ++_u)", [&]() {
                cg.addLine("const u64 _slot = _t0 + (u64)_u * blockDim.x + (u64)threadIdx.x;");
                cg.addBlock("if (_slot < " + n + ")", [&]() {
                    if (slotWords_ % 2 == 0) {
                        cg.addLine("const u64 _w0 = _sv[_u][0].x;");
                        for (int w = 1; w < slotWords_; ++w)
                            cg.addLine("const u64 " + prefix_ + "_w" + s(w) + " = _sv[_u][" + s(w / 2) + "]." +
                                       (w % 2 ? "y" : "x") + ";");
                    } else {
                        cg.addLine("const u64* _sw = " + table_ + " + _slot * " + W + ";");
                        cg.addLine("const u64 _w0 = __builtin_nontemporal_load(_sw);");
                        for (int w = 1; w < slotWords_; ++w)
                            cg.addLine("const u64 " + prefix_ + "_w" + s(w) + " = __builtin_nontemporal_load(_sw + " + s(w) +
                                       ");");
                    }
                    cg.addBlock("if ((u32)_w0 != GDB_EMPTY_KEY)", [&]() {
                        cg.addLine("const u32 " + prefix_ + "_key = (u32)_w0;");
                        cg.addLine("const u32 " + prefix_ + "_payload = (u32)(_w0 >> 32);");
                        consume();
                    });
                });
            });
        });
    });
}

std::string HipHashTableScan::describe() const { return "HashTableScan(" + table_ + ", slots=" + slotsExpr_ + ")"; }

// ===========================================================================
// HipSelection / HipComputeExpr
// ===========================================================================

HipSelection::HipSelection(std::unique_ptr<HipOperator> child, const std::string& predicate)
    : HipUnaryOperator(std::move(child)), predicate_(predicate) {}

void HipSelection::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    child_->produce(cg, [&]() {
        if (cg.staged()) {
            cg.rowLoop([&]() { cg.addLine("_act[_u] = (" + predicate_ + ");"); });
            consume();
            return;
        }
        cg.addIf(predicate_, [&]() { consume(); });
    });
}

std::string HipSelection::describe() const { return "Selection(" + predicate_ + ")"; }

HipComputeExpr::HipComputeExpr(std::unique_ptr<HipOperator> child, const std::string& varName,
                               const std::string& varType, const std::string& expression)
    : HipUnaryOperator(std::move(child)), varName_(varName), varType_(varType), expression_(expression) {}

void HipComputeExpr::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    child_->produce(cg, [&]() {
        const std::string decl = "const " + varType_ + " " + varName_ + " = " + expression_ + ";";
        if (cg.staged()) {
            cg.addRowBinding(decl);
            consume();
            cg.rowBindings().pop_back();
            return;
        }
        cg.addLine(decl);
        consume();
    });
}

std::string HipComputeExpr::describe() const { return "ComputeExpr(" + varName_ + " = " + expression_ + ")"; }

// ===========================================================================
// HipWaveCompact
// ===========================================================================

HipWaveCompact::HipWaveCompact(std::unique_ptr<HipOperator> child, std::vector<Var> vars)
    : HipUnaryOperator(std::move(child)), vars_(std::move(vars)) {
    if (vars_.empty()) throw std::runtime_error("HipWaveCompact: no values to queue");
}

void HipWaveCompact::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    if (!cg.phaseStaged()) throw std::runtime_error("HipWaveCompact: the phase must be staged");
    // one queue region of 64 x ITEMS entries per wave the launch bound allows
    const int lb = cg.getPhases().back().launchBound;
    const std::string cap = s((lb + 63) / 64 * 64) + " * ITEMS";
    std::vector<std::string> q;
    for (const auto& v : vars_) {
        q.push_back(cg.uniqueName("_wq"));
        cg.declareLds(v.type, q.back(), cap);
    }
    child_->produce(cg, [&]() {
        if (!cg.staged()) throw std::runtime_error("HipWaveCompact: the child must hand over a staged batch");
        const std::string N = cg.uniqueName("_wn"), B = cg.uniqueName("_wb");
        cg.addLine("const u32 " + B + " = (u32)gdb_wave() * 64u * ITEMS;");
        cg.addLine("u32 " + N + " = 0;");
        cg.addComment("append the batch's active rows to this wave's queue (slot order: row slot, then lane)");
        cg.addLine("#pragma unroll");
        cg.addBlock("for (int _u = 0; _u < ITEMS; ++_u)", [&]() {
            cg.addLine("const u64 _m = __ballot(_act[_u]);");
            cg.addBlock("if (_act[_u])", [&]() {
                const std::vector<std::string> binds = cg.rowBindings();
                for (const auto& b : binds) cg.addLine(b);
                cg.addLine("const u32 _p = " + B + " + " + N + " + (u32)__builtin_popcountll(_m & gdb_lanemask_lt());");
                for (size_t k = 0; k < vars_.size(); ++k) cg.addLine(q[k] + "[_p] = " + vars_[k].name + ";");
            });
            cg.addLine(N + " += (u32)__builtin_popcountll(_m);");
        });
        cg.addLine("gdb_wave_sync();");
        cg.addComment("the consumer runs once per queued row, lanes dense");
        cg.addBlock("for (u32 _j = (u32)gdb_lane(); _j < " + N + "; _j += 64u)", [&]() {
            for (size_t k = 0; k < vars_.size(); ++k)
                cg.addLine("const " + vars_[k].type + " " + vars_[k].name + " = " + q[k] + "[" + B + " + _j];");
            cg.setStaged(false);
            consume();
            cg.setStaged(true);
        });
        cg.addLine("gdb_wave_sync();  // the next tile's appends reuse the queue");
    });
}

std::string HipWaveCompact::describe() const {
    std::string d = "WaveCompact(";
    for (size_t k = 0; k < vars_.size(); ++k) d += (k ? ", " : "") + vars_[k].type + " " + vars_[k].name;
    return d + ")";
}

// ===========================================================================
// Bitmaps
// ===========================================================================

HipBitmapBuild::HipBitmapBuild(std::unique_ptr<HipOperator> child, const std::string& bitmapName,
                               const std::string& keyExpr, const std::string& bitsExpr, bool skewed)
    : HipUnaryOperator(std::move(child)), bitmap_(bitmapName), key_(keyExpr), bits_(bitsExpr), skewed_(skewed) {}

void HipBitmapBuild::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    cg.addBitmapWriteParam(bitmap_, bits_);
    // Per-block direct-mapped cache of recently set bitmap words (word index
    // in the high half, the bits this block set in the low half, one 64-bit
    // LDS word so an entry is never torn): repeated keys (skewed or foreign-
    // key columns) skip the global atomic after their first set in a block.
    const std::string C = "_bmc_" + bitmap_;
    cg.declareLds("u64", C, "1024");
    cg.addLine("for (int _c = (int)threadIdx.x; _c < 1024; _c += (int)blockDim.x) " + C + "[_c] = ~0ull;");
    cg.addLine("__syncthreads();");
    child_->produce(cg, [&]() {
        if (cg.staged()) {
            // LDS cache check for the whole batch, then the atomics (keys
            // near-unique: no global read)
            const std::string k = "((u64)(i64)(" + key_ + "))";
            const std::string nb = "nbits_" + bitmap_;
            cg.rowLoop([&]() {
                cg.addBlock("if (" + k + " < " + nb + ")", [&]() {
                    cg.addLine("const u32 _w = (u32)(" + k + " >> 5), _m = 1u << ((u32)" + k + " & 31u);");
                    cg.addLine("u64* _e = " + C + " + (_w & 1023u);");
                    cg.addLine("const u64 _c = __hip_atomic_load(_e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);");
                    cg.addLine("const bool _same = (u32)(_c >> 32) == _w;");
                    cg.addBlock("if (!(_same && ((u32)_c & _m)))", [&]() {
                        cg.addLine("gdb_atomic_or_u32(" + bitmap_ + " + _w, _m);");
                        cg.addLine("gdb_bitmap_cache_note(_e, _c, _same, _w, _m);");
                    });
                });
            });
            consume();
            return;
        }
        cg.addLine(std::string("gdb_bitmap_set_cached<") + (skewed_ ? "true" : "false") + ">(" + bitmap_ + ", nbits_" +
                   bitmap_ + ", (i64)(" + key_ + "), " + C + ");");
        consume();
    });
}

std::string HipBitmapBuild::describe() const { return "BitmapBuild(" + bitmap_ + ", key=" + key_ + ")"; }

HipBitmapProbe::HipBitmapProbe(std::unique_ptr<HipOperator> child, const std::string& bitmapName,
                               const std::string& keyExpr, const std::string& bitsExpr, bool anti)
    : HipUnaryOperator(std::move(child)), bitmap_(bitmapName), key_(keyExpr), bits_(bitsExpr), anti_(anti) {}

void HipBitmapProbe::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    cg.addBitmapReadParam(bitmap_, bits_);
    const bool cache = ldsWords_ > 0 && (ldsPlain_ || cg.rangeMode());
    const std::string C = "_bmc_" + bitmap_, CN = "_bmn_" + bitmap_;
    if (cache) {
        cg.addComment(cg.rangeMode() ? "the bitmap's first words in LDS for this work item"
                                     : "the bitmap's first words in LDS, once per (persistent) block");
        cg.declareLds("u32", C, s(ldsWords_));
        cg.addLine("const u64 " + CN + " = (nbits_" + bitmap_ + " + 31ull) >> 5 < " + s(ldsWords_) + "ull ? (nbits_" +
                   bitmap_ + " + 31ull) >> 5 : " + s(ldsWords_) + "ull;");
        cg.addLine("for (u64 _w = threadIdx.x; _w < " + CN + "; _w += blockDim.x) " + C + "[_w] = " + bitmap_ + "[_w];");
        cg.addLine("__syncthreads();");
    }
    // the word holding bit k (k below nbits)
    auto word = [&](const std::string& k) {
        return cache ? "(" + k + " >> 5 < " + CN + " ? " + C + "[" + k + " >> 5] : " + bitmap_ + "[" + k + " >> 5])"
                     : bitmap_ + "[" + k + " >> 5]";
    };
    child_->produce(cg, [&]() {
        if (cg.staged()) {
            // all bitmap words of the batch first, then the bit tests
            const std::string A = cg.uniqueName("_bw");
            const std::string k = "((u64)(i64)(" + key_ + "))";
            cg.addLine("u32 " + A + "[ITEMS];");
            cg.rowLoop([&]() {
                cg.addLine(A + "[_u] = " + k + " < nbits_" + bitmap_ + " ? " + word(k) + " : 0u;");
            });
            cg.rowLoop([&]() {
                cg.addLine("_act[_u] = ((" + A + "[_u] >> ((u32)" + k + " & 31u)) & 1u) " + (anti_ ? "==" : "!=") +
                           " 0u;");
            });
            consume();
            return;
        }
        if (cache) {
            const std::string k = "((u64)(i64)(" + key_ + "))";
            cg.addIf("((" + k + " < nbits_" + bitmap_ + " ? " + word(k) + " : 0u) >> ((u32)" + k + " & 31u) & 1u) " +
                         (anti_ ? "==" : "!=") + " 0u",
                     [&]() { consume(); });
            return;
        }
        cg.addIf(std::string(anti_ ? "!" : "") + "gdb_bitmap_test(" + bitmap_ + ", nbits_" + bitmap_ + ", (i64)(" +
                     key_ + "))",
                 [&]() { consume(); });
    });
}

std::string HipBitmapProbe::describe() const {
    return std::string(anti_ ? "AntiBitmapProbe(" : "BitmapProbe(") + bitmap_ + ", key=" + key_ + ")";
}

// ===========================================================================
// Hash build / probe
// ===========================================================================

HipHashBuild::HipHashBuild(std::unique_ptr<HipOperator> child, const std::string& table, const std::string& keyExpr,
                           const std::string& payloadExpr, const std::string& slotsExpr, int slotWords)
    : HipUnaryOperator(std::move(child)),
      table_(table),
      key_(keyExpr),
      payload_(payloadExpr),
      slots_(slotsExpr),
      slotWords_(slotWords) {}

void HipHashBuild::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    std::vector<uint64_t> pat(static_cast<size_t>(slotWords_), 0);
    pat[0] = kEmptyWord;
    cg.addBufferParam(table_, "u64", 8, slots_ + " * " + s(slotWords_), InitKind::Pattern, wordsPattern(pat), false,
                      true);
    cg.addSymbolParam("nslots_" + table_, slots_);
    const std::string st = table_ + "_status";
    cg.addBufferParam(st, "u64", 8, "4", InitKind::Zero, {}, false, true);
    cg.setPhaseStatusBuffer(st);
    cg.addLine("const u32 _bits_" + table_ + " = (u32)__builtin_ctzll(nslots_" + table_ + ");");
    cg.addLine("const u32 _maxp_" + table_ + " = (u32)(nslots_" + table_ + " < 65536ull ? nslots_" + table_ +
               " : 65536ull);");
    cg.addLine("u32 _ins_" + table_ + " = 0, _ovf_" + table_ + " = 0, _dup_" + table_ + " = 0;");
    child_->produce(cg, [&]() {
        if (cg.staged()) {
            // Home-slot CAS of every row of the batch in flight together (the
            // returned occupants are only examined in the next loop); rows
            // whose home slot holds another key probe on in lock-step rounds.
            const std::string T = table_, W = s(slotWords_);
            const std::string HS = cg.uniqueName("_bs"), CU = cg.uniqueName("_bc"), PD = cg.uniqueName("_bp"),
                              ANY = cg.uniqueName("_ba");
            const std::string K = "(u32)(" + key_ + ")";
            const std::string MINE = "((u64)" + K + " | ((u64)(u32)(" + payload_ + ") << 32))";
            cg.addLine("u64 " + HS + "[ITEMS];");
            cg.addLine("u64 " + CU + "[ITEMS];");
            cg.addLine("bool " + PD + "[ITEMS];");
            cg.addLine("bool " + ANY + " = false;");
            cg.addLine("#pragma unroll");
            cg.addLine("for (int _u = 0; _u < ITEMS; ++_u) " + PD + "[_u] = false;");
            cg.rowLoop([&]() {
                cg.addLine(HS + "[_u] = gdb_hash_slot(" + K + ", _bits_" + T + ");");
                cg.addLine("u64 _c = GDB_EMPTY_WORD;");
                cg.addLine("(void)__hip_atomic_compare_exchange_strong(" + T + " + " + HS + "[_u] * " + W + ", &_c, " +
                           MINE + ", __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);");
                cg.addLine(CU + "[_u] = _c;");
            });
            auto resolve = [&]() {
                cg.addLine("const u64 _c = " + CU + "[_u];");
                cg.addLine("if (_c == GDB_EMPTY_WORD) { ++_ins_" + T + "; " + PD + "[_u] = false; }");
                cg.addLine("else if ((u32)_c == " + K + ") { ++_dup_" + T + "; " + PD + "[_u] = false; }");
                cg.addLine("else { " + PD + "[_u] = true; " + ANY + " = true; }");
            };
            cg.rowLoop(resolve);
            cg.addBlock("for (u32 _p = 1; " + ANY + " && _p < _maxp_" + T + "; ++_p)", [&]() {
                cg.addLine(ANY + " = false;");
                cg.rowLoop([&]() {
                    cg.addLine(HS + "[_u] = (" + HS + "[_u] + 1ull) & ((1ull << _bits_" + T + ") - 1ull);");
                    cg.addLine("u64 _c = GDB_EMPTY_WORD;");
                    cg.addLine("(void)__hip_atomic_compare_exchange_strong(" + T + " + " + HS + "[_u] * " + W +
                               ", &_c, " + MINE + ", __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);");
                    cg.addLine(CU + "[_u] = _c;");
                }, PD);
                cg.rowLoop(resolve, PD);
            });
            cg.rowLoop([&]() { cg.addLine("if (" + PD + "[_u]) ++_ovf_" + T + ";"); });
            consume();
            return;
        }
        cg.addLine("const i64 _hs = gdb_ht_insert_unique(" + table_ + ", " + s(slotWords_) + ", _bits_" + table_ +
                   ", (u32)(" + key_ + "), (u32)(" + payload_ + "), _maxp_" + table_ + ");");
        cg.addLine("if (_hs >= 0) ++_ins_" + table_ + "; else if (_hs == -1) ++_ovf_" + table_ + "; else ++_dup_" +
                   table_ + ";");
        consume();
    });
    cg.addComment("build status: overflow / duplicate flags and inserted-key count");
    emitBlockStatus(cg, st, {{"_ins_" + table_, 2, false}, {"_ovf_" + table_, 0, true}, {"_dup_" + table_, 1, true}});
}

std::string HipHashBuild::describe() const {
    return "HashBuild(" + table_ + ", key=" + key_ + ", payload=" + payload_ + ", slots=" + slots_ + ")";
}

HipHashProbe::HipHashProbe(std::unique_ptr<HipOperator> child, const std::string& table, const std::string& keyExpr,
                           const std::string& slotsExpr, int slotWords, const std::string& slotVar)
    : HipUnaryOperator(std::move(child)),
      table_(table),
      key_(keyExpr),
      slots_(slotsExpr),
      slotVar_(slotVar),
      slotWords_(slotWords) {}

void HipHashProbe::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    cg.addBufferParam(table_, "u64", 8, slots_ + " * " + s(slotWords_), InitKind::None, {}, true, true);
    cg.addSymbolParam("nslots_" + table_, slots_);
    cg.addLine("const u32 _bits_" + table_ + " = (u32)__builtin_ctzll(nslots_" + table_ + ");");
    cg.addLine("const u32 _maxp_" + table_ + " = (u32)(nslots_" + table_ + " < 65536ull ? nslots_" + table_ +
               " : 65536ull);");
    child_->produce(cg, [&]() {
        if (cg.staged()) {
            // First slots of the whole batch in flight together; rows whose
            // first slot holds another key continue probing in lock-step
            // rounds (one load per pending row per round). Every array is
            // indexed only inside unrolled loops, so all stay in registers.
            const std::string T = table_, W = s(slotWords_);
            const std::string HW = cg.uniqueName("_hw"), HS = cg.uniqueName("_hs"), HP = cg.uniqueName("_hp");
            const std::string PD = cg.uniqueName("_pd"), ANY = cg.uniqueName("_any");
            const std::string K = "(u32)(" + key_ + ")";
            cg.addLine("u64 " + HW + "[ITEMS];");
            cg.addLine("i64 " + HS + "[ITEMS];");
            cg.addLine("u32 " + HP + "[ITEMS];");
            cg.addLine("bool " + PD + "[ITEMS];");
            cg.addLine("bool " + ANY + " = false;");
            cg.rowLoop([&]() {
                cg.addLine(HS + "[_u] = (i64)gdb_hash_slot(" + K + ", _bits_" + T + ");");
                cg.addLine(HW + "[_u] = " + T + "[(u64)" + HS + "[_u] * " + W + "];");
            });
            cg.addLine("#pragma unroll");
            cg.addLine("for (int _u = 0; _u < ITEMS; ++_u) " + PD + "[_u] = false;");
            cg.rowLoop([&]() {
                cg.addLine("const u64 _w = " + HW + "[_u];");
                cg.addLine(HP + "[_u] = (u32)(_w >> 32);");
                cg.addBlock("if ((u32)_w != " + K + ")", [&]() {
                    cg.addLine("_act[_u] = false;");
                    cg.addLine(PD + "[_u] = (u32)_w != GDB_EMPTY_KEY;");
                    cg.addLine(ANY + " = " + ANY + " || " + PD + "[_u];");
                });
            });
            cg.addBlock("for (u32 _p = 1; " + ANY + " && _p < _maxp_" + T + "; ++_p)", [&]() {
                cg.addLine(ANY + " = false;");
                cg.rowLoop([&]() {
                    cg.addLine(HS + "[_u] = (i64)(((u64)" + HS + "[_u] + 1ull) & ((1ull << _bits_" + T + ") - 1ull));");
                    cg.addLine(HW + "[_u] = " + T + "[(u64)" + HS + "[_u] * " + W + "];");
                }, PD);
                cg.rowLoop([&]() {
                    cg.addLine("const u64 _w = " + HW + "[_u];");
                    cg.addLine("if ((u32)_w == " + K + ") { _act[_u] = true; " + PD + "[_u] = false; " + HP +
                               "[_u] = (u32)(_w >> 32); }");
                    cg.addLine("else if ((u32)_w == GDB_EMPTY_KEY) " + PD + "[_u] = false;");
                    cg.addLine("else " + ANY + " = true;");
                }, PD);
            });
            cg.addRowBinding("const i64 " + slotVar_ + " = " + HS + "[_u];");
            cg.addRowBinding("const u32 " + slotVar_ + "_payload = " + HP + "[_u];");
            consume();
            cg.rowBindings().resize(cg.rowBindings().size() - 2);
            return;
        }
        cg.addLine("const i64 " + slotVar_ + " = gdb_ht_find(" + table_ + ", " + s(slotWords_) + ", _bits_" + table_ +
                   ", (u32)(" + key_ + "), _maxp_" + table_ + ");");
        cg.addIf(slotVar_ + " >= 0", [&]() {
            cg.addLine("const u32 " + slotVar_ + "_payload = (u32)(" + table_ + "[" + slotVar_ + " * " +
                       s(slotWords_) + "] >> 32);");
            consume();
        });
    });
}

std::string HipHashProbe::describe() const { return "HashProbe(" + table_ + ", key=" + key_ + ")"; }

HipSlotAtomicAdd::HipSlotAtomicAdd(std::unique_ptr<HipOperator> child, const std::string& table, int slotWords,
                                   const std::string& slotVar, int word, const std::string& valueExpr)
    : HipUnaryOperator(std::move(child)),
      table_(table),
      slotVar_(slotVar),
      value_(valueExpr),
      slotWords_(slotWords),
      word_(word) {}

void HipSlotAtomicAdd::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    // The probe registered the table read-only; the atomic makes it writable.
    cg.addBufferParam(table_, "u64", 8, "", InitKind::None, {}, false, true);
    child_->produce(cg, [&]() {
        cg.addLine("gdb_atomic_add_u64(" + table_ + " + " + slotVar_ + " * " + s(slotWords_) + " + " + s(word_) +
                   ", (u64)(i64)(" + value_ + "));");
        consume();
    });
}

std::string HipSlotAtomicAdd::describe() const {
    return "SlotAtomicAdd(" + table_ + "[" + slotVar_ + "].w" + s(word_) + " += " + value_ + ")";
}

HipArrayLookup::HipArrayLookup(std::unique_ptr<HipOperator> child, const std::string& arrayName,
                               const std::string& keyExpr, const std::string& sizeExpr, const std::string& resultVar,
                               long long sentinel, const std::string& elemType)
    : HipUnaryOperator(std::move(child)),
      array_(arrayName),
      key_(keyExpr),
      size_(sizeExpr),
      var_(resultVar),
      elem_(elemType),
      sentinel_(sentinel) {
    if (elem_ != "u32" && elem_ != "i64") throw std::runtime_error("HipArrayLookup: element type must be u32 or i64");
}

void HipArrayLookup::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    cg.addBufferParam(array_, elem_, elem_ == "u32" ? 4 : 8, size_, InitKind::None, {}, true, true);
    cg.addSymbolParam("nelem_" + array_, size_);
    child_->produce(cg, [&]() {
        cg.addLine("const u64 _ak_" + var_ + " = (u64)(i64)(" + key_ + ");");
        cg.addLine("const i64 " + var_ + " = _ak_" + var_ + " < nelem_" + array_ + " ? (i64)" + array_ + "[_ak_" +
                   var_ + "] : (i64)(" + s(sentinel_) + ");");
        cg.addIf(var_ + " != (i64)(" + s(sentinel_) + ")", [&]() { consume(); });
    });
}

std::string HipArrayLookup::describe() const { return "ArrayLookup(" + var_ + " = " + array_ + "[" + key_ + "])"; }

// ===========================================================================
// HipLikeFilter
// ===========================================================================

HipLikeFilter::HipLikeFilter(std::unique_ptr<HipOperator> child, const std::string& column, int width,
                             const std::string& wordsBuffer, const std::string& l1Scalar, const std::string& l2Scalar,
                             bool negate)
    : HipUnaryOperator(std::move(child)),
      column_(column),
      width_(width),
      words_(wordsBuffer),
      l1_(l1Scalar),
      l2_(l2Scalar),
      negate_(negate) {}

void HipLikeFilter::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    const std::string t1 = cg.uniqueName("_lk1"), t2 = cg.uniqueName("_lk2");
    cg.declareLds("u32", t1, "256");
    cg.declareLds("u32", t2, "256");
    cg.addLine("gdb_like2_tables(" + t1 + ", " + t2 + ", (const char*)" + words_ + ", (int)" + l1_ + ", (const char*)" +
               words_ + " + 32, (int)" + l2_ + ");");
    cg.addLine("__syncthreads();");
    child_->produce(cg, [&]() {
        cg.addIf(std::string(negate_ ? "!" : "") + "gdb_like2_bp<" + s(width_) + ">(" + column_ + ", " + t1 + ", " + t2 +
                     ", (int)" + l1_ + ", (int)" + l2_ + ")",
                 [&]() { consume(); });
    });
}

std::string HipLikeFilter::describe() const {
    return std::string(negate_ ? "NotLike(" : "Like(") + column_ + ", '%w1%w2%')";
}

// ===========================================================================
// HipBitmapFlag, HipBufferScalars, HipDenseAtomicAdd, HipAppendRows
// ===========================================================================

HipBitmapFlag::HipBitmapFlag(std::unique_ptr<HipOperator> child, const std::string& bitmapName,
                             const std::string& keyExpr, const std::string& bitsExpr, const std::string& var)
    : HipUnaryOperator(std::move(child)), bitmap_(bitmapName), key_(keyExpr), bits_(bitsExpr), var_(var) {}

void HipBitmapFlag::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    cg.addBitmapReadParam(bitmap_, bits_);
    child_->produce(cg, [&]() {
        cg.addLine("const u32 " + var_ + " = gdb_bitmap_test(" + bitmap_ + ", nbits_" + bitmap_ + ", (i64)(" + key_ +
                   ")) ? 1u : 0u;");
        consume();
    });
}

std::string HipBitmapFlag::describe() const {
    return "BitmapFlag(" + var_ + " = " + key_ + " in " + bitmap_ + ")";
}

HipByteFlagBuild::HipByteFlagBuild(std::unique_ptr<HipOperator> child, const std::string& flagsName,
                                   const std::string& keyExpr, const std::string& sizeExpr,
                                   const std::string& valueExpr, bool tested)
    : HipUnaryOperator(std::move(child)),
      flags_(flagsName),
      key_(keyExpr),
      size_(sizeExpr),
      value_(valueExpr),
      tested_(tested) {}

void HipByteFlagBuild::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    cg.addBufferParam(flags_, "u8", 1, size_, InitKind::Zero);
    cg.addSymbolParam("nflags_" + flags_, size_);
    const std::string C = "_fc_" + flags_;
    if (tested_) {
        // Per-block direct-mapped LDS cache of keys whose flag this block has
        // seen set or set itself: a hit skips the global read (the random
        // byte reads, not HBM, bound this loop). A key enters the cache only
        // after its lane has read the flag set or issued the store, so a hit
        // never hides a missing flag; the cache holds key + 1 (0 = empty).
        cg.declareLds("u32", C, "4096");
        cg.addLine("for (int _c = (int)threadIdx.x; _c < 4096; _c += (int)blockDim.x) " + C + "[_c] = 0u;");
        cg.addLine("__syncthreads();");
    }
    child_->produce(cg, [&]() {
        if (tested_) {
            cg.addBlock("", [&]() {
                cg.addLine("const u64 _fk = (u64)(i64)(" + key_ + ");");
                cg.addLine("const u8 _fv = (u8)(" + value_ + ");");
                cg.addBlock("if (_fk < nflags_" + flags_ + ")", [&]() {
                    cg.addLine("u32* _ce = " + C + " + (((u32)_fk ^ ((u32)_fk >> 12)) & 4095u);");
                    cg.addBlock("if (*_ce != (u32)_fk + 1u)", [&]() {
                        cg.addLine("if (" + flags_ + "[_fk] != _fv) " + flags_ + "[_fk] = _fv;");
                        cg.addLine("*_ce = (u32)_fk + 1u;");
                    });
                });
            });
        } else {
            cg.addLine("{ const u64 _fk = (u64)(i64)(" + key_ + "); if (_fk < nflags_" + flags_ + ") " + flags_ +
                       "[_fk] = (u8)(" + value_ + "); }");
        }
        consume();
    });
}

std::string HipByteFlagBuild::describe() const {
    return "ByteFlagBuild(" + flags_ + "[" + key_ + "] = " + value_ + (tested_ ? ", tested" : "") + ")";
}

void HipFlagsToBitmap::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    cg.addBufferParam(flags_, "u8", 1, bits_, InitKind::None, {}, true, true);
    cg.addBitmapWriteParam(bitmap_, bits_);
    cg.setPhaseRowsExpr("(" + bits_ + " + 31) / 32");
    const std::string nb = "nbits_" + bitmap_;
    cg.addComment("pack 32 byte flags per bitmap word");
    cg.addLine("const u64 _nw = (" + nb + " + 31) / 32;");
    cg.addBlock("for (u64 _w = (u64)blockIdx.x * blockDim.x + threadIdx.x; _w < _nw; _w += (u64)gridDim.x * blockDim.x)",
                [&]() {
                    cg.addLine("const u64 _b0 = _w * 32;");
                    cg.addLine("u32 _word = 0;");
                    cg.addBlock("if (_b0 + 32 <= " + nb + ")", [&]() {
                        cg.addLine("const u32x4 _x = *reinterpret_cast<const u32x4*>(" + flags_ + " + _b0);");
                        cg.addLine("const u32x4 _y = *reinterpret_cast<const u32x4*>(" + flags_ + " + _b0 + 16);");
                        cg.addLine("_word = gdb_pack_flags16(_x) | (gdb_pack_flags16(_y) << 16);");
                    });
                    cg.addBlock("else", [&]() {
                        cg.addLine("for (u32 _k = 0; _k < 32 && _b0 + _k < " + nb + "; ++_k) _word |= (u32)(" + flags_ +
                                   "[_b0 + _k] != 0) << _k;");
                    });
                    cg.addLine(bitmap_ + "[_w] = _word;");
                });
    consume();
}

HipDenseStore::HipDenseStore(std::unique_ptr<HipOperator> child, const std::string& arrayName,
                             const std::string& keyExpr, const std::string& sizeExpr, const std::string& valueExpr)
    : HipUnaryOperator(std::move(child)), array_(arrayName), key_(keyExpr), size_(sizeExpr), value_(valueExpr) {}

void HipDenseStore::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    cg.addBufferParam(array_, "u32", 4, size_, InitKind::Zero);
    cg.addSymbolParam("nelem_" + array_, size_);
    child_->produce(cg, [&]() {
        cg.addLine("{ const u64 _sk = (u64)(i64)(" + key_ + "); if (_sk < nelem_" + array_ + ") " + array_ +
                   "[_sk] = (u32)(" + value_ + "); }");
        consume();
    });
}

std::string HipDenseStore::describe() const { return "DenseStore(" + array_ + "[" + key_ + "] = " + value_ + ")"; }

HipByteFlagProbe::HipByteFlagProbe(std::unique_ptr<HipOperator> child, const std::string& flagsName,
                                   const std::string& keyExpr, const std::string& sizeExpr, bool anti,
                                   const std::string& var, bool keepAll)
    : HipUnaryOperator(std::move(child)),
      flags_(flagsName),
      key_(keyExpr),
      size_(sizeExpr),
      var_(var),
      anti_(anti),
      keepAll_(keepAll) {
    if (anti_ && !var_.empty()) throw std::runtime_error("HipByteFlagProbe: an anti probe defines no value");
    if (keepAll_ && (anti_ || var_.empty()))
        throw std::runtime_error("HipByteFlagProbe: keepAll needs a value variable and no anti");
}

void HipByteFlagProbe::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    cg.addBufferParam(flags_, "u8", 1, size_, InitKind::None, {}, true, true);
    cg.addSymbolParam("nflags_" + flags_, size_);
    child_->produce(cg, [&]() {
        const std::string k = "(u64)(i64)(" + key_ + ")";
        const std::string v = var_.empty() ? cg.uniqueName("_fv") : var_;
        cg.addLine("const u32 " + v + " = " + k + " < nflags_" + flags_ + " ? (u32)" + flags_ + "[" + k + "] : 0u;");
        if (keepAll_) {
            consume();
            return;
        }
        cg.addIf(anti_ ? v + " == 0u" : v + " != 0u", [&]() { consume(); });
    });
}

std::string HipByteFlagProbe::describe() const {
    return std::string(anti_ ? "AntiByteFlagProbe(" : "ByteFlagProbe(") + flags_ + ", key=" + key_ + ")";
}

HipBufferScalars::HipBufferScalars(std::unique_ptr<HipOperator> child, const std::string& buffer,
                                   const std::string& sizeExpr, std::vector<std::pair<std::string, int>> vars)
    : HipUnaryOperator(std::move(child)), buffer_(buffer), size_(sizeExpr), vars_(std::move(vars)) {}

void HipBufferScalars::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    cg.addBufferParam(buffer_, "u64", 8, size_, InitKind::None, {}, true, true);
    for (const auto& v : vars_) cg.addLine("const i64 " + v.first + " = (i64)" + buffer_ + "[" + s(v.second) + "];");
    child_->produce(cg, consume);
}

std::string HipBufferScalars::describe() const {
    std::string d = "BufferScalars(" + buffer_ + ":";
    for (const auto& v : vars_) d += " " + v.first + "=[" + s(v.second) + "]";
    return d + ")";
}

HipDenseAtomicAdd::HipDenseAtomicAdd(std::unique_ptr<HipOperator> child, const std::string& arrayName,
                                     const std::string& indexExpr, const std::string& sizeExpr,
                                     const std::string& valueExpr)
    : HipUnaryOperator(std::move(child)), array_(arrayName), index_(indexExpr), size_(sizeExpr), value_(valueExpr) {}

void HipDenseAtomicAdd::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    cg.addBufferParam(array_, "u64", 8, size_, InitKind::Zero);
    cg.addSymbolParam("nelem_" + array_, size_);
    // Run combining: a thread keeps one pending (index, sum) and issues the
    // atomic only when the index changes (and once at the end), so rows of a
    // clustered key column (a fact table stored in key order, e.g. lineitem
    // by l_orderkey) cost one atomic per run instead of one per row.
    const std::string PK = "_pk_" + array_, PS = "_ps_" + array_, N = "nelem_" + array_;
    cg.addLine("u64 " + PK + " = ~0ull;");
    cg.addLine("u64 " + PS + " = 0;");
    child_->produce(cg, [&]() {
        cg.addBlock("", [&]() {
            cg.addLine("const u64 _di = (u64)(i64)(" + index_ + ");");
            cg.addLine("const u64 _dv = (u64)(i64)(" + value_ + ");");
            cg.addBlock("if (_di == " + PK + ")", [&]() { cg.addLine(PS + " += _dv;"); }, " else {");
            {
                IndentGuard g(cg);
                cg.addLine("if (" + PK + " < " + N + ") gdb_atomic_add_u64(" + array_ + " + " + PK + ", " + PS + ");");
                cg.addLine(PK + " = _di;");
                cg.addLine(PS + " = _dv;");
            }
            cg.addLine("}");
        });
        consume();
    });
    cg.addLine("if (" + PK + " < " + N + ") gdb_atomic_add_u64(" + array_ + " + " + PK + ", " + PS + ");");
}

std::string HipDenseAtomicAdd::describe() const {
    return "DenseAtomicAdd(" + array_ + "[" + index_ + "] += " + value_ + ")";
}

HipHash64Build::HipHash64Build(std::unique_ptr<HipOperator> child, const std::string& table,
                               const std::string& keyExpr, const std::string& payloadExpr, const std::string& slotsExpr,
                               bool distinct, const std::string& slotVar)
    : HipUnaryOperator(std::move(child)),
      table_(table),
      key_(keyExpr),
      payload_(payloadExpr),
      slots_(slotsExpr),
      distinct_(distinct),
      slotVar_(slotVar) {}

void HipHash64Build::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    cg.addBufferParam(table_, "u64", 8, slots_ + " * 2", InitKind::Pattern, wordsPattern({~0ull, 0ull}), false, true);
    cg.addSymbolParam("nslots_" + table_, slots_);
    const std::string st = table_ + "_status";
    cg.addBufferParam(st, "u64", 8, "4", InitKind::Zero, {}, false, true);
    cg.setPhaseStatusBuffer(st);
    cg.addLine("const u32 _bits_" + table_ + " = (u32)__builtin_ctzll(nslots_" + table_ + ");");
    cg.addLine("const u32 _maxp_" + table_ + " = (u32)(nslots_" + table_ + " < 65536ull ? nslots_" + table_ +
               " : 65536ull);");
    cg.addLine("u32 _ins_" + table_ + " = 0, _ovf_" + table_ + " = 0;");
    if (!uniqueMsg_.empty()) {
        const std::string ck = table_ + "_unique";
        cg.addBufferParam(ck, "u64", 8, "4", InitKind::Zero, {}, false, true);
        cg.setPhaseCheck(ck, uniqueMsg_);
        cg.addLine("u32 _dup_" + table_ + " = 0;");
    }
    if (blockFirst_) {
        if (!distinct_) throw std::runtime_error("HipHash64Build(" + table_ + "): the block filter needs distinct");
        cg.declareLds("u64", "_seen_" + table_, "512");
        cg.addLine("for (int _k = (int)threadIdx.x; _k < 512; _k += (int)blockDim.x) _seen_" + table_ + "[_k] = ~0ull;");
        cg.addLine("__syncthreads();");
    }
    child_->produce(cg, [&]() {
        const std::string r = cg.uniqueName("_hr");
        const std::string slot = slotVar_.empty() ? cg.uniqueName("_hs") : slotVar_;
        const std::string k = cg.uniqueName("_hk");
        cg.addLine("i64 " + slot + " = -1;");
        cg.addLine("const u64 " + k + " = (u64)(" + key_ + ");");
        const std::string call = "gdb_ht64_insert_at(" + table_ + ", _bits_" + table_ + ", " + k + ", (u64)(" +
                                 payload_ + "), _maxp_" + table_ + ", &" + slot + ")";
        if (blockFirst_)
            cg.addLine("const int " + r + " = gdb_block_first(_seen_" + table_ + ", " + k + ") ? " + call + " : 0;");
        else
            cg.addLine("const int " + r + " = " + call + ";");
        if (!uniqueMsg_.empty()) cg.addLine("_dup_" + table_ + " += " + r + " == 0 ? 1u : 0u;");
        cg.addLine("_ins_" + table_ + " += " + r + " == 1 ? 1u : 0u;");
        cg.addLine("_ovf_" + table_ + " |= " + r + " < 0 ? 1u : 0u;");
        if (distinct_)
            cg.addIf(r + " == 1", [&]() { consume(); });
        else if (!slotVar_.empty())
            cg.addIf(r + " >= 0", [&]() { consume(); });
        else
            consume();
    });
    emitBlockStatus(cg, st, {{"_ins_" + table_, 2, false}, {"_ovf_" + table_, 0, true}});
    if (!uniqueMsg_.empty())
        emitBlockStatus(cg, table_ + "_unique", {{"_dup_" + table_, 2, false}, {"_dup_" + table_, 0, true}});
}

HipBindColumns::HipBindColumns(std::unique_ptr<HipOperator> child, const std::string& table,
                               std::vector<std::pair<std::string, ColumnType>> columns)
    : HipUnaryOperator(std::move(child)), table_(table), columns_(std::move(columns)) {}

void HipBindColumns::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    for (const auto& c : columns_) cg.addColumnParam(table_, c.first, c.second);
    child_->produce(cg, consume);
}

std::string HipBindColumns::describe() const {
    std::string d = "BindColumns(" + table_ + ":";
    for (const auto& c : columns_) d += " " + c.first;
    return d + ")";
}

std::string HipHash64Build::describe() const {
    return std::string(distinct_ ? "Hash64BuildDistinct(" : "Hash64Build(") + table_ + ", key=" + key_ +
           ", payload=" + payload_ + ", slots=" + slots_ + ")";
}

HipDenseGroupId::HipDenseGroupId(std::unique_ptr<HipOperator> child, const std::string& table,
                                 const std::string& slotVar, const std::string& slotsExpr, const std::string& idArray,
                                 long long cap)
    : HipUnaryOperator(std::move(child)), table_(table), slotVar_(slotVar), slots_(slotsExpr), ids_(idArray), cap_(cap) {}

void HipDenseGroupId::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    const std::string st = table_ + "_status";
    cg.addBufferParam(st, "u64", 8, "4", InitKind::Zero, {}, false, true);
    cg.addBufferParam(ids_, "u32", 4, slots_, InitKind::Zero, {}, false, true);
    child_->produce(cg, [&]() {
        cg.addBlock("", [&]() {
            cg.addLine("const u64 _id = gdb_wave_append(" + st + " + 3);  // one counter atomic per wave");
            cg.addLine("if (_id < " + s(cap_) + "ull) " + ids_ + "[(u64)" + slotVar_ + "] = (u32)(_id + 1);");
            cg.addLine("else gdb_atomic_max_u64(" + st + ", 1ull);");
        });
        consume();
    });
}

std::string HipDenseGroupId::describe() const {
    return "DenseGroupId(" + ids_ + "[" + slotVar_ + "] = next id + 1, cap " + s(cap_) + ")";
}

HipDirectGroupId::HipDirectGroupId(std::unique_ptr<HipOperator> child, const std::string& keyExpr,
                                   const std::string& sizeExpr, const std::string& idArray, const std::string& status,
                                   long long cap, bool blockFirst)
    : HipUnaryOperator(std::move(child)),
      key_(keyExpr),
      size_(sizeExpr),
      ids_(idArray),
      status_(status),
      cap_(cap),
      blockFirst_(blockFirst) {}

void HipDirectGroupId::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    cg.addBufferParam(ids_, "u32", 4, size_, InitKind::Zero, {}, false, true);
    cg.addSymbolParam("nids_" + ids_, size_);
    cg.addBufferParam(status_, "u64", 8, "4", InitKind::Zero, {}, false, true);
    cg.setPhaseStatusBuffer(status_);
    if (blockFirst_) {
        cg.declareLds("u64", "_seen_" + ids_, "512");
        cg.addLine("for (int _k = (int)threadIdx.x; _k < 512; _k += (int)blockDim.x) _seen_" + ids_ + "[_k] = ~0ull;");
        cg.addLine("__syncthreads();");
    }
    child_->produce(cg, [&]() {
        cg.addBlock("", [&]() {
            cg.addLine("const u64 _dk = (u64)(" + key_ + ");");
            // the id is read past the CU cache: a stale zero there would send
            // every later row of the group to the compare-and-swap
            cg.addBlock("if (_dk < nids_" + ids_ + (blockFirst_ ? " && gdb_block_first(_seen_" + ids_ + ", _dk)" : "") +
                            " && __hip_atomic_load(" + ids_ + " + _dk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u)",
                        [&]() {
                            cg.addLine("u32 _exp = 0u;");
                            cg.addBlock("if (__hip_atomic_compare_exchange_strong(" + ids_ +
                                            " + _dk, &_exp, 0xFFFFFFFFu, __ATOMIC_RELAXED, __ATOMIC_RELAXED, "
                                            "__HIP_MEMORY_SCOPE_AGENT))",
                                        [&]() {
                                            cg.addLine("const u64 _id = gdb_wave_append(" + status_ +
                                                       " + 3);  // one counter atomic per wave");
                                            cg.addLine("if (_id < " + s(cap_) + "ull) __hip_atomic_store(" + ids_ +
                                                       " + _dk, (u32)(_id + 1), __ATOMIC_RELAXED, "
                                                       "__HIP_MEMORY_SCOPE_AGENT);");
                                            cg.addLine("else gdb_atomic_max_u64(" + status_ + ", 1ull);");
                                        });
                        });
        });
        consume();
    });
}

HipDenseUniqueStore::HipDenseUniqueStore(std::unique_ptr<HipOperator> child, const std::string& arrayName,
                                         const std::string& keyExpr, const std::string& sizeExpr,
                                         const std::string& valueExpr, const std::string& message)
    : HipUnaryOperator(std::move(child)),
      array_(arrayName),
      key_(keyExpr),
      size_(sizeExpr),
      value_(valueExpr),
      message_(message) {}

void HipDenseUniqueStore::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    cg.addBufferParam(array_, "u32", 4, size_, InitKind::Zero, {}, false, true);
    cg.addSymbolParam("nelem_" + array_, size_);
    const std::string ck = array_ + "_unique";
    cg.addBufferParam(ck, "u64", 8, "4", InitKind::Zero, {}, false, true);
    cg.setPhaseCheck(ck, message_);
    const std::string dup = "_dup_" + array_;
    cg.addLine("u32 " + dup + " = 0;");
    child_->produce(cg, [&]() {
        cg.addBlock("", [&]() {
            cg.addLine("const u64 _sk = (u64)(i64)(" + key_ + ");");
            cg.addBlock("if (_sk < nelem_" + array_ + ")", [&]() {
                cg.addLine("u32 _e = 0u;");
                cg.addLine(dup + " += __hip_atomic_compare_exchange_strong(" + array_ + " + _sk, &_e, (u32)(" + value_ +
                           "), __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ? 0u : 1u;");
            });
        });
        consume();
    });
    emitBlockStatus(cg, ck, {{dup, 2, false}, {dup, 0, true}});
}

std::string HipDenseUniqueStore::describe() const {
    return "DenseUniqueStore(" + array_ + "[" + key_ + "] = " + value_ + ")";
}

std::string HipDirectGroupId::describe() const {
    return "DirectGroupId(" + ids_ + "[" + key_ + "] = next id + 1, cap " + s(cap_) + ")";
}

HipRowCheck::HipRowCheck(std::unique_ptr<HipOperator> child, const std::string& name, const std::string& condition,
                         const std::string& message)
    : HipUnaryOperator(std::move(child)), name_(name), cond_(condition), message_(message) {}

void HipRowCheck::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    const std::string st = name_ + "_status";
    cg.addBufferParam(st, "u64", 8, "4", InitKind::Zero, {}, false, true);
    cg.setPhaseCheck(st, message_);
    const std::string bad = "_bad_" + name_;
    cg.addLine("u32 " + bad + " = 0;");
    child_->produce(cg, [&]() {
        cg.addLine(bad + " += (" + cond_ + ") ? 0u : 1u;");
        consume();
    });
    emitBlockStatus(cg, st, {{bad, 2, false}, {bad, 0, true}});
}

std::string HipRowCheck::describe() const { return "RowCheck(" + cond_ + ")"; }

HipHash64Probe::HipHash64Probe(std::unique_ptr<HipOperator> child, const std::string& table,
                               const std::string& keyExpr, const std::string& slotsExpr, const std::string& slotVar,
                               bool anti)
    : HipUnaryOperator(std::move(child)), table_(table), key_(keyExpr), slots_(slotsExpr), slotVar_(slotVar),
      anti_(anti) {}

void HipHash64Probe::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    cg.addBufferParam(table_, "u64", 8, slots_ + " * 2", InitKind::None, {}, false, true);
    cg.addSymbolParam("nslots_" + table_, slots_);
    cg.addLine("const u32 _pbits_" + table_ + " = (u32)__builtin_ctzll(nslots_" + table_ + ");");
    cg.addLine("const u32 _pmaxp_" + table_ + " = (u32)(nslots_" + table_ + " < 65536ull ? nslots_" + table_ +
               " : 65536ull);");
    child_->produce(cg, [&]() {
        cg.addLine("const i64 " + slotVar_ + " = gdb_ht64_find(" + table_ + ", _pbits_" + table_ + ", (u64)(" + key_ +
                   "), _pmaxp_" + table_ + ");");
        if (anti_) {
            cg.addIf(slotVar_ + " < 0", [&]() { consume(); });
            return;
        }
        cg.addIf(slotVar_ + " >= 0", [&]() {
            cg.addLine("const u64 " + slotVar_ + "_payload = " + table_ + "[(u64)" + slotVar_ + " * 2 + 1];");
            consume();
        });
    });
}

std::string HipHash64Probe::describe() const {
    return std::string(anti_ ? "Hash64AntiProbe(" : "Hash64Probe(") + table_ + ", key=" + key_ + ")";
}

HipDenseAtomicMinMax::HipDenseAtomicMinMax(std::unique_ptr<HipOperator> child, const std::string& arrayName,
                                           const std::string& indexExpr, const std::string& sizeExpr,
                                           const std::string& valueExpr, bool isMax)
    : HipUnaryOperator(std::move(child)),
      array_(arrayName),
      index_(indexExpr),
      size_(sizeExpr),
      value_(valueExpr),
      max_(isMax) {}

void HipDenseAtomicMinMax::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    const uint64_t neutral = max_ ? 0x8000000000000000ull : kI64Max;
    cg.addBufferParam(array_, "i64", 8, size_, InitKind::Pattern, wordsPattern({neutral}));
    cg.addSymbolParam("nelem_" + array_, size_);
    const std::string PK = "_pk_" + array_, PV = "_pv_" + array_, N = "nelem_" + array_;
    const std::string op = max_ ? "gdb_max64" : "gdb_min64";
    const std::string atomic = max_ ? "gdb_atomic_max_i64" : "gdb_atomic_min_i64";
    cg.addLine("u64 " + PK + " = ~0ull;");
    cg.addLine("i64 " + PV + " = " + std::string(max_ ? "GDB_I64_MIN" : "GDB_I64_MAX") + ";");
    child_->produce(cg, [&]() {
        cg.addBlock("", [&]() {
            cg.addLine("const u64 _mi = (u64)(i64)(" + index_ + ");");
            cg.addLine("const i64 _mv = (i64)(" + value_ + ");");
            cg.addBlock("if (_mi == " + PK + ")", [&]() { cg.addLine(PV + " = " + op + "(" + PV + ", _mv);"); }, " else {");
            {
                IndentGuard g(cg);
                cg.addLine("if (" + PK + " < " + N + ") " + atomic + "(" + array_ + " + " + PK + ", " + PV + ");");
                cg.addLine(PK + " = _mi;");
                cg.addLine(PV + " = _mv;");
            }
            cg.addLine("}");
        });
        consume();
    });
    cg.addLine("if (" + PK + " < " + N + ") " + atomic + "(" + array_ + " + " + PK + ", " + PV + ");");
}

std::string HipDenseAtomicMinMax::describe() const {
    return "DenseAtomic" + std::string(max_ ? "Max(" : "Min(") + array_ + "[" + index_ + "], " + value_ + ")";
}

HipAppendRows::HipAppendRows(std::unique_ptr<HipOperator> child, const std::string& out, const std::string& capExpr)
    : HipUnaryOperator(std::move(child)), out_(out), cap_(capExpr) {}

void HipAppendRows::addField(const std::string& name, const std::string& expr, int scale, bool date) {
    fields_.push_back({name, expr, scale, date});
}

void HipAppendRows::orderBy(const std::string& field, bool desc) { order_.push_back({field, desc}); }

void HipAppendRows::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    if (fields_.empty()) throw std::runtime_error("HipAppendRows(" + out_ + "): no fields");
    const std::string W = s(fields_.size());
    const std::string st = out_ + "_status";  // word 0 overflow flag, word 2 records appended
    cg.addBufferParam(out_, "u64", 8, "(" + cap_ + ") * " + W, InitKind::None);
    cg.addBufferParam(st, "u64", 8, "4", InitKind::Zero, {}, false, true);
    cg.setPhaseStatusBuffer(st);
    cg.addSymbolParam("cap_" + out_, cap_);
    child_->produce(cg, [&]() {
        cg.addBlock("", [&]() {
            cg.addLine("const u64 _ap = gdb_wave_append(" + st + " + 2);");
            cg.addBlock("if (_ap < cap_" + out_ + ")", [&]() {
                for (size_t f = 0; f < fields_.size(); ++f)
                    cg.addLine(out_ + "[_ap * " + W + " + " + s(f) + "] = (u64)(i64)(" + fields_[f].expr + ");");
            }, " else {");
            {
                IndentGuard g(cg);
                cg.addLine("gdb_atomic_max_u64(" + st + ", 1ull);");
            }
            cg.addLine("}");
        });
        consume();
    });
    auto& rs = cg.resultSchema();
    rs.kind = HipResultSchema::ROWS;
    rs.rows.bufferName = out_;
    rs.rows.ctlBuffer = st;
    rs.rows.capExpr = cap_;
    rs.rows.fields.clear();
    for (const auto& f : fields_) rs.rows.fields.push_back({f.name, f.scale, f.date});
    rs.rows.orderBy.clear();
    for (const auto& o : order_) {
        int idx = -1;
        for (size_t f = 0; f < fields_.size(); ++f)
            if (fields_[f].name == o.first) idx = static_cast<int>(f);
        if (idx < 0) throw std::runtime_error("HipAppendRows(" + out_ + "): order by unknown field " + o.first);
        rs.rows.orderBy.push_back({idx, o.second});
    }
    rs.rows.limit = limit_;
}

std::string HipAppendRows::describe() const {
    std::string d = "AppendRows(" + out_ + ":";
    for (const auto& f : fields_) d += " " + f.name;
    return d + ")";
}

// ===========================================================================
// HipKeyedAgg
// ===========================================================================

HipKeyedAgg::HipKeyedAgg(std::unique_ptr<HipOperator> child, const std::string& outputName,
                         const std::string& bucketExpr, int numBuckets)
    : HipUnaryOperator(std::move(child)), out_(outputName), bucket_(bucketExpr), numBuckets_(numBuckets) {
    if (numBuckets_ <= 0) throw std::runtime_error("HipKeyedAgg: numBuckets must be positive");
}

void HipKeyedAgg::addAggregate(const std::string& name, AggKind kind, const std::string& valueExpr,
                               int scaleDigits) {
    aggs_.push_back({name, kind, valueExpr, scaleDigits, 0});
}

void HipKeyedAgg::setBucketLabels(const std::vector<std::string>& labels, const std::string& keyColumn) {
    labels_ = labels;
    keyColumn_ = keyColumn;
}

void HipKeyedAgg::addAverage(const std::string& name, const std::string& numerator, const std::string& denominator) {
    avgs_.push_back({name, {numerator, denominator}});
}

// GDB_SCALAR_ATOMIC=1 (read when a plan is generated): the reduction
// ablation of small-bucket KeyedAgg, every row updating the output with
// global atomics instead of registers + wave/LDS reduction + one atomic per
// block (the reference's GPUDB_SCALAR_ATOMIC, metal_operators.cpp:7-13).
static bool scalarAtomicMode() {
    const char* e = std::getenv("GDB_SCALAR_ATOMIC");
    return e && e[0] && e[0] != '0';
}

void HipKeyedAgg::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    if (aggs_.empty()) throw std::runtime_error("HipKeyedAgg(" + out_ + "): no aggregates");
    // Word layout per bucket; up to 4096 buckets are padded to 128 bytes so
    // the per-block atomics of different buckets land on different cache
    // lines (more buckets spread their atomics by themselves: packed).
    int words = 0;
    for (auto& a : aggs_) {
        a.wordOffset = words;
        words += (a.kind == AggKind::SumI128) ? 2 : 1;
    }
    const int wpb = numBuckets_ > 4096 ? words : ((words + 15) / 16) * 16;
    std::vector<uint64_t> pat(static_cast<size_t>(wpb), 0);
    for (const auto& a : aggs_) {
        if (a.kind == AggKind::MinI64) pat[a.wordOffset] = kI64Max;
        if (a.kind == AggKind::MaxI64) pat[a.wordOffset] = kI64Min;
    }
    const std::string NB = s(numBuckets_);
    cg.addBufferParam(out_, "u64", 8, s((long long)numBuckets_ * wpb), InitKind::Pattern, wordsPattern(pat));

    auto neutral = [](AggKind k) -> std::string {
        if (k == AggKind::MinI64) return "GDB_I64_MAX";
        if (k == AggKind::MaxI64) return "GDB_I64_MIN";
        return "0";
    };

    const bool regs = numBuckets_ <= 8;
    const long long ldsBytes = (long long)numBuckets_ * words * 8;
    const bool lds = !regs && ldsBytes <= 48 * 1024;

    if (regs && scalarAtomicMode()) {
        // Ablation (GDB_SCALAR_ATOMIC=1): no register accumulation and no
        // block reduction; every row issues its own global atomics, as the
        // reference's GPUDB_SCALAR_ATOMIC (metal_operators.cpp:7-13, :259).
        cg.addComment("KeyedAgg " + out_ + ": one global atomic per row and slot (reduction ablation)");
        child_->produce(cg, [&]() {
            cg.addLine("u64* _ka_row = " + out_ + " + (u64)(" + (numBuckets_ > 1 ? "(int)(" + bucket_ + ")" : std::string("0")) +
                       ") * " + s(wpb) + "ull;");
            for (const auto& a : aggs_) {
                const std::string dst = "_ka_row + " + s(a.wordOffset);
                const std::string v = "(i64)(" + a.valueExpr + ")";
                switch (a.kind) {
                    case AggKind::SumI64: cg.addLine("gdb_atomic_add_u64(" + dst + ", (u64)" + v + ");"); break;
                    case AggKind::SumI128:
                        cg.addBlock("", [&]() {
                            cg.addLine("const i64 _sv = " + v + ";");
                            cg.addLine("gdb_atomic_add_u64(" + dst + ", (u64)(u32)_sv);");
                            cg.addLine("gdb_atomic_add_u64(_ka_row + " + s(a.wordOffset + 1) + ", (u64)(_sv >> 32));");
                        });
                        break;
                    case AggKind::Count: cg.addLine("gdb_atomic_add_u64(" + dst + ", 1ull);"); break;
                    case AggKind::MinI64: cg.addLine("gdb_atomic_min_i64((i64*)(" + dst + "), " + v + ");"); break;
                    case AggKind::MaxI64: cg.addLine("gdb_atomic_max_i64((i64*)(" + dst + "), " + v + ");"); break;
                }
            }
            consume();
        });
    } else if (regs) {
        cg.addComment("KeyedAgg " + out_ + ": register accumulators, " + NB + " bucket(s)");
        for (int b = 0; b < numBuckets_; ++b) {
            for (const auto& a : aggs_) {
                const std::string base = "_ka_" + out_ + "_" + a.name + "_b" + s(b);
                switch (a.kind) {
                    case AggKind::SumI64: cg.addLine("i64 " + base + " = 0;"); break;
                    case AggKind::SumI128:
                        cg.addLine("u64 " + base + "_lo = 0;");
                        cg.addLine("i64 " + base + "_hi = 0;");
                        break;
                    case AggKind::Count: cg.addLine("u32 " + base + " = 0;"); break;
                    case AggKind::MinI64: cg.addLine("i64 " + base + " = GDB_I64_MAX;"); break;
                    case AggKind::MaxI64: cg.addLine("i64 " + base + " = GDB_I64_MIN;"); break;
                }
            }
        }
        child_->produce(cg, [&]() {
            if (numBuckets_ > 1) cg.addLine("const int _bk = (int)(" + bucket_ + ");");
            for (const auto& a : aggs_)
                if (a.kind != AggKind::Count)
                    cg.addLine("const i64 _v_" + a.name + " = (i64)(" + a.valueExpr + ");");
            for (int b = 0; b < numBuckets_; ++b) {
                auto body = [&]() {
                    for (const auto& a : aggs_) {
                        const std::string base = "_ka_" + out_ + "_" + a.name + "_b" + s(b);
                        const std::string v = "_v_" + a.name;
                        switch (a.kind) {
                            case AggKind::SumI64: cg.addLine(base + " += " + v + ";"); break;
                            case AggKind::SumI128:
                                cg.addLine(base + "_lo += (u64)(u32)" + v + ";");
                                cg.addLine(base + "_hi += (" + v + " >> 32);");
                                break;
                            case AggKind::Count: cg.addLine(base + " += 1u;"); break;
                            case AggKind::MinI64: cg.addLine(base + " = gdb_min64(" + base + ", " + v + ");"); break;
                            case AggKind::MaxI64: cg.addLine(base + " = gdb_max64(" + base + ", " + v + ");"); break;
                        }
                    }
                };
                if (numBuckets_ > 1)
                    cg.addIf("_bk == " + s(b), body);
                else
                    body();
            }
            consume();
        });
        // Slot table: (bucket, word) -> op; op 0 add, 1 min, 2 max.
        std::ostringstream ops, wordIdx;
        int ns = 0;
        std::vector<std::pair<std::string, int>> slotExprs;  // (value expr, op)
        for (int b = 0; b < numBuckets_; ++b) {
            for (const auto& a : aggs_) {
                const std::string base = "_ka_" + out_ + "_" + a.name + "_b" + s(b);
                auto add = [&](const std::string& expr, int op, int w) {
                    if (ns) {
                        ops << ",";
                        wordIdx << ",";
                    }
                    ops << op;
                    wordIdx << (b * wpb + w);
                    slotExprs.push_back({expr, op});
                    ++ns;
                };
                switch (a.kind) {
                    case AggKind::SumI64: add("(u64)" + base, 0, a.wordOffset); break;
                    case AggKind::SumI128:
                        add(base + "_lo", 0, a.wordOffset);
                        add("(u64)" + base + "_hi", 0, a.wordOffset + 1);
                        break;
                    case AggKind::Count: add("(u64)" + base, 0, a.wordOffset); break;
                    case AggKind::MinI64: add(base, 1, a.wordOffset); break;
                    case AggKind::MaxI64: add(base, 2, a.wordOffset); break;
                }
            }
        }
        const std::string NS = s(ns);
        cg.addHelper("__constant__ int " + out_ + "_ops[" + NS + "] = {" + ops.str() + "};\n__constant__ int " + out_ +
                     "_words[" + NS + "] = {" + wordIdx.str() + "};");
        cg.addComment("block reduction: wave DPP sums -> LDS -> one global atomic per slot and block");
        cg.addBlock("", [&]() {
            cg.declareLds("u64", "_red", NS + " * 16");
            cg.addLine("const int _lane = gdb_lane(), _wave = gdb_wave(), _nw = (int)(blockDim.x >> 6);");
            for (int k = 0; k < ns; ++k) {
                const auto& se = slotExprs[static_cast<size_t>(k)];
                std::string red;
                if (se.second == 0)
                    red = "gdb_wave_sum_u64(" + se.first + ")";
                else if (se.second == 1)
                    red = "(u64)gdb_wave_min_i64(" + se.first + ")";
                else
                    red = "(u64)gdb_wave_max_i64(" + se.first + ")";
                cg.addLine("{ const u64 _w = " + red + "; if (_lane == 0) _red[_wave * " + NS + " + " + s(k) +
                           "] = _w; }");
            }
            cg.addLine("__syncthreads();");
            cg.addBlock("for (int _s = (int)threadIdx.x; _s < " + NS + "; _s += (int)blockDim.x)", [&]() {
                cg.addLine("const int _op = " + out_ + "_ops[_s];");
                cg.addLine("u64 _acc = _red[_s];");
                cg.addBlock("for (int _w2 = 1; _w2 < _nw; ++_w2)", [&]() {
                    cg.addLine("const u64 _x = _red[_w2 * " + NS + " + _s];");
                    cg.addLine("_acc = _op == 0 ? _acc + _x : (_op == 1 ? (u64)gdb_min64((i64)_acc, (i64)_x) : "
                               "(u64)gdb_max64((i64)_acc, (i64)_x));");
                });
                cg.addLine("u64* _dst = " + out_ + " + " + out_ + "_words[_s];");
                cg.addLine("if (_op == 0) { if (_acc != 0) gdb_atomic_add_u64(_dst, _acc); }");
                cg.addLine("else if (_op == 1) { if ((i64)_acc != GDB_I64_MAX) gdb_atomic_min_i64((i64*)_dst, (i64)_acc); }");
                cg.addLine("else { if ((i64)_acc != GDB_I64_MIN) gdb_atomic_max_i64((i64*)_dst, (i64)_acc); }");
            });
        });
    } else if (lds) {
        const std::string W = s(words);
        cg.addComment("KeyedAgg " + out_ + ": block-shared LDS table, " + NB + " buckets x " + W + " words");
        cg.declareLds("u64", "_kl_" + out_, s((long long)numBuckets_ * words));
        cg.addBlock("for (int _s = (int)threadIdx.x; _s < " + s((long long)numBuckets_ * words) +
                        "; _s += (int)blockDim.x)",
                    [&]() {
                        cg.addLine("const int _w = _s % " + W + ";");
                        std::string init = "0";
                        for (const auto& a : aggs_) {
                            if (a.kind == AggKind::MinI64)
                                init = "(_w == " + s(a.wordOffset) + " ? (u64)GDB_I64_MAX : " + init + ")";
                            if (a.kind == AggKind::MaxI64)
                                init = "(_w == " + s(a.wordOffset) + " ? (u64)GDB_I64_MIN : " + init + ")";
                        }
                        cg.addLine("_kl_" + out_ + "[_s] = " + init + ";");
                    });
        cg.addLine("__syncthreads();");
        child_->produce(cg, [&]() {
            cg.addLine("const int _bk = (int)(" + bucket_ + ");");
            cg.addBlock("if ((unsigned)_bk < " + NB + "u)", [&]() {
                cg.addLine("u64* _kb = _kl_" + out_ + " + _bk * " + W + ";");
                for (const auto& a : aggs_) {
                    const std::string v = "(i64)(" + a.valueExpr + ")";
                    const std::string w = s(a.wordOffset);
                    switch (a.kind) {
                        case AggKind::SumI64: cg.addLine("gdb_lds_add_u64(_kb + " + w + ", (u64)" + v + ");"); break;
                        case AggKind::SumI128:
                            cg.addBlock("", [&]() {
                                cg.addLine("const i64 _v = " + v + ";");
                                cg.addLine("gdb_lds_add_u64(_kb + " + w + ", (u64)(u32)_v);");
                                cg.addLine("gdb_lds_add_u64(_kb + " + s(a.wordOffset + 1) + ", (u64)(_v >> 32));");
                            });
                            break;
                        case AggKind::Count: cg.addLine("gdb_lds_add_u64(_kb + " + w + ", 1ull);"); break;
                        case AggKind::MinI64: cg.addLine("gdb_lds_min_i64((i64*)(_kb + " + w + "), " + v + ");"); break;
                        case AggKind::MaxI64: cg.addLine("gdb_lds_max_i64((i64*)(_kb + " + w + "), " + v + ");"); break;
                    }
                }
            });
            consume();
        });
        cg.addLine("__syncthreads();");
        cg.addBlock("for (int _s = (int)threadIdx.x; _s < " + s((long long)numBuckets_ * words) +
                        "; _s += (int)blockDim.x)",
                    [&]() {
                        cg.addLine("const int _b = _s / " + W + ", _w = _s % " + W + ";");
                        cg.addLine("const u64 _x = _kl_" + out_ + "[_s];");
                        cg.addLine("u64* _dst = " + out_ + " + _b * " + s(wpb) + " + _w;");
                        std::string minCond = "false", maxCond = "false";
                        for (const auto& a : aggs_) {
                            if (a.kind == AggKind::MinI64) minCond += " || _w == " + s(a.wordOffset);
                            if (a.kind == AggKind::MaxI64) maxCond += " || _w == " + s(a.wordOffset);
                        }
                        cg.addLine("if (" + minCond + ") { if ((i64)_x != GDB_I64_MAX) gdb_atomic_min_i64((i64*)_dst, (i64)_x); }");
                        cg.addLine("else if (" + maxCond + ") { if ((i64)_x != GDB_I64_MIN) gdb_atomic_max_i64((i64*)_dst, (i64)_x); }");
                        cg.addLine("else if (_x != 0) gdb_atomic_add_u64(_dst, _x);");
                    });
    } else {
        cg.addComment("KeyedAgg " + out_ + ": per-row global atomics, " + NB + " buckets");
        child_->produce(cg, [&]() {
            cg.addLine("const int _bk = (int)(" + bucket_ + ");");
            cg.addBlock("if ((unsigned)_bk < " + NB + "u)", [&]() {
                cg.addLine("u64* _kb = " + out_ + " + (u64)_bk * " + s(wpb) + ";");
                for (const auto& a : aggs_) {
                    const std::string v = "(i64)(" + a.valueExpr + ")";
                    const std::string w = s(a.wordOffset);
                    switch (a.kind) {
                        case AggKind::SumI64: cg.addLine("gdb_atomic_add_u64(_kb + " + w + ", (u64)" + v + ");"); break;
                        case AggKind::SumI128:
                            cg.addBlock("", [&]() {
                                cg.addLine("const i64 _v = " + v + ";");
                                cg.addLine("gdb_atomic_add_u64(_kb + " + w + ", (u64)(u32)_v);");
                                cg.addLine("gdb_atomic_add_u64(_kb + " + s(a.wordOffset + 1) + ", (u64)(_v >> 32));");
                            });
                            break;
                        case AggKind::Count: cg.addLine("gdb_atomic_add_u64(_kb + " + w + ", 1ull);"); break;
                        case AggKind::MinI64: cg.addLine("gdb_atomic_min_i64((i64*)(_kb + " + w + "), " + v + ");"); break;
                        case AggKind::MaxI64: cg.addLine("gdb_atomic_max_i64((i64*)(_kb + " + w + "), " + v + ");"); break;
                    }
                }
            });
            consume();
        });
    }

    auto& rs = cg.resultSchema();
    rs.kind = HipResultSchema::KEYED_AGG;
    rs.keyed.bufferName = out_;
    rs.keyed.numBuckets = numBuckets_;
    rs.keyed.wordsPerBucket = wpb;
    rs.keyed.slots.clear();
    for (const auto& a : aggs_) rs.keyed.slots.push_back({a.name, a.kind, a.wordOffset, a.scaleDigits});
    rs.keyed.bucketLabels = labels_;
    rs.keyed.keyColumn = keyColumn_;
    rs.keyed.averages = avgs_;
}

std::string HipKeyedAgg::describe() const {
    return "KeyedAgg(" + out_ + ", " + s(numBuckets_) + " buckets, " + s((long long)aggs_.size()) + " aggs, bucket=" +
           bucket_ + ")";
}

// ===========================================================================
// HipAtomicCount
// ===========================================================================

HipAtomicCount::HipAtomicCount(std::unique_ptr<HipOperator> child, const std::string& arrayName,
                               const std::string& bucketExpr, const std::string& sizeExpr, int staticSize)
    : HipUnaryOperator(std::move(child)), array_(arrayName), bucket_(bucketExpr), size_(sizeExpr), static_(staticSize) {}

void HipAtomicCount::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    cg.addBufferParam(array_, "u32", 4, size_, InitKind::Zero, {}, false, true);
    cg.addSymbolParam("nelem_" + array_, size_);
    if (static_ > 0 && static_ <= 8192) {
        const std::string S = s(static_);
        cg.addComment("AtomicCount " + array_ + ": LDS histogram of " + S + " bins");
        cg.declareLds("u32", "_hist_" + array_, S);
        cg.addLine("for (int _h = (int)threadIdx.x; _h < " + S + "; _h += (int)blockDim.x) _hist_" + array_ + "[_h] = 0;");
        cg.addLine("__syncthreads();");
        child_->produce(cg, [&]() {
            cg.addLine("const u64 _hb = (u64)(i64)(" + bucket_ + ");");
            cg.addLine("if (_hb < " + S + "ull && _hb < nelem_" + array_ + ") gdb_lds_add_u32(&_hist_" + array_ +
                       "[_hb], 1u);");
            consume();
        });
        cg.addLine("__syncthreads();");
        cg.addLine("for (int _h = (int)threadIdx.x; _h < " + S + "; _h += (int)blockDim.x) if ((u64)_h < nelem_" +
                   array_ + " && _hist_" + array_ + "[_h]) gdb_atomic_add_u32(" + array_ + " + _h, _hist_" + array_ +
                   "[_h]);");
    } else {
        cg.addComment("AtomicCount " + array_ + ": per-row global atomics");
        child_->produce(cg, [&]() {
            cg.addLine("const u64 _hb = (u64)(i64)(" + bucket_ + ");");
            cg.addLine("if (_hb < nelem_" + array_ + ") gdb_atomic_add_u32(" + array_ + " + _hb, 1u);");
            consume();
        });
    }
    if (isResult_) {
        auto& rs = cg.resultSchema();
        rs.kind = HipResultSchema::COUNTS;
        rs.counts.bufferName = array_;
        rs.counts.sizeExpr = size_;
        rs.counts.keyName = keyName_;
        rs.counts.valueName = valueName_;
        rs.counts.sortByCountDesc = sortDesc_;
    }
}

void HipAtomicCount::setResult(const std::string& keyName, const std::string& valueName, bool sortByCountDesc) {
    isResult_ = true;
    keyName_ = keyName;
    valueName_ = valueName;
    sortDesc_ = sortByCountDesc;
}

std::string HipAtomicCount::describe() const { return "AtomicCount(" + array_ + "[" + bucket_ + "])"; }

// ===========================================================================
// HipHashGroupBy
// ===========================================================================

HipHashGroupBy::HipHashGroupBy(std::unique_ptr<HipOperator> child, const std::string& table,
                               const std::string& keyExpr, const std::string& valueExpr, const std::string& slotsExpr,
                               Variant variant, int ldsSlots)
    : HipUnaryOperator(std::move(child)),
      table_(table),
      key_(keyExpr),
      value_(valueExpr),
      slots_(slotsExpr),
      variant_(variant),
      ldsSlots_(ldsSlots) {
    if (variant_ == LDS && (ldsSlots_ < 64 || (ldsSlots_ & (ldsSlots_ - 1)) != 0 || ldsSlots_ > 4096))
        throw std::runtime_error("HipHashGroupBy: LDS slots must be a power of two in [64, 4096]");
}

namespace {

// Slots in use of a group-by LDS table: the compiled capacity, or 1 << scalar
// (clamped to it) when the slot count comes from a run-time scalar. Returns
// the (slots, bits) expressions.
std::pair<std::string, std::string> ldsTableSize(HipCodegen& cg, const std::string& T, int cap,
                                                 const std::string& bitsScalar) {
    int lbits = 0;
    while ((1 << lbits) < cap) ++lbits;
    if (bitsScalar.empty()) return {s(cap), s(lbits) + "u"};
    cg.addScalarParam(bitsScalar);
    cg.addLine("const u32 _lb_" + T + " = (u32)(" + bitsScalar + " < 1 ? 1 : (" + bitsScalar + " > " + s(lbits) + " ? " +
               s(lbits) + " : " + bitsScalar + "));");
    cg.addLine("const int _ls_" + T + " = 1 << _lb_" + T + ";");
    return {"_ls_" + T, "_lb_" + T};
}

// Word-major LDS group table (word j of slot s at j * LS + s): empty slots.
void emitLdsGbInit(HipCodegen& cg, const std::string& tab, const std::string& LS) {
    cg.addBlock("for (int _s = (int)threadIdx.x; _s < " + LS + "; _s += (int)blockDim.x)", [&]() {
        cg.addLine(tab + "[_s] = GDB_EMPTY_WORD; " + tab + "[" + LS + " + _s] = 0;");
        cg.addLine(tab + "[2 * " + LS + " + _s] = (u64)GDB_I64_MAX; " + tab + "[3 * " + LS + " + _s] = (u64)GDB_I64_MIN;");
    });
}

// Merge every occupied slot of an LDS group table into the global table T.
void emitLdsGbFlush(HipCodegen& cg, const std::string& T, const std::string& tab, const std::string& LS) {
    cg.addBlock("for (int _s = (int)threadIdx.x; _s < " + LS + "; _s += (int)blockDim.x)", [&]() {
        cg.addLine("const u64 _w0 = " + tab + "[_s];");
        cg.addBlock("if ((u32)_w0 != GDB_EMPTY_KEY)", [&]() {
            cg.addLine("const int _r = gdb_gb_merge(" + T + ", _bits_" + T + ", (u32)_w0, (u32)(_w0 >> 32), (i64)" + tab +
                       "[" + LS + " + _s], (i64)" + tab + "[2 * " + LS + " + _s], (i64)" + tab + "[3 * " + LS +
                       " + _s], _maxp_" + T + ");");
            cg.addLine("if (_r < 0) ++_ovf_" + T + "; else _ins_" + T + " += (u32)_r;");
        });
    });
}

// The global group table T, its slot count symbol and status words, and the
// per-thread locals every group-by phase on T uses.
std::string bindGroupTable(HipCodegen& cg, const std::string& T, const std::string& slots) {
    cg.addBufferParam(T, "u64", 8, slots + " * 4", InitKind::Pattern, wordsPattern({kEmptyWord, 0, kI64Max, kI64Min}),
                      false, true);
    cg.addSymbolParam("nslots_" + T, slots);
    const std::string st = T + "_status";
    cg.addBufferParam(st, "u64", 8, "4", InitKind::Zero, {}, false, true);
    cg.setPhaseStatusBuffer(st);
    cg.addLine("const u32 _bits_" + T + " = (u32)__builtin_ctzll(nslots_" + T + ");");
    cg.addLine("const u32 _maxp_" + T + " = (u32)(nslots_" + T + " < 65536ull ? nslots_" + T + " : 65536ull);");
    cg.addLine("u32 _ins_" + T + " = 0, _ovf_" + T + " = 0, _spill_" + T + " = 0;");
    return st;
}

void setHashAggSchema(HipCodegen& cg, const std::string& T, const std::string& slots) {
    auto& rs = cg.resultSchema();
    rs.kind = HipResultSchema::HASH_AGG;
    rs.hash.tableName = T;
    rs.hash.slotsSymbol = slots;
    rs.hash.slotWords = 4;
    rs.hash.slots = {{"count", AggKind::Count, 0, 0},
                     {"sum", AggKind::SumI64, 1, 0},
                     {"min", AggKind::MinI64, 2, 0},
                     {"max", AggKind::MaxI64, 3, 0}};
}

}  // namespace

void HipHashGroupBy::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    const std::string T = table_;
    const std::string st = bindGroupTable(cg, T, slots_);
    const std::string merge1 = "gdb_gb_merge(" + T + ", _bits_" + T + ", _gk, 1u, _gv, _gv, _gv, _maxp_" + T + ")";
    if (variant_ == LDS) {
        const std::string CAP = s(ldsSlots_), tab = "_gl_" + T;
        cg.addComment("HashGroupBy " + T + ": LDS table of " + CAP + " slots absorbs hot keys; spills go global");
        cg.declareLds("u64", tab, CAP + " * 4");
        std::string LS, LB;
        std::tie(LS, LB) = ldsTableSize(cg, T, ldsSlots_, ldsBitsScalar_);
        emitLdsGbInit(cg, tab, LS);
        cg.addLine("__syncthreads();");
        child_->produce(cg, [&]() {
            cg.addLine("const u32 _gk = (u32)(" + key_ + ");");
            cg.addLine("const i64 _gv = (i64)(" + value_ + ");");
            cg.addBlock("if (gdb_lds_gb_update(" + tab + ", " + LB + ", _gk, _gv, " + s(window_) + "u) < 0)", [&]() {
                cg.addLine("const int _r = " + merge1 + ";");
                cg.addLine("if (_r < 0) ++_ovf_" + T + "; else _ins_" + T + " += (u32)_r;");
                cg.addLine("++_spill_" + T + ";");
            });
            consume();
        });
        cg.addLine("__syncthreads();");
        if (partialsMax_.empty()) {
            emitLdsGbFlush(cg, T, tab, LS);
        } else {
            const std::string PT = T + "_pt", PN = T + "_ptn", NP = "nparts_" + T;
            cg.addBufferParam(PT, "u64", 8, "(" + partialsMax_ + ") * " + CAP + " * 4", InitKind::None, {}, false, true);
            cg.addBufferParam(PN, "u64", 8, "1", InitKind::Zero, {}, false, true);
            cg.addSymbolParam(NP, partialsMax_);
            cg.addComment("the LDS table becomes partial table _pi (word-major, " + CAP + " slots apart), if one is left");
            cg.declareLds("u64", "_pti_" + T, "1");
            cg.addLine("if (threadIdx.x == 0) _pti_" + T + "[0] = __hip_atomic_fetch_add(" + PN +
                       ", 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);");
            cg.addLine("__syncthreads();");
            cg.addLine("const u64 _pi_" + T + " = _pti_" + T + "[0];");
            cg.addBlock("if (_pi_" + T + " < " + NP + ")", [&]() {
                cg.addComment("key words of every slot; the other words of occupied slots only");
                cg.addLine("u64* const _pd = " + PT + " + _pi_" + T + " * " + CAP + "ull * 4ull;");
                cg.addBlock("for (int _s = (int)threadIdx.x; _s < " + LS + "; _s += (int)blockDim.x)", [&]() {
                    cg.addLine("const u64 _w0 = " + tab + "[_s];");
                    cg.addLine("_pd[_s] = _w0;");
                    cg.addBlock("if ((u32)_w0 != GDB_EMPTY_KEY)", [&]() {
                        cg.addLine("_pd[" + LS + " + _s] = " + tab + "[" + LS + " + _s]; _pd[2 * " + LS + " + _s] = " + tab +
                                   "[2 * " + LS + " + _s]; _pd[3 * " + LS + " + _s] = " + tab + "[3 * " + LS + " + _s];");
                    });
                });
            });
            cg.addBlock("else", [&]() { emitLdsGbFlush(cg, T, tab, LS); });
        }
    } else {
        cg.addComment("HashGroupBy " + T + ": every row merges into the global table");
        child_->produce(cg, [&]() {
            cg.addLine("const u32 _gk = (u32)(" + key_ + ");");
            cg.addLine("const i64 _gv = (i64)(" + value_ + ");");
            cg.addLine("const int _r = " + merge1 + ";");
            cg.addLine("if (_r < 0) ++_ovf_" + T + "; else _ins_" + T + " += (u32)_r;");
            consume();
        });
    }
    cg.addComment("group-by status: [0] overflow flag, [2] keys inserted, [3] rows spilled from LDS");
    emitBlockStatus(cg, st, {{"_ins_" + T, 2, false}, {"_ovf_" + T, 0, true}, {"_spill_" + T, 3, false}});
    setHashAggSchema(cg, T, slots_);
}

void HipGroupByMerge::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    const std::string T = table_, CAP = s(ldsSlots_), tab = "_gm_" + T;
    if (ldsSlots_ < 64 || (ldsSlots_ & (ldsSlots_ - 1)))
        throw std::runtime_error("HipGroupByMerge: LDS slots must be a power of two >= 64");
    const std::string st = bindGroupTable(cg, T, slots_);
    const std::string PT = T + "_pt", PN = T + "_ptn", NP = "nparts_" + T;
    cg.addBufferParam(PT, "u64", 8, "(" + partialsMax_ + ") * " + CAP + " * 4", InitKind::None, {}, true, true);
    cg.addBufferParam(PN, "u64", 8, "1", InitKind::Zero, {}, true, true);
    cg.addSymbolParam(NP, partialsMax_);
    cg.setPhaseRowsExpr("(" + partialsMax_ + ") * " + CAP);
    cg.addComment("GroupByMerge " + T + ": partial LDS tables -> one LDS table per block -> the global table");
    cg.declareLds("u64", tab, CAP + " * 4");
    std::string LS, LB;
    std::tie(LS, LB) = ldsTableSize(cg, T, ldsSlots_, ldsBitsScalar_);
    emitLdsGbInit(cg, tab, LS);
    cg.addLine("__syncthreads();");
    cg.addLine("const u64 _np = " + PN + "[0] < " + NP + " ? " + PN + "[0] : " + NP + ";");
    cg.addLine("const u64 _flat = _np * " + CAP + "ull;  // partial slots written");
    auto body = [&]() {
        cg.addLine("const u64 _s = _r & " + s(ldsSlots_ - 1) + "ull;");
        cg.addBlock("if (_s < (u64)" + LS + ")", [&]() {
            cg.addLine("const u64* const _p = " + PT + " + (_r - _s) * 4ull + _s;");
            cg.addLine("const u64 _w0 = _p[0];");
            cg.addBlock("if ((u32)_w0 != GDB_EMPTY_KEY)", [&]() {
                cg.addLine("const u32 _k = (u32)_w0, _c = (u32)(_w0 >> 32);");
                cg.addLine("const i64 _sm = (i64)_p[" + LS + "], _mn = (i64)_p[2 * " + LS + "], _mx = (i64)_p[3 * " + LS +
                           "];");
                // a bounded probe window: with more groups than LDS slots a
                // full-table probe per missing key would cost thousands of
                // LDS loads; such keys merge into the global table instead
                cg.addBlock("if (gdb_lds_gb_merge_at(" + tab + ", (u32)" + LS + " - 1u, (u32)gdb_hash_slot(_k, " + LB +
                                "), _k, _c, _sm, _mn, _mx, " + s(kMergeWindow) + "u) < 0)",
                            [&]() {
                                cg.addLine("const int _g = gdb_gb_merge(" + T + ", _bits_" + T +
                                           ", _k, _c, _sm, _mn, _mx, _maxp_" + T + ");");
                                cg.addLine("if (_g < 0) ++_ovf_" + T + "; else _ins_" + T + " += (u32)_g;");
                            });
            });
        });
    };
    if (cg.rangeMode()) {
        cg.addComment("work item: partial slots [_rb, _re)");
        cg.addBlock("for (u64 _r = _rb + threadIdx.x; _r < (_re < _flat ? _re : _flat); _r += blockDim.x)", body);
    } else {
        cg.addComment("block: contiguous tiles of partial slots, tile-strided over the grid");
        cg.addLine("const u64 _tile = (u64)blockDim.x * ITEMS;");
        cg.addBlock("for (u64 _t0 = (u64)blockIdx.x * _tile; _t0 < _flat; _t0 += (u64)gridDim.x * _tile)", [&]() {
            cg.addLine("const u64 _te = _t0 + _tile < _flat ? _t0 + _tile : _flat;");
            cg.addBlock("for (u64 _r = _t0 + threadIdx.x; _r < _te; _r += blockDim.x)", body);
        });
    }
    cg.addLine("__syncthreads();");
    emitLdsGbFlush(cg, T, tab, LS);
    cg.addComment("group-by status: [0] overflow flag, [2] keys inserted");
    emitBlockStatus(cg, st, {{"_ins_" + T, 2, false}, {"_ovf_" + T, 0, true}});
    consume();
    setHashAggSchema(cg, T, slots_);
}

std::string HipGroupByMerge::describe() const {
    return "GroupByMerge(" + table_ + ", partials<=" + partialsMax_ + ", lds" + s(ldsSlots_) + ")";
}

std::string HipHashGroupBy::describe() const {
    return std::string("HashGroupBy(") + table_ + ", key=" + key_ + ", value=" + value_ + ", slots=" + slots_ +
           ", variant=" + (variant_ == LDS ? "lds" + s(ldsSlots_) : "global") + ")";
}

// ===========================================================================
// Partitioned hash group-by (five phases, no global atomics)
// ===========================================================================

int PartitionConfig::maxParts() const {
    // LDS: pass-1 table (32 B per slot) + one u32 counter per partition, within
    // 150 KiB of the CU's 160 KiB; a power of two, at most 16384.
    const long long room = (150LL * 1024 - 32LL * ldsSlots - 256) / 4;
    int p = 1;
    while (2LL * p <= room && p < 16384) p *= 2;
    return p;
}

namespace {

struct PartNames {
    std::string P, NP, PC, NPC, rows, hot, ccnt, pcnt, coff, poff, meta, rec, part;
};

// Which partition buffers a phase binds: 0 = none, 'r' = read, 'w' = written.
struct PartUse {
    char hot = 0, cnt = 0, off = 0, meta = 0, stream = 0;
    bool np = false;   // binds the table's partition count (needs the table size)
    bool npc = false;  // binds the first pass's count-partition count
};

// Bind the partition buffers and symbols a phase uses. Sizes are expressions
// of the group-table slots, the count partitions and the scanned row count,
// the same in every phase; the first pass binds nothing that depends on the
// table size, so a pre-pass fused into it can size the table afterwards.
PartNames bindPartition(HipCodegen& cg, const PartitionConfig& c, const PartUse& u) {
    PartNames n;
    const std::string px = c.prefix;
    const std::string LB = s(PartitionConfig::kLogicalBlocks), CAP = s(c.ldsSlots);
    n.P = "(" + c.slots + ") / " + s(PartitionConfig::kPartSlots);
    n.PC = c.countParts.empty() ? n.P : c.countParts;
    n.NP = "nparts_" + px;
    n.NPC = "ncparts_" + px;
    n.rows = "nrows_" + px;
    n.hot = px + "_hot";
    n.ccnt = px + "_ccnt";
    n.pcnt = px + "_pcnt";
    n.coff = px + "_coff";
    n.poff = px + "_poff";
    n.meta = px + "_meta";
    n.rec = px + "_rec";
    n.part = px + "_part";
    auto buf = [&](char use, const std::string& name, const char* type, size_t bytes, const std::string& size) {
        if (use) cg.addBufferParam(name, type, bytes, size, InitKind::None, {}, use == 'r', true);
    };
    if (u.np) cg.addSymbolParam(n.NP, n.P);
    if (u.npc) cg.addSymbolParam(n.NPC, n.PC);
    cg.addSymbolParam(n.rows, c.rowsSym);
    buf(u.hot, n.hot, "u64", 8, LB + " * " + CAP + " * 4");
    // cold-row and partial counts per (logical block, count partition); their
    // exclusive offsets per (logical block, table partition)
    buf(u.cnt, n.ccnt, "u32", 4, LB + " * (" + n.PC + ")");
    buf(u.cnt, n.pcnt, "u32", 4, LB + " * (" + n.PC + ")");
    buf(u.off, n.coff, "u32", 4, LB + " * (" + n.P + ")");
    buf(u.off, n.poff, "u32", 4, LB + " * (" + n.P + ")");
    buf(u.meta, n.meta, "u32", 4, "4 * (" + n.P + ")");
    // cold rows as 16-byte records (key, value): one store per row
    buf(u.stream, n.rec, "u64", 8, "2 * (" + c.rowsSym + " + 1)");
    buf(u.stream, n.part, "u64", 8, LB + " * " + CAP + " * 4");
    return n;
}

// The logical-block loop shared by pass 1 and pass 4: logical block _lb
// covers rows [_rb, _re) of the scanned table, a whole number of tiles.
void openLogicalBlocks(HipCodegen& cg, const std::string& nRows, const std::string& cond) {
    cg.addLine("const u64 _ltile = (u64)blockDim.x * ITEMS;");
    cg.addLine("const u64 _chunk = ((" + nRows + " + " + s(PartitionConfig::kLogicalBlocks - 1) + ") / " +
               s(PartitionConfig::kLogicalBlocks) + " + _ltile - 1) / _ltile * _ltile;");
    cg.addLine("for (u64 _lb = blockIdx.x; " + cond + " && _lb < " + s(PartitionConfig::kLogicalBlocks) +
               "; _lb += gridDim.x) {");
    cg.increaseIndent();
    cg.addLine("const u64 _rb = _lb * _chunk;");
    cg.addLine("const u64 _re = _rb + _chunk < " + nRows + " ? _rb + _chunk : " + nRows + ";");
}

void closeLogicalBlocks(HipCodegen& cg) {
    cg.decreaseIndent();
    cg.addLine("}");
}

}  // namespace

HipPartitionCount::HipPartitionCount(std::unique_ptr<HipOperator> child, PartitionConfig cfg,
                                     const std::string& keyExpr, const std::string& valueExpr)
    : HipUnaryOperator(std::move(child)), c_(std::move(cfg)), key_(keyExpr), value_(valueExpr) {
    if (c_.ldsSlots < 64 || (c_.ldsSlots & (c_.ldsSlots - 1)) || c_.ldsSlots > 4096)
        throw std::runtime_error("HipPartitionCount: LDS slots must be a power of two in [64, 4096]");
}

void HipPartitionCount::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    PartUse u;
    u.hot = 'w';
    u.cnt = 'w';
    u.npc = true;
    const PartNames n = bindPartition(cg, c_, u);
    const std::string st = c_.table + "_status";
    cg.addBufferParam(st, "u64", 8, "4", InitKind::Zero, {}, false, true);
    cg.setPhaseStatusBuffer(st);
    const std::string nRows = cg.addTableSizeParam(c_.scanTable);
    const std::string CAP = s(c_.ldsSlots), PMAX = s(c_.maxParts());
    const bool flags = !c_.flags.empty();
    std::string F, NF;
    if (flags) {
        // the exact distinct-count pre-pass, fused: one byte per key of the range
        F = c_.flags;
        NF = "nflags_" + F;
        cg.addBufferParam(F, "u8", 1, c_.flagBits, InitKind::Zero);
        cg.addSymbolParam(NF, c_.flagBits);
        cg.addScalarParam(c_.flagMin);
    }
    // flag of u32 key k (a key of the signed 32-bit column): plain read, then
    // a plain store (every writer stores the same byte)
    auto setFlag = [&](const std::string& k) {
        cg.addLine("{ const u64 _fk = (u64)((i64)(i32)(" + k + ") - " + c_.flagMin + "); if (_fk < " + NF + " && " + F +
                   "[_fk] == 0) " + F + "[_fk] = 1; }");
    };
    int lbits = 0;
    while ((1 << lbits) < c_.ldsSlots) ++lbits;
    cg.addComment("pass 1: LDS table of " + CAP + " slots per logical block; cold rows counted per partition" +
                  (flags ? "; key flags set (cold rows' keys here, the table's keys at the end)" : ""));
    cg.declareLds("u64", "_gl", CAP + " * 4");
    cg.declareLds("u32", "_pc", PMAX);
    cg.addLine("const u64 _P = " + n.NPC + ";  // count partitions");
    cg.addLine("const u32 _pb = (u32)__builtin_ctzll(_P);");
    cg.addLine("const bool _fits = _P <= " + PMAX + "ull;");
    cg.addLine("u32 _ovf = (!_fits && blockIdx.x == 0 && threadIdx.x == 0) ? 1u : 0u, _spill = 0;");
    openLogicalBlocks(cg, nRows, "_fits");
    cg.addBlock("for (int _s = (int)threadIdx.x; _s < " + CAP + "; _s += (int)blockDim.x)", [&]() {  // word-major
        cg.addLine("_gl[_s] = GDB_EMPTY_WORD; _gl[" + CAP + " + _s] = 0;");
        cg.addLine("_gl[2 * " + CAP + " + _s] = (u64)GDB_I64_MAX; _gl[3 * " + CAP + " + _s] = (u64)GDB_I64_MIN;");
    });
    cg.addLine("for (u64 _q = threadIdx.x; _q < _P; _q += blockDim.x) _pc[_q] = 0;");
    cg.addLine("__syncthreads();");
    cg.setScanSegment(true);
    child_->produce(cg, [&]() {
        cg.addLine("const u32 _gk = (u32)(" + key_ + ");");
        cg.addLine("const i64 _gv = (i64)(" + value_ + ");");
        cg.addBlock("if (gdb_lds_gb_update(_gl, " + s(lbits) + "u, _gk, _gv, " + s(c_.window) + "u) < 0)", [&]() {
            cg.addLine("gdb_lds_add_u32(&_pc[gdb_partition(_gk, _pb)], 1u);");
            cg.addLine("++_spill;");
            if (flags) setFlag("_gk");
        });
        consume();
    });
    cg.setScanSegment(false);
    cg.addLine("__syncthreads();");
    cg.addComment("the block's final LDS table, and its cold-row counts per partition");
    cg.addLine("u64* _h = " + n.hot + " + _lb * " + CAP + " * 4;");
    cg.addLine("for (int _s = (int)threadIdx.x; _s < " + CAP + " * 4; _s += (int)blockDim.x) _h[_s] = _gl[(_s & 3) * " +
               CAP + " + (_s >> 2)];  // slot-major in HBM");
    cg.addLine("for (u64 _q = threadIdx.x; _q < _P; _q += blockDim.x) { " + n.ccnt + "[_lb * _P + _q] = _pc[_q]; _pc[_q] = 0; }");
    cg.addLine("__syncthreads();");
    cg.addComment("partial-aggregate counts per partition");
    cg.addBlock("for (int _s = (int)threadIdx.x; _s < " + CAP + "; _s += (int)blockDim.x)", [&]() {
        cg.addLine("const u64 _w0 = _gl[_s];");
        cg.addBlock("if ((u32)_w0 != GDB_EMPTY_KEY)", [&]() {
            cg.addLine("gdb_lds_add_u32(&_pc[gdb_partition((u32)_w0, _pb)], 1u);");
            if (flags) setFlag("(u32)_w0");
        });
    });
    cg.addLine("__syncthreads();");
    cg.addLine("for (u64 _q = threadIdx.x; _q < _P; _q += blockDim.x) " + n.pcnt + "[_lb * _P + _q] = _pc[_q];");
    cg.addLine("__syncthreads();");
    closeLogicalBlocks(cg);
    cg.addComment("status: [0]+[1] partition count beyond the LDS counters, [3] cold rows");
    emitBlockStatus(cg, st, {{"_ovf", 0, true}, {"_ovf", 1, true}, {"_spill", 3, false}});
}

std::string HipPartitionCount::describe() const {
    return "PartitionCount(" + c_.prefix + ", key=" + key_ + ", value=" + value_ + ", slots=" + c_.slots + ", lds" +
           s(c_.ldsSlots) + (c_.countParts.empty() ? "" : ", count parts " + c_.countParts) +
           (c_.flags.empty() ? "" : ", flags " + c_.flags) + ")";
}

void HipPartitionOffsets::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    PartUse u;
    u.cnt = 'r';
    u.off = 'w';
    u.meta = 'w';
    u.np = true;
    u.npc = true;
    const PartNames n = bindPartition(cg, c_, u);
    const std::string st = c_.table + "_status";
    cg.addBufferParam(st, "u64", 8, "4", InitKind::Zero, {}, false, true);
    cg.setPhaseStatusBuffer(st);
    const std::string LB = s(PartitionConfig::kLogicalBlocks);
    cg.setPhaseRowsExpr("(" + n.P + ") * 64");
    cg.addComment("counts -> exclusive offsets over the logical blocks, per partition, and totals; a partition's "
                  "count is the sum of its consecutive count partitions");
    cg.addLine("const u64 _P = " + n.NP + ", _PC = " + n.NPC + ";");
    cg.addLine("const bool _ok = _P <= _PC && _P <= " + s(c_.maxParts()) + "ull;");
    cg.addLine("const u64 _f = _ok ? _PC / _P : 1;");
    cg.addLine("const u64 _gt = (u64)blockIdx.x * blockDim.x + threadIdx.x, _nt = (u64)gridDim.x * blockDim.x;");
    cg.addBlock("if (!_ok)", [&]() {
        cg.addComment("more table partitions than count partitions or LDS counters: [0]+[1], and zero offsets "
                      "and totals, so the later passes of this (failed) run stay within their buffers");
        cg.addLine("if (_gt == 0) { gdb_atomic_max_u64(" + st + " + 0, 1ull); gdb_atomic_max_u64(" + st + " + 1, 1ull); }");
        cg.addLine("for (u64 _i = _gt; _i < " + LB + " * _P; _i += _nt) { " + n.coff + "[_i] = 0; " + n.poff + "[_i] = 0; }");
        cg.addLine("for (u64 _i = _gt; _i < 2 * _P; _i += _nt) " + n.meta + "[_i] = 0;");
    });
    cg.addBlock("if (_f == 1)", [&]() {
        cg.addComment("one thread per partition: consecutive lanes read consecutive counters");
        cg.addBlock("for (u64 _q = _gt; _ok && _q < _P; _q += _nt)", [&]() {
            cg.addLine("u32 _c = 0, _r = 0;");
            cg.addBlock("for (u64 _lb = 0; _lb < " + LB + "; ++_lb)", [&]() {
                cg.addLine("const u32 _a = " + n.ccnt + "[_lb * _PC + _q], _b = " + n.pcnt + "[_lb * _PC + _q];");
                cg.addLine(n.coff + "[_lb * _P + _q] = _c; " + n.poff + "[_lb * _P + _q] = _r;");
                cg.addLine("_c += _a; _r += _b;");
            });
            cg.addLine(n.meta + "[_q] = _c; " + n.meta + "[_P + _q] = _r;");
        });
    });
    cg.addBlock("else", [&]() {
        cg.addComment("one wave per partition: the lanes sum its count partitions");
        cg.addLine("const u64 _lane = threadIdx.x & 63u;");
        cg.addBlock("for (u64 _q = _gt >> 6; _ok && _q < _P; _q += _nt >> 6)", [&]() {
            cg.addLine("u64 _c = 0, _r = 0;");
            cg.addBlock("for (u64 _lb = 0; _lb < " + LB + "; ++_lb)", [&]() {
                cg.addLine("u64 _a = 0, _b = 0;");
                cg.addLine("for (u64 _j = _lane; _j < _f; _j += 64) { _a += " + n.ccnt + "[_lb * _PC + _q * _f + _j]; _b += " +
                           n.pcnt + "[_lb * _PC + _q * _f + _j]; }");
                cg.addLine("_a = gdb_wave_sum_u64(_a); _b = gdb_wave_sum_u64(_b);");
                cg.addLine("if (_lane == 0) { " + n.coff + "[_lb * _P + _q] = (u32)_c; " + n.poff + "[_lb * _P + _q] = (u32)_r; }");
                cg.addLine("_c += _a; _r += _b;");
            });
            cg.addLine("if (_lane == 0) { " + n.meta + "[_q] = (u32)_c; " + n.meta + "[_P + _q] = (u32)_r; }");
        });
    });
    consume();
}

void HipPartitionBases::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    PartUse u;
    u.meta = 'w';
    u.np = true;
    const PartNames n = bindPartition(cg, c_, u);
    cg.setPhaseRowsExpr("1");
    cg.setPhaseMaxThreadgroups(1);
    cg.addComment("one block: exclusive scans of the partition totals -> partition bases");
    cg.declareLds("u32", "_sa", "1024");
    cg.declareLds("u32", "_sb", "1024");
    cg.addBlock("if (blockIdx.x == 0)", [&]() {
        cg.addLine("const u64 _P = " + n.NP + ";");
        cg.addLine("const u64 _seg = (_P + blockDim.x - 1) / blockDim.x;");
        cg.addLine("const u64 _q0 = threadIdx.x * _seg, _q1 = _q0 + _seg < _P ? _q0 + _seg : _P;");
        cg.addLine("u32 _a = 0, _b = 0;");
        cg.addLine("for (u64 _q = _q0; _q < _q1; ++_q) { _a += " + n.meta + "[_q]; _b += " + n.meta + "[_P + _q]; }");
        cg.addLine("_sa[threadIdx.x] = _a; _sb[threadIdx.x] = _b;");
        cg.addLine("__syncthreads();");
        cg.addComment("inclusive Hillis-Steele scan of the per-thread sums");
        cg.addBlock("for (u32 _d = 1; _d < blockDim.x; _d <<= 1)", [&]() {
            cg.addLine("const u32 _xa = threadIdx.x >= _d ? _sa[threadIdx.x - _d] : 0u;");
            cg.addLine("const u32 _xb = threadIdx.x >= _d ? _sb[threadIdx.x - _d] : 0u;");
            cg.addLine("__syncthreads();");
            cg.addLine("_sa[threadIdx.x] += _xa; _sb[threadIdx.x] += _xb;");
            cg.addLine("__syncthreads();");
        });
        cg.addLine("u32 _ra = _sa[threadIdx.x] - _a, _rb2 = _sb[threadIdx.x] - _b;");
        cg.addBlock("for (u64 _q = _q0; _q < _q1; ++_q)", [&]() {
            cg.addLine("const u32 _ta = " + n.meta + "[_q], _tb = " + n.meta + "[_P + _q];");
            cg.addLine(n.meta + "[2 * _P + _q] = _ra; " + n.meta + "[3 * _P + _q] = _rb2;");
            cg.addLine("_ra += _ta; _rb2 += _tb;");
        });
    });
    consume();
}

HipPartitionScatter::HipPartitionScatter(std::unique_ptr<HipOperator> child, PartitionConfig cfg,
                                         const std::string& keyExpr, const std::string& valueExpr)
    : HipUnaryOperator(std::move(child)), c_(std::move(cfg)), key_(keyExpr), value_(valueExpr) {}

void HipPartitionScatter::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    PartUse u;
    u.hot = 'r';
    u.off = 'r';
    u.meta = 'r';
    u.stream = 'w';
    u.np = true;
    const PartNames n = bindPartition(cg, c_, u);
    const std::string nRows = cg.addTableSizeParam(c_.scanTable);
    const std::string CAP = s(c_.ldsSlots), PMAX = s(c_.maxParts());
    const std::string rst = c_.prefix + "_rec_status";
    if (c_.narrow) {
        cg.addBufferParam(rst, "u64", 8, "4", InitKind::Zero, {}, false, true);
        cg.setPhaseStatusBuffer(rst);
        cg.addLine("u32 _wide = 0;  // values that do not fit an 8-byte record");
    }
    int lbits = 0;
    while ((1 << lbits) < c_.ldsSlots) ++lbits;
    cg.addComment(std::string("pass 4: rows whose key is not in the block's final LDS table are written, packed by "
                              "partition, as ") + (c_.narrow ? "8-byte records (key | value << 32)" : "16-byte records"));
    cg.declareLds("u32", "_hk", CAP);
    cg.declareLds("u32", "_pc", PMAX);
    cg.addLine("const u64 _P = " + n.NP + ";");
    cg.addLine("const u32 _pb = (u32)__builtin_ctzll(_P);");
    openLogicalBlocks(cg, nRows, "_P <= " + PMAX + "ull");
    cg.addLine("const u64* _h = " + n.hot + " + _lb * " + CAP + " * 4;");
    cg.addLine("for (int _s = (int)threadIdx.x; _s < " + CAP + "; _s += (int)blockDim.x) _hk[_s] = (u32)_h[4 * _s];");
    cg.addLine("for (u64 _q = threadIdx.x; _q < _P; _q += blockDim.x) _pc[_q] = " + n.meta + "[2 * _P + _q] + " + n.coff +
               "[_lb * _P + _q];");
    cg.addLine("__syncthreads();");
    cg.setScanSegment(true);
    child_->produce(cg, [&]() {
        cg.addLine("const u32 _gk = (u32)(" + key_ + ");");
        cg.addBlock("if (!gdb_lds_key_member(_hk, " + s(lbits) + "u, _gk, " + s(c_.window) + "u))", [&]() {
            cg.addLine("const u32 _i = __hip_atomic_fetch_add(&_pc[gdb_partition(_gk, _pb)], 1u, __ATOMIC_RELAXED, "
                       "__HIP_MEMORY_SCOPE_WORKGROUP);");
            if (c_.narrow) {
                cg.addLine("const i64 _v = (i64)(" + value_ + ");");
                cg.addLine("if (_v != (i64)(i32)_v) ++_wide;");
                cg.addLine(n.rec + "[_i] = (u64)_gk | ((u64)(u32)(i32)_v << 32);");
            } else {
                cg.addLine("u64x2 _rv; _rv.x = (u64)_gk; _rv.y = (u64)(i64)(" + value_ + ");");
                cg.addLine("*reinterpret_cast<u64x2*>(" + n.rec + " + 2 * (u64)_i) = _rv;");
            }
        });
        consume();
    });
    cg.setScanSegment(false);
    cg.addLine("__syncthreads();");
    cg.addComment("the block's hot partial aggregates, packed by partition");
    cg.addLine("for (u64 _q = threadIdx.x; _q < _P; _q += blockDim.x) _pc[_q] = " + n.meta + "[3 * _P + _q] + " + n.poff +
               "[_lb * _P + _q];");
    cg.addLine("__syncthreads();");
    cg.addBlock("for (int _s = (int)threadIdx.x; _s < " + CAP + "; _s += (int)blockDim.x)", [&]() {
        cg.addLine("const u64 _w0 = _h[4 * _s];");
        cg.addBlock("if ((u32)_w0 != GDB_EMPTY_KEY)", [&]() {
            cg.addLine("const u32 _j = __hip_atomic_fetch_add(&_pc[gdb_partition((u32)_w0, _pb)], 1u, __ATOMIC_RELAXED, "
                       "__HIP_MEMORY_SCOPE_WORKGROUP);");
            cg.addLine("u64* _d = " + n.part + " + (u64)_j * 4;");
            cg.addLine("_d[0] = _w0; _d[1] = _h[4 * _s + 1]; _d[2] = _h[4 * _s + 2]; _d[3] = _h[4 * _s + 3];");
        });
    });
    cg.addLine("__syncthreads();");
    closeLogicalBlocks(cg);
    if (c_.narrow) {
        cg.addComment("status: [0] a value needs a 16-byte record");
        emitBlockStatus(cg, rst, {{"_wide", 0, true}});
    }
}

std::string HipPartitionScatter::describe() const {
    return "PartitionScatter(" + c_.prefix + ", key=" + key_ + ", lds" + s(c_.ldsSlots) +
           (c_.narrow ? ", 8-byte records" : "") + ")";
}

void HipPartitionAgg::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    PartUse u;
    u.meta = 'r';
    u.stream = 'r';
    u.np = true;
    const PartNames n = bindPartition(cg, c_, u);
    const std::string T = c_.table;
    const std::string PS = s(PartitionConfig::kPartSlots), PM = s(PartitionConfig::kPartSlots - 1);
    cg.setPhaseRowsExpr(c_.slots);  // one tile (blockDim x ITEMS = 4096) of the slot range per partition
    cg.addBufferParam(T, "u64", 8, c_.slots + " * 4", InitKind::None, {}, false, true);
    cg.addSymbolParam("nslots_" + T, c_.slots);
    const std::string st = T + "_status";
    cg.addBufferParam(st, "u64", 8, "4", InitKind::Zero, {}, false, true);
    cg.setPhaseStatusBuffer(st);
    cg.addLine("const u64 _P = " + n.NP + ";");
    cg.addLine("const u32 _pb = (u32)__builtin_ctzll(_P);");
    cg.addLine("const u32 _sh = 32u - _pb - 12u;  // home slot: the 12 hash bits after the partition bits");
    cg.addLine("u32 _ins = 0, _ovf = 0;");
    cg.addComment("pass 5: each partition's group-table slice is built in LDS, then stored whole");
    cg.declareLds("u64", "_pl", PS + " * 4");
    cg.addBlock("for (u64 _p = blockIdx.x; _p < _P; _p += gridDim.x)", [&]() {
        cg.addBlock("for (int _s = (int)threadIdx.x; _s < " + PS + "; _s += (int)blockDim.x)", [&]() {
            cg.addLine("_pl[_s] = GDB_EMPTY_WORD; _pl[" + PS + " + _s] = 0;  // word-major");
            cg.addLine("_pl[2 * " + PS + " + _s] = (u64)GDB_I64_MAX; _pl[3 * " + PS + " + _s] = (u64)GDB_I64_MIN;");
        });
        cg.addLine("__syncthreads();");
        cg.addLine("u64 _cb = " + n.meta + "[2 * _P + _p], _cn = " + n.meta + "[_p];");
        cg.addComment("records beyond the buffers can only come from a failed earlier pass: reported, not read");
        cg.addLine("if (_cb + _cn > " + n.rows + ") { _cn = 0; if (threadIdx.x == 0) ++_ovf; }");
        cg.addBlock("for (u64 _i = threadIdx.x; _i < _cn; _i += blockDim.x)", [&]() {
            if (c_.narrow) {
                cg.addLine("const u64 _w = " + n.rec + "[_cb + _i];");
                cg.addLine("const u32 _k = (u32)_w;");
                cg.addLine("const i64 _v = (i64)(i32)(u32)(_w >> 32);");
            } else {
                cg.addLine("const u64x2 _rv = *reinterpret_cast<const u64x2*>(" + n.rec + " + 2 * (_cb + _i));");
                cg.addLine("const u32 _k = (u32)_rv.x;");
                cg.addLine("const i64 _v = (i64)_rv.y;");
            }
            cg.addLine("if (gdb_lds_gb_merge_at(_pl, " + PM + "u, ((_k * GDB_KNUTH) >> _sh) & " + PM +
                       "u, _k, 1u, _v, _v, _v, " + PS + "u) < 0) ++_ovf;");
        });
        cg.addLine("u64 _pbase = " + n.meta + "[3 * _P + _p], _pn = " + n.meta + "[_P + _p];");
        cg.addLine("if (_pbase + _pn > " + s(PartitionConfig::kLogicalBlocks) + "ull * " + s(c_.ldsSlots) +
                   ") { _pn = 0; if (threadIdx.x == 0) ++_ovf; }");
        cg.addBlock("for (u64 _j = threadIdx.x; _j < _pn; _j += blockDim.x)", [&]() {
            cg.addLine("const u64* _d = " + n.part + " + (_pbase + _j) * 4;");
            cg.addLine("const u64 _w0 = _d[0];");
            cg.addLine("const u32 _k = (u32)_w0;");
            cg.addLine("if (gdb_lds_gb_merge_at(_pl, " + PM + "u, ((_k * GDB_KNUTH) >> _sh) & " + PM +
                       "u, _k, (u32)(_w0 >> 32), (i64)_d[1], (i64)_d[2], (i64)_d[3], " + PS + "u) < 0) ++_ovf;");
        });
        cg.addLine("__syncthreads();");
        cg.addBlock("for (int _s = (int)threadIdx.x; _s < " + PS + "; _s += (int)blockDim.x)", [&]() {
            cg.addLine("u64* _g = " + T + " + (_p * " + PS + " + (u64)_s) * 4;");
            cg.addLine("const u64 _w0 = _pl[_s];");
            cg.addLine("_g[0] = _w0; _g[1] = _pl[" + PS + " + _s]; _g[2] = _pl[2 * " + PS + " + _s]; _g[3] = _pl[3 * " + PS +
                       " + _s];");
            cg.addLine("if ((u32)_w0 != GDB_EMPTY_KEY) ++_ins;");
        });
        cg.addLine("__syncthreads();");
    });
    cg.addComment("status: [0] overflow (a partition holds more groups than its slice), [2] groups");
    emitBlockStatus(cg, st, {{"_ovf", 0, true}, {"_ins", 2, false}});
    consume();

    auto& rs = cg.resultSchema();
    rs.kind = HipResultSchema::HASH_AGG;
    rs.hash.tableName = T;
    rs.hash.slotsSymbol = c_.slots;
    rs.hash.slotWords = 4;
    rs.hash.slots = {{"count", AggKind::Count, 0, 0},
                     {"sum", AggKind::SumI64, 1, 0},
                     {"min", AggKind::MinI64, 2, 0},
                     {"max", AggKind::MaxI64, 3, 0}};
}

// ===========================================================================
// HipTopK
// ===========================================================================

HipTopK::HipTopK(std::unique_ptr<HipOperator> child, const std::string& outputName, int k, const std::string& rankHiExpr,
                 const std::string& rankLoExpr)
    : HipUnaryOperator(std::move(child)), out_(outputName), hi_(rankHiExpr), lo_(rankLoExpr), k_(k) {
    if (k_ < 1 || k_ > 32) throw std::runtime_error("HipTopK: k must be in [1, 32]");
}

void HipTopK::addColumn(const std::string& name, int word, int shift, int bits, long long complementOf, int scale,
                        bool date) {
    cols_.push_back({name, word, shift, bits, complementOf, scale, date});
}

void HipTopK::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    const std::string K = s(k_);
    const std::string cand = out_ + "_cand";
    // Each block writes its top K to its own candidate slots; a final
    // one-block phase (HipTopKMerge) merges them.
    cg.setPhaseMaxThreadgroups(kMaxBlocks);
    cg.addBufferParam(cand, "u64", 8, s(2LL * k_ * kMaxBlocks), InitKind::Zero, {}, false, true);
    cg.addLine("GdbTopK<" + K + "> _tk;");
    cg.addLine("gdb_topk_init<" + K + ">(_tk);");
    child_->produce(cg, [&]() {
        cg.addLine("gdb_topk_push<" + K + ">(_tk, (u64)(" + hi_ + "), (u64)(" + lo_ + "));");
        consume();
    });
    cg.addComment("TopK " + out_ + ": block selection into this block's candidate slots");
    cg.addBlock("", [&]() {
        cg.declareLds("u64", "_tk_lds", "GDB_TOPK_LDS_WORDS(" + K + ")");
        cg.addLine("gdb_topk_block_merge<" + K + ">(_tk, _tk_lds);");
        cg.addBlock("if (threadIdx.x < " + s(2LL * k_) + ")", [&]() {
            cg.addLine(cand + "[(u64)blockIdx.x * " + s(2LL * k_) + " + threadIdx.x] = _tk_lds[32 * " + K +
                       " + threadIdx.x];");
        });
    });
}

std::unique_ptr<HipOperator> HipTopK::makeMerge() const { return std::make_unique<HipTopKMerge>(out_, k_, cols_); }

HipTopKMerge::HipTopKMerge(const std::string& outputName, int k, std::vector<HipTopK::Column> cols)
    : out_(outputName), k_(k), cols_(std::move(cols)) {}

void HipTopKMerge::produceImpl(HipCodegen& cg, ConsumerFn consume) {
    const std::string K = s(k_);
    const std::string ctl = out_ + "_ctl";
    const std::string cand = out_ + "_cand";
    const std::string NC = s(static_cast<long long>(k_) * HipTopK::kMaxBlocks);
    cg.setPhaseRowsExpr("1");
    cg.setPhaseMaxThreadgroups(1);
    cg.addBufferParam(out_, "u64", 8, s(2LL * k_), InitKind::Zero, {}, false, true);
    cg.addBufferParam(ctl, "u64", 8, "4", InitKind::Zero, {}, false, true);
    cg.addBufferParam(cand, "u64", 8, s(2LL * k_ * HipTopK::kMaxBlocks), InitKind::None, {}, true, true);
    cg.addComment("TopK " + out_ + ": one block merges every block's candidate list");
    cg.addBlock("if (blockIdx.x == 0)", [&]() {
        cg.declareLds("u64", "_tk_lds", "GDB_TOPK_LDS_WORDS(" + K + ")");
        cg.addLine("GdbTopK<" + K + "> _f;");
        cg.addLine("gdb_topk_init<" + K + ">(_f);");
        // Candidates in batches of 8 per thread: the batch's 16-byte loads
        // are all issued before the first insertion, so the (cross-XCD,
        // uncached) candidate reads cost one round trip per batch instead of
        // one per candidate.
        cg.addBlock("for (u64 _c0 = 0; _c0 < " + NC + "; _c0 += 8ull * blockDim.x)", [&]() {
            cg.addLine("u64x2 _cv[8];");
            cg.addLine("#pragma unroll");
            cg.addLine("for (int _j = 0; _j < 8; ++_j) { const u64 _c = _c0 + (u64)_j * blockDim.x + threadIdx.x; _cv[_j] = _c < " +
                       NC + " ? reinterpret_cast<const u64x2*>(" + cand + ")[_c] : u64x2{0ull, 0ull}; }");
            cg.addLine("#pragma unroll");
            cg.addLine("for (int _j = 0; _j < 8; ++_j) if ((_cv[_j].x | _cv[_j].y) != 0) gdb_topk_push<" + K +
                       ">(_f, _cv[_j].x, _cv[_j].y);");
        });
        cg.addLine("gdb_topk_block_merge<" + K + ">(_f, _tk_lds);");
        cg.addLine("const u64* _tk_res = _tk_lds + 32 * " + K + ";");
        cg.addBlock("if (threadIdx.x == 0)", [&]() {
            cg.addLine("u64 _valid = 0;");
            cg.addLine("for (int _j = 0; _j < " + K + "; ++_j) { " + out_ + "[2 * _j] = _tk_res[2 * _j]; " + out_ +
                       "[2 * _j + 1] = _tk_res[2 * _j + 1]; if ((_tk_res[2 * _j] | _tk_res[2 * _j + 1]) != 0) ++_valid; }");
            cg.addLine(ctl + "[3] = _valid;");
        });
    });
    consume();
    auto& rs = cg.resultSchema();
    rs.kind = HipResultSchema::TOPK;
    rs.topk.bufferName = out_;
    rs.topk.ctlBuffer = ctl;
    rs.topk.k = k_;
    rs.topk.columns.clear();
    for (const auto& c : cols_)
        rs.topk.columns.push_back({c.name, c.word, c.shift, c.bits, c.complementOf, c.scale, c.date});
}

std::string HipTopKMerge::describe() const { return "TopKMerge(" + out_ + ", k=" + s(k_) + ")"; }

std::string HipTopK::describe() const { return "TopK(" + out_ + ", k=" + s(k_) + ", rank=(" + hi_ + ", " + lo_ + "))"; }

}  // namespace codegen
}  // namespace gdb

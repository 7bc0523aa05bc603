#pragma once
// HIP operator emitters — producer/consumer code generation for gfx950.
//
// Counterpart of codegen/operators/metal_operators.h:26-353. Each operator's
// produce() emits HIP C++ into a HipCodegen and calls its child's produce()
// with a consumer callback, so a tree of operators becomes one fused loop.
// The leaf scan owns the loop; unary operators wrap the per-row consumer;
// terminal (aggregating) operators emit declarations before the loop and a
// block/grid reduction after it.
//
// Row values: inside the loop every scanned column is a local named after the
// column (e.g. `f_qty`), loaded by the scan with 16-byte vector loads; CHAR
// columns are `const char*` into an LDS-staged copy of the tile. Expressions
// passed to operators are HIP C++ over those names (and over scalar
// parameters registered on the phase).
//
// Staged phases (item-batched emission): the scan hands its consumer all
// ITEMS rows of a thread at once, with an active mask `_act[ITEMS]`. A
// stageable operator (selection, compute, bitmap probe, hash probe) emits
// each step as its own unrolled loop over the ITEMS rows, so the random loads
// of ITEMS rows (bitmap words, first hash slots) are in flight together
// instead of one dependent chain per row; the first operator that is not
// stageable gets the usual per-row scope, wrapped in one loop over the active
// rows. Results are identical to per-row emission.

#include "../core/hip_codegen_base.h"

#include <memory>
#include <string>
#include <vector>

namespace gdb {
namespace codegen {

class HipOperator {
public:
    virtual ~HipOperator() = default;
    // Emit this operator and its subtree; consume() emits the consumer's code
    // for each row (or, in a staged phase, for the batch of active rows when
    // the consumer is stageable).
    void produce(HipCodegen& cg, ConsumerFn consume);
    virtual std::string describe() const = 0;
    virtual void describeTree(std::string& out, int depth) const;
    // True when the operator can consume a staged batch of rows.
    virtual bool stageable() const { return false; }

protected:
    virtual void produceImpl(HipCodegen& cg, ConsumerFn consume) = 0;
};

class HipUnaryOperator : public HipOperator {
protected:
    std::unique_ptr<HipOperator> child_;

public:
    explicit HipUnaryOperator(std::unique_ptr<HipOperator> child) : child_(std::move(child)) {}
    void describeTree(std::string& out, int depth) const override;
};

// ---------------------------------------------------------------------------
// Leaves
// ---------------------------------------------------------------------------

// Tiled grid-stride scan over a columnar table (metal_operators.cpp:28-49).
// A tile is blockDim.x * ITEMS rows; each thread owns ITEMS consecutive rows
// and loads them per column with one vector load; full tiles skip bounds
// checks. The index variable (`i` by default) is the absolute row number.
class HipGridStrideScan : public HipOperator {
public:
    struct ColumnDesc {
        std::string name;
        ColumnType type;
        bool lazy = false;  // char columns: read in place from HBM, not staged in LDS
    };
    explicit HipGridStrideScan(const std::string& table, const std::string& idxVar = "i");
    void addColumn(const std::string& column, ColumnType type);
    // A char column whose rows are read only in part or only for some rows
    // (a 2-byte code, a prefix, a compare after a selective filter): bound as
    // a pointer into the column in HBM instead of being staged in LDS with
    // the whole tile.
    void setLazy(const std::string& column);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;
    const std::string& tableName() const { return table_; }

private:
    std::string table_;
    std::string idxVar_;
    std::vector<ColumnDesc> columns_;
};

// Scan over the slots of a hash table built by an earlier phase; consumes
// occupied slots only. Exposes <prefix>_key (u32), <prefix>_payload (u32,
// high half of word 0) and <prefix>_w<k> (u64 words k >= 1) plus the slot
// index `_slot`.
class HipHashTableScan : public HipOperator {
public:
    HipHashTableScan(const std::string& table, const std::string& slotsExpr, int slotWords,
                     const std::string& prefix);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string table_, slotsExpr_, prefix_;
    int slotWords_;
};

// ---------------------------------------------------------------------------
// Pipeline operators
// ---------------------------------------------------------------------------

// WHERE filter: if (predicate) { consume(); } (metal_operators.cpp:63-69).
class HipSelection : public HipUnaryOperator {
public:
    bool stageable() const override { return true; }
    HipSelection(std::unique_ptr<HipOperator> child, const std::string& predicate);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string predicate_;
};

// const type var = expr; consume(); (metal_operators.cpp:81-95).
class HipComputeExpr : public HipUnaryOperator {
public:
    bool stageable() const override { return true; }
    HipComputeExpr(std::unique_ptr<HipOperator> child, const std::string& varName, const std::string& varType,
                   const std::string& expression);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string varName_, varType_, expression_;
};

// Wave-level stream compaction of a staged batch (no reference counterpart:
// the Metal pipelines run one row per thread). The active rows of the wave's
// 64 x ITEMS batch append the named values to a per-wave LDS queue (ballot +
// prefix popcount, no atomics); the consumer then runs once per queued row
// with the wave's lanes dense, so a selective probe issues one dependent
// random load per lane instead of one per row slot. Needs a staged phase
// whose launch bound sets the queue size (launchBound x ITEMS entries).
class HipWaveCompact : public HipUnaryOperator {
public:
    struct Var {
        std::string type, name;
    };
    bool stageable() const override { return true; }
    HipWaveCompact(std::unique_ptr<HipOperator> child, std::vector<Var> vars);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::vector<Var> vars_;
};

// Semi-join bitmap build: set bit `key` (metal_operators.cpp:108-116).
class HipBitmapBuild : public HipUnaryOperator {
public:
    // skewed: keys with heavy repeats across blocks (e.g. a Zipf column); the
    // set then also skips atomics whose bit a plain read already sees.
    bool stageable() const override { return !skewed_; }
    HipBitmapBuild(std::unique_ptr<HipOperator> child, const std::string& bitmapName, const std::string& keyExpr,
                   const std::string& bitsExpr, bool skewed = false);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string bitmap_, key_, bits_;
    bool skewed_ = false;
};

// Semi-join / anti-join probe (metal_operators.cpp:132-165).
class HipBitmapProbe : public HipUnaryOperator {
public:
    bool stageable() const override { return true; }
    HipBitmapProbe(std::unique_ptr<HipOperator> child, const std::string& bitmapName, const std::string& keyExpr,
                   const std::string& bitsExpr, bool anti = false);
    // Copy the bitmap's first `words` words into LDS and test keys below
    // them there (keys above read the global words): in a batched (range-
    // mode) body at the start of each work item, and with `plain` also in a
    // plain launch, once per block before its grid-stride loop (plain grids
    // are capped at the resident blocks, so each block copies once). A
    // batched kernel's LDS is sized by its largest body (a group-by table),
    // so a join item uses it without lowering occupancy; a plain launch with
    // a large prefix runs one block per CU and wants large blocks. 0 (the
    // default): no copy.
    void setLdsCache(int words, bool plain = false) {
        ldsWords_ = words;
        ldsPlain_ = plain;
    }
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string bitmap_, key_, bits_;
    bool anti_;
    int ldsWords_ = 0;
    bool ldsPlain_ = false;
};

// Open-addressing hash build of unique keys with a 32-bit payload. Replaces
// the reference's direct-address ArrayStore join map (metal_operators.cpp:185-200)
// and its host-built tables (query_preprocessing.cpp:352-378).
class HipHashBuild : public HipUnaryOperator {
public:
    bool stageable() const override { return true; }
    HipHashBuild(std::unique_ptr<HipOperator> child, const std::string& table, const std::string& keyExpr,
                 const std::string& payloadExpr, const std::string& slotsExpr, int slotWords);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string table_, key_, payload_, slots_;
    int slotWords_;
};

// Probe a table built by HipHashBuild; consumes rows whose key is present with
// `slotVar` (i64) and `<slotVar>_payload` (u32) in scope. Replaces
// ArrayLookup (metal_operators.cpp:213-225) and q9_ht_probe
// (metal_plan_builder.cpp:1598-1610).
class HipHashProbe : public HipUnaryOperator {
public:
    bool stageable() const override { return true; }
    HipHashProbe(std::unique_ptr<HipOperator> child, const std::string& table, const std::string& keyExpr,
                 const std::string& slotsExpr, int slotWords, const std::string& slotVar);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string table_, key_, slots_, slotVar_;
    int slotWords_;
};

// Open-addressing table with 64-bit keys (composite keys such as
// (ps_partkey, ps_suppkey) packed as partkey << 32 | suppkey) and a u64
// payload: slots [key, payload], empty key all ones. Build inserts each key
// once (first payload wins; overflow sets the phase status); probe consumes
// rows whose key is present with <slotVar> (i64 slot) and <slotVar>_payload.
// The reference hashes 64-bit keys by folding them the same way
// (metal_param_binding.h:20-26); its tables are host-built.
// distinct: only the row whose insert created the key is consumed (a key's
// first arrival), so a count downstream is COUNT(DISTINCT key). slotVar set:
// const i64 <slotVar> holds the key's slot for the consumer (a dense group
// id), and rows whose insert overflowed are not consumed.
class HipHash64Build : public HipUnaryOperator {
public:
    HipHash64Build(std::unique_ptr<HipOperator> child, const std::string& table, const std::string& keyExpr,
                   const std::string& payloadExpr, const std::string& slotsExpr, bool distinct = false,
                   const std::string& slotVar = "");
    // Distinct builds with a few hot keys (GROUP BY flags): an LDS set of the
    // keys each block already sent (gdb_block_first) keeps all but a block's
    // first arrival per key away from the global table.
    void setBlockFirst() { blockFirst_ = true; }
    // Build side of a key join: a second row with a key already present
    // fails the run with `message` after the phase (the join would
    // otherwise drop rows).
    void setUniqueKeys(const std::string& message) { uniqueMsg_ = message; }
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string table_, key_, payload_, slots_;
    bool distinct_ = false;
    std::string slotVar_;
    bool blockFirst_ = false;
    std::string uniqueMsg_;
};

// Binds columns of another table as kernel parameters (p_<column>), for
// expressions that read a joined row by its index (p_<column>[row]).
class HipBindColumns : public HipUnaryOperator {
public:
    HipBindColumns(std::unique_ptr<HipOperator> child, const std::string& table,
                   std::vector<std::pair<std::string, ColumnType>> columns);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string table_;
    std::vector<std::pair<std::string, ColumnType>> columns_;
};

// Dense group numbering over a HipHash64Build(distinct, slotVar) in the same
// phase: each first arrival takes the next id from word 3 of the table's
// status words and stores id + 1 at its key's slot in `idArray` (u32, one
// per slot); an id at or beyond `cap` flags the table's overflow word, so
// the caller can rebuild with a larger cap.
class HipDenseGroupId : public HipUnaryOperator {
public:
    HipDenseGroupId(std::unique_ptr<HipOperator> child, const std::string& table, const std::string& slotVar,
                    const std::string& slotsExpr, const std::string& idArray, long long cap);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string table_, slotVar_, slots_, ids_;
    long long cap_;
};

// Dense group numbering over a small direct-addressed key domain (a packed
// GROUP BY key below `size`): the block's first arrival per key (an LDS set)
// claims ids[key] with a compare-and-swap from 0, takes the next id from word
// 3 of `status` and stores id + 1; ids at or beyond `cap` leave the claim
// mark (~0) and flag word 0, so the caller rebuilds with the exact count.
class HipDirectGroupId : public HipUnaryOperator {
public:
    // sizeExpr: the key domain (a constant, or a symbol such as a column's
    // largest value + 1).
    // blockFirst: the LDS first-arrival filter (few hot keys); without it an
    // existing group costs one plain read of its id.
    HipDirectGroupId(std::unique_ptr<HipOperator> child, const std::string& keyExpr, const std::string& sizeExpr,
                     const std::string& idArray, const std::string& status, long long cap, bool blockFirst = true);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string key_, size_, ids_, status_;
    long long cap_;
    bool blockFirst_;
};

// Packs key-indexed byte flags (nonzero = set) into a bitmap (bit k of word
// k >> 5, the layout HipBitmapProbe reads): one thread per bitmap word, two
// 16-byte loads of its 32 flags. With HipByteFlagBuild in front it builds a
// bitmap without atomic ORs (plain byte stores, then this pass).
class HipFlagsToBitmap : public HipOperator {
public:
    HipFlagsToBitmap(const std::string& flagsName, const std::string& bitmapName, const std::string& bitsExpr)
        : flags_(flagsName), bitmap_(bitmapName), bits_(bitsExpr) {}
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override { return "FlagsToBitmap(" + flags_ + " -> " + bitmap_ + ")"; }

private:
    std::string flags_, bitmap_, bits_;
};

// Key-indexed u32 store array[key] = value (value != 0) for keys that must
// be unique (a join's key side): a compare-and-swap from 0 claims the
// entry, and a key seen twice fails the run with `message` after the phase.
// Keys outside [0, size) are skipped.
class HipDenseUniqueStore : public HipUnaryOperator {
public:
    HipDenseUniqueStore(std::unique_ptr<HipOperator> child, const std::string& arrayName, const std::string& keyExpr,
                        const std::string& sizeExpr, const std::string& valueExpr, const std::string& message);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string array_, key_, size_, value_, message_;
};

// Row check: counts the rows for which `condition` is false and fails the
// run after the phase with `message` when any did (e.g. a fingerprint whose
// group representative holds other key bytes). Every row is consumed.
class HipRowCheck : public HipUnaryOperator {
public:
    HipRowCheck(std::unique_ptr<HipOperator> child, const std::string& name, const std::string& condition,
                const std::string& message);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string name_, cond_, message_;
};

// anti: rows whose key is absent are consumed (NOT EXISTS / NOT IN); the
// slot variable is then not defined.
class HipHash64Probe : public HipUnaryOperator {
public:
    HipHash64Probe(std::unique_ptr<HipOperator> child, const std::string& table, const std::string& keyExpr,
                   const std::string& slotsExpr, const std::string& slotVar, bool anti = false);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string table_, key_, slots_, slotVar_;
    bool anti_ = false;
};

// Atomic add of an i64 value into word `word` of a hash-table slot
// (the group-by-join-key aggregate; counterpart of MetalAtomicAgg,
// metal_operators.cpp:525-533, with the slot as the bucket).
class HipSlotAtomicAdd : public HipUnaryOperator {
public:
    HipSlotAtomicAdd(std::unique_ptr<HipOperator> child, const std::string& table, int slotWords,
                     const std::string& slotVar, int word, const std::string& valueExpr);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string table_, slotVar_, value_;
    int slotWords_, word_;
};

// Direct-address lookup with a sentinel guard (metal_operators.cpp:203-225).
// elemType: "u32" (counters) or "i64" (sums).
class HipArrayLookup : public HipUnaryOperator {
public:
    HipArrayLookup(std::unique_ptr<HipOperator> child, const std::string& arrayName, const std::string& keyExpr,
                   const std::string& sizeExpr, const std::string& resultVar, long long sentinel,
                   const std::string& elemType = "u32");
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string array_, key_, size_, var_, elem_;
    long long sentinel_;
};

// LIKE '%w1%w2%' (NOT LIKE with negate) on a char[width] column staged in
// LDS, byte-exact (q13_comment_match, metal_plan_builder.cpp:1249-1266, for
// any two words): bit-parallel Shift-And automata over the row with per-block
// LDS byte tables built from the words buffer (w1 at byte 0, w2 at byte 32;
// lengths in two scalars). The phase must register the words buffer and the
// length scalars.
class HipLikeFilter : public HipUnaryOperator {
public:
    HipLikeFilter(std::unique_ptr<HipOperator> child, const std::string& column, int width,
                  const std::string& wordsBuffer, const std::string& l1Scalar, const std::string& l2Scalar, bool negate);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string column_;
    int width_;
    std::string words_, l1_, l2_;
    bool negate_;
};

// Semi-join membership as a value: const u32 <var> = bit `key` is set;
// every row is consumed (CASE WHEN <key in build side> THEN ... expressions,
// e.g. TPC-H Q12's priority class and Q14's promotional parts).
class HipBitmapFlag : public HipUnaryOperator {
public:
    HipBitmapFlag(std::unique_ptr<HipOperator> child, const std::string& bitmapName, const std::string& keyExpr,
                  const std::string& bitsExpr, const std::string& var);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string bitmap_, key_, bits_, var_;
};

// Key-indexed byte flags: flags[key] = 1 with plain byte stores (no
// read-modify-write; every writer stores the same value, so the races are
// benign). For a small key domain hit by many rows from every block (e.g. the
// customers that have orders, ~10 orders per customer), where a bitmap's
// atomic ORs would serialise on a few thousand words. Keys outside
// [0, size) are skipped; the flags are zeroed before the phase.
// valueExpr (default 1, must be 1..255 for a present key and the same for
// every row of a key) makes the flags a direct-address byte payload (e.g. an
// order's priority class), as the reference's ArrayStore
// (metal_operators.cpp:185-200) with byte values.
// tested=true reads the flag first and skips the store when it already
// holds the value, behind a per-block LDS cache of recently seen keys: for
// skewed keys (a Zipf column), whose hot keys then cost an LDS read instead
// of a random global read, and whose lines are not rewritten by every block.
class HipByteFlagBuild : public HipUnaryOperator {
public:
    HipByteFlagBuild(std::unique_ptr<HipOperator> child, const std::string& flagsName, const std::string& keyExpr,
                     const std::string& sizeExpr, const std::string& valueExpr = "1", bool tested = false);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string flags_, key_, size_, value_;
    bool tested_ = false;
};

// Key-indexed u32 store array[key] = value (zeroed before its phase, so 0
// reads as absent; store value + 1 for values that may be 0): the
// reference's direct-address ArrayStore join map (metal_operators.cpp:185-200)
// for keys that are unique on the build side (e.g. o_orderkey), read back
// with HipArrayLookup. Keys outside [0, size) are skipped.
class HipDenseStore : public HipUnaryOperator {
public:
    HipDenseStore(std::unique_ptr<HipOperator> child, const std::string& arrayName, const std::string& keyExpr,
                  const std::string& sizeExpr, const std::string& valueExpr);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string array_, key_, size_, value_;
};

// Semi-/anti-join probe of byte flags built by HipByteFlagBuild: rows whose
// key has a non-zero flag (anti: a zero flag) are consumed; with `var` set,
// const u32 <var> holds the flag's value (semi-join form only).
// keepAll: every row is consumed and <var> is the flag (0 when absent), a
// CASE WHEN <key in build side> value rather than a join.
class HipByteFlagProbe : public HipUnaryOperator {
public:
    HipByteFlagProbe(std::unique_ptr<HipOperator> child, const std::string& flagsName, const std::string& keyExpr,
                     const std::string& sizeExpr, bool anti, const std::string& var = "", bool keepAll = false);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string flags_, key_, size_, var_;
    bool anti_, keepAll_;
};

// Direct-address i64 minimum (or maximum) array[index] = min(array[index],
// value) with global atomics (the array starts at INT64_MAX, or INT64_MIN);
// rows whose index is outside [0, size) are skipped; runs of equal indices in
// a thread's rows combine before the atomic, as in HipDenseAtomicAdd. A
// per-key MIN/MAX subquery (TPC-H Q2's minimum supply cost per part, Q21's
// smallest and largest supplier per order).
class HipDenseAtomicMinMax : public HipUnaryOperator {
public:
    HipDenseAtomicMinMax(std::unique_ptr<HipOperator> child, const std::string& arrayName,
                         const std::string& indexExpr, const std::string& sizeExpr, const std::string& valueExpr,
                         bool isMax);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string array_, index_, size_, value_;
    bool max_;
};

// Values an earlier phase left in a buffer, read once at kernel start:
// const i64 <name> = (i64)buffer[word]. Carries a scalar subquery's result
// (TPC-H Q22's average balance as sum and count) to a later phase without a
// host round trip.
class HipBufferScalars : public HipUnaryOperator {
public:
    HipBufferScalars(std::unique_ptr<HipOperator> child, const std::string& buffer, const std::string& sizeExpr,
                     std::vector<std::pair<std::string, int>> vars);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string buffer_, size_;
    std::vector<std::pair<std::string, int>> vars_;
};

// Direct-address i64 accumulation array[index] += value with global atomics
// (the reference's per-key atomic aggregation into a key-indexed array,
// MetalAtomicAgg, metal_operators.cpp:525-533); rows whose index is outside
// [0, size) are skipped. The array is zeroed before its phase.
class HipDenseAtomicAdd : public HipUnaryOperator {
public:
    HipDenseAtomicAdd(std::unique_ptr<HipOperator> child, const std::string& arrayName, const std::string& indexExpr,
                      const std::string& sizeExpr, const std::string& valueExpr);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string array_, index_, size_, value_;
};

// Append one record of u64 fields per consumed row to `out` (capacity
// capExpr records; a wave-aggregated cursor in <out>_ctl word 0). The
// records are the plan's result (ROWS): ordered on the host by the given
// fields, then cut to `limit` rows (0 = all). Rows beyond the capacity set the
// phase's overflow status.
class HipAppendRows : public HipUnaryOperator {
public:
    HipAppendRows(std::unique_ptr<HipOperator> child, const std::string& out, const std::string& capExpr);
    void addField(const std::string& name, const std::string& expr, int scale = 0, bool date = false);
    void orderBy(const std::string& field, bool desc);
    void setLimit(int limit) { limit_ = limit; }
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    struct Field {
        std::string name, expr;
        int scale;
        bool date;
    };
    std::string out_, cap_;
    std::vector<Field> fields_;
    std::vector<std::pair<std::string, bool>> order_;
    int limit_ = 0;
};

// ---------------------------------------------------------------------------
// Aggregation (terminal) operators
// ---------------------------------------------------------------------------

// Fused group-by over a small dense bucket domain (MetalKeyedAgg,
// metal_operators.cpp:375-502, which also covers MetalTGReduce with one
// bucket). Exact integer aggregates: SumI64, SumI128 (two 64-bit limbs:
// sum of the low 32 bits and sum of the arithmetic high 32 bits), Count,
// MinI64, MaxI64. Strategy by bucket count:
//   <= 8 buckets: per-thread register accumulators, wave DPP reduction, LDS
//                 block reduction, one global atomic per block and slot;
//   <= LDS budget: block-shared LDS table with LDS atomics, flushed once;
//   otherwise:   one global atomic per row.
class HipKeyedAgg : public HipUnaryOperator {
public:
    struct Aggregate {
        std::string name;
        AggKind kind;
        std::string valueExpr;
        int scaleDigits;
        int wordOffset = 0;
    };
    HipKeyedAgg(std::unique_ptr<HipOperator> child, const std::string& outputName, const std::string& bucketExpr,
                int numBuckets);
    void addAggregate(const std::string& name, AggKind kind, const std::string& valueExpr, int scaleDigits = 0);
    void setBucketLabels(const std::vector<std::string>& labels, const std::string& keyColumn);
    void addAverage(const std::string& name, const std::string& numerator, const std::string& denominator);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string out_, bucket_;
    int numBuckets_;
    std::vector<Aggregate> aggs_;
    std::vector<std::string> labels_;
    std::string keyColumn_ = "bucket";
    std::vector<std::pair<std::string, std::pair<std::string, std::string>>> avgs_;
};

// Count per bucket into u32 counters (MetalAtomicCount,
// metal_operators.cpp:550-594): LDS histogram when the static size fits,
// global atomics otherwise.
class HipAtomicCount : public HipUnaryOperator {
public:
    HipAtomicCount(std::unique_ptr<HipOperator> child, const std::string& arrayName, const std::string& bucketExpr,
                   const std::string& sizeExpr, int staticSize);
    // Make the counters the plan's result: rows (key, count) for non-zero
    // counters, optionally ordered count desc then key desc (Q13).
    void setResult(const std::string& keyName, const std::string& valueName, bool sortByCountDesc);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string array_, bucket_, size_;
    int static_;
    bool isResult_ = false;
    std::string keyName_, valueName_;
    bool sortDesc_ = false;
};

// Hash group-by with sum/count/min/max of an i64 value per 32-bit key, into
// a global open-addressing table of 32-byte slots. Variants:
//   GLOBAL — every row merges into the global table;
//   LDS    — a block-local LDS table absorbs repeated (hot) keys; rows whose
//            LDS probe window is full spill to the global table; the LDS
//            table is flushed once at the end of the block.
class HipHashGroupBy : public HipUnaryOperator {
public:
    enum Variant { GLOBAL = 0, LDS = 1 };
    HipHashGroupBy(std::unique_ptr<HipOperator> child, const std::string& table, const std::string& keyExpr,
                   const std::string& valueExpr, const std::string& slotsExpr, Variant variant, int ldsSlots);
    // LDS variant: take the number of LDS slots in use (log2) from a scalar
    // at run time, up to the compiled capacity ldsSlots (autotuned per
    // pipeline without recompiling).
    void setRuntimeLdsBits(const std::string& scalarName) { ldsBitsScalar_ = scalarName; }
    // LDS probe window: rows whose key finds no slot within it spill to the
    // global table (a long window absorbs more keys, a short one makes the
    // misses of high-cardinality keys cheap).
    void setWindow(int w) { window_ = w; }
    // LDS variant: at the end of a block (or batched work item) store the
    // LDS table whole as one of at most maxExpr partial tables instead of
    // merging it into the global table slot by slot — with many blocks over
    // few groups that merge is the same keys' atomics repeated once per
    // block. HipGroupByMerge folds the partials into the table in the next
    // phase; a block past the last partial merges directly as before.
    void setPartials(const std::string& maxExpr) { partialsMax_ = maxExpr; }
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string table_, key_, value_, slots_;
    Variant variant_;
    int ldsSlots_;
    std::string ldsBitsScalar_;
    int window_ = 16;
    std::string partialsMax_;
};

// Second phase of HipHashGroupBy with partials: the partial LDS tables
// (buffer <table>_pt, count in <table>_ptn) are folded into a block-local LDS
// table — a block or work item covers a contiguous run of partials — which
// is then merged into the global table, so a key costs one global merge per
// block of this phase instead of one per block of the scan. Rows of the phase
// = partial slots (maxExpr x ldsSlots); slots beyond the partials written, or
// beyond the LDS slots in use, are skipped.
class HipGroupByMerge : public HipOperator {
public:
    HipGroupByMerge(const std::string& table, const std::string& slotsExpr, int ldsSlots, const std::string& partialsMax)
        : table_(table), slots_(slotsExpr), ldsSlots_(ldsSlots), partialsMax_(partialsMax) {}
    void setRuntimeLdsBits(const std::string& scalarName) { ldsBitsScalar_ = scalarName; }
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    static constexpr int kMergeWindow = 16;
    std::string table_, slots_;
    int ldsSlots_;
    std::string partialsMax_, ldsBitsScalar_;
};

// Partitioned hash group-by: no counterpart in the reference, whose
// high-cardinality aggregation is one atomic per row into a direct-addressed
// array (metal_operators.cpp:525-533). Five phases, no global atomics:
//   1. HipPartitionCount   — per logical block (a fixed row range), hot keys
//      are aggregated in an LDS table; a row whose key the table cannot hold
//      is "cold" and only counted, per hash partition, in LDS. The final LDS
//      table (hot partial aggregates) and both count vectors (cold rows,
//      partials) are written out.
//   2. HipPartitionOffsets — per partition, counts -> exclusive offsets over
//      the logical blocks, and totals.
//   3. HipPartitionBases   — one block scans the totals into each partition's
//      base in the packed cold-row and partial streams.
//   4. HipPartitionScatter — the same logical blocks re-read their rows; a
//      row is cold iff its key is absent from the block's final LDS table (a
//      full probe window never admits a key later, so this is exactly the
//      pass-1 decision), and cold rows and partials are written to exact
//      positions (LDS counters), packed by partition.
//   5. HipPartitionAgg     — one block per partition aggregates its cold rows
//      and partials in an LDS table that is the partition's slice of the
//      group table (4096 slots of [key|count, sum, min, max]; home slot = the
//      12 hash bits after the partition bits), and stores the slice whole.
// Partition of a key = top bits of key * 2654435769; partitions = slots / 4096.
struct PartitionConfig {
    std::string prefix;     // buffer/symbol prefix, e.g. "d_gbp"
    std::string table;      // group table, e.g. "d_gb_ht"
    std::string slots;      // slots expression of the group table
    std::string rowsSym;    // symbol: rows of the scanned table (all of them)
    std::string scanTable;  // scanned table name
    int ldsSlots = 4096;    // pass-1 LDS table capacity (power of two, 64..4096)
    // Probe window of the pass-1 LDS table (pass 4 uses the same): a key
    // whose window is full is cold. Short, because most rows of a
    // high-cardinality column are cold and pay the whole window in LDS loads.
    int window = 2;
    // Partitions pass 1 counts cold rows and partials at (an expression;
    // empty = the table's partitions). When it is larger, the offsets pass
    // sums each table partition's consecutive count partitions (a key's
    // partition at 2^b partitions is its partition at 2^c >= 2^b shifted
    // right by c - b), so pass 1 need not know the table size.
    std::string countParts;
    // Pass 1 also sets a byte flag (buffer `flags`, index key - flagMin, a
    // scalar; flagBits flags) for every key it sees — cold rows' keys as they
    // come, the LDS table's keys at the end: the exact distinct-count pre-pass
    // fused into pass 1. Empty = off.
    std::string flags, flagMin, flagBits;
    // Cold rows as 8-byte records (key | value << 32) instead of 16-byte
    // (key, value) ones: half the scatter's writes and the fold's reads. Exact
    // only while every value fits in 32 bits, so the scatter checks each one
    // and reports a value that does not in <prefix>_rec_status[0]; the caller
    // then reruns with wide records.
    bool narrow = false;
    static constexpr int kPartSlots = 4096;
    static constexpr int kLogicalBlocks = 256;
    // Largest partition count whose LDS counters fit beside the LDS table.
    int maxParts() const;
};

class HipPartitionCount : public HipUnaryOperator {
public:
    HipPartitionCount(std::unique_ptr<HipOperator> child, PartitionConfig cfg, const std::string& keyExpr,
                      const std::string& valueExpr);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    PartitionConfig c_;
    std::string key_, value_;
};

class HipPartitionOffsets : public HipOperator {
public:
    explicit HipPartitionOffsets(PartitionConfig cfg) : c_(std::move(cfg)) {}
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override { return "PartitionOffsets(" + c_.prefix + ")"; }

private:
    PartitionConfig c_;
};

class HipPartitionBases : public HipOperator {
public:
    explicit HipPartitionBases(PartitionConfig cfg) : c_(std::move(cfg)) {}
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override { return "PartitionBases(" + c_.prefix + ")"; }

private:
    PartitionConfig c_;
};

class HipPartitionScatter : public HipUnaryOperator {
public:
    HipPartitionScatter(std::unique_ptr<HipOperator> child, PartitionConfig cfg, const std::string& keyExpr,
                        const std::string& valueExpr);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    PartitionConfig c_;
    std::string key_, value_;
};

class HipPartitionAgg : public HipOperator {
public:
    explicit HipPartitionAgg(PartitionConfig cfg) : c_(std::move(cfg)) {}
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override { return "PartitionAgg(" + c_.prefix + " -> " + c_.table + ")"; }

private:
    PartitionConfig c_;
};

// Top-k by a 128-bit rank (hi, lo), larger first, rank (0,0) reserved.
// Per-thread register lists, wave then block selection (DPP argmax rounds),
// one candidate list per block; a following one-block phase (HipTopKMerge,
// from makeMerge) merges the lists.
// Replaces the reference's CPU full sort (codegen/codegen_main.cpp:1211-1248).
class HipTopK : public HipUnaryOperator {
public:
    struct Column {
        std::string name;
        int word;            // 0 = rank hi, 1 = rank lo
        int shift;
        int bits;
        long long complementOf;  // value = complementOf - field when != 0
        int scale;               // fixed-point digits of the shown value
        bool date;               // YYYYMMDD shown as a date
    };
    HipTopK(std::unique_ptr<HipOperator> child, const std::string& outputName, int k, const std::string& rankHiExpr,
            const std::string& rankLoExpr);
    void addColumn(const std::string& name, int word, int shift, int bits, long long complementOf = 0, int scale = 0,
                   bool date = false);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;
    // The operator of the final phase: one block merges every block's
    // candidates (the phase boundary orders the candidate writes, so no
    // fences or tickets) and defines the result.
    std::unique_ptr<HipOperator> makeMerge() const;
    static constexpr int kMaxBlocks = 512;

private:
    std::string out_, hi_, lo_;
    int k_;
    std::vector<Column> cols_;
};

// Final top-k phase (leaf): one block of the merge of kMaxBlocks candidate
// lists written by a HipTopK phase.
class HipTopKMerge : public HipOperator {
public:
    HipTopKMerge(const std::string& outputName, int k, std::vector<HipTopK::Column> cols);
    void produceImpl(HipCodegen& cg, ConsumerFn consume) override;
    std::string describe() const override;

private:
    std::string out_;
    int k_;
    std::vector<HipTopK::Column> cols_;
};

}  // namespace codegen
}  // namespace gdb

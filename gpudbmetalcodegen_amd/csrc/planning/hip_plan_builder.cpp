#include "hip_plan_builder.h"
#include "plan_util.h"
#include "tpch_plans.h"

#include <sstream>
#include <stdexcept>

namespace gdb {
namespace codegen {

// ---------------------------------------------------------------------------
// Catalog. Column names carry a table prefix so they are unique inside a
// kernel; the reference's TPC-H catalog (codegen/planning/tpch_schema.h:47-166)
// has no int64 or uint8 types, which these workloads need.
// ---------------------------------------------------------------------------

namespace {

ColumnType I32{ColType::I32, 0};
ColumnType I64{ColType::I64, 0};
ColumnType U8{ColType::U8, 0};
ColumnType CHAR25{ColType::CHAR, 25};
ColumnType chr(int width) { return ColumnType{ColType::CHAR, width}; }

const std::vector<CatalogTable>& catalog() {
    static const std::vector<CatalogTable> tables = {
        {"F",
         {{"f_qty", I32},
          {"f_price", I64},
          {"f_disc", I32},
          {"f_tax", I32},
          {"f_flag", U8},
          {"f_status", U8},
          {"f_date", I32},
          {"f_fkey", I32}}},
        {"A", {{"a_key", I32}, {"a_seg", U8}}},
        {"B", {{"b_key", I32}, {"b_fkey", I32}, {"b_date", I32}}},
        {"S", {{"s_fkey", I32}, {"s_text", CHAR25}}},
        {"G", {{"g_key", I32}, {"g_val", I64}}},
        {"GBFL", {{"gbfl_flag", U8}}},
        {"P", {{"p_key", I32}, {"p_date", I32}, {"p_val", I64}, {"p_flag", U8}}},
        // TPC-H tables as loaded from .colbin (tpch_schema.h:66-162): INT and
        // DATE (YYYYMMDD) int32, FLOAT as exact int64 hundredths, CHAR1 one
        // byte, CHAR_FIXED char[width].
        {"lineitem",
         {{"l_orderkey", I32}, {"l_partkey", I32}, {"l_suppkey", I32}, {"l_linenumber", I32}, {"l_quantity", I64},
          {"l_extendedprice", I64}, {"l_discount", I64}, {"l_tax", I64}, {"l_returnflag", U8}, {"l_linestatus", U8},
          {"l_shipdate", I32}, {"l_commitdate", I32}, {"l_receiptdate", I32}, {"l_shipinstruct", chr(25)},
          {"l_shipmode", chr(2)}, {"l_comment", chr(44)}}},
        {"orders",
         {{"o_orderkey", I32}, {"o_custkey", I32}, {"o_orderstatus", U8}, {"o_totalprice", I64}, {"o_orderdate", I32},
          {"o_orderpriority", U8}, {"o_clerk", chr(15)}, {"o_shippriority", I32}, {"o_comment", chr(79)}}},
        {"customer",
         {{"c_custkey", I32}, {"c_name", chr(25)}, {"c_address", chr(40)}, {"c_nationkey", I32}, {"c_phone", chr(15)},
          {"c_acctbal", I64}, {"c_mktsegment", U8}, {"c_comment", chr(117)}}},
        {"part",
         {{"p_partkey", I32}, {"p_name", chr(55)}, {"p_mfgr", chr(25)}, {"p_brand", chr(10)}, {"p_type", chr(25)},
          {"p_size", I32}, {"p_container", chr(10)}, {"p_retailprice", I64}, {"p_comment", chr(23)}}},
        {"supplier",
         {{"s_suppkey", I32}, {"s_name", chr(25)}, {"s_address", chr(40)}, {"s_nationkey", I32}, {"s_phone", chr(15)},
          {"s_acctbal", I64}, {"s_comment", chr(101)}}},
        {"partsupp",
         {{"ps_partkey", I32}, {"ps_suppkey", I32}, {"ps_availqty", I32}, {"ps_supplycost", I64},
          {"ps_comment", chr(199)}}},
        {"nation", {{"n_nationkey", I32}, {"n_name", chr(25)}, {"n_regionkey", I32}, {"n_comment", chr(152)}}},
        {"region", {{"r_regionkey", I32}, {"r_name", chr(25)}, {"r_comment", chr(152)}}},
    };
    return tables;
}


// ---------------------------------------------------------------------------
// Q1 — filter + group-by (flag, status) with exact fixed-point aggregates.
// Bucket encoding follows codegen/planning/metal_plan_builder.cpp:335
// (flag A->0, N->2, R->4; status F->0, else 1). Aggregates follow the
// reference's Q1 slot list (:342-349) in exact integers: price in cents,
// discount/tax in hundredths, so disc_price carries 4 fixed-point digits and
// charge 6 (int128).
// ---------------------------------------------------------------------------
HipQueryPlan buildQ1(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q1";
    auto scan = makeScan("F", {"f_qty", "f_price", "f_disc", "f_tax", "f_flag", "f_status", "f_date"});
    auto sel = std::make_unique<HipSelection>(std::move(scan), "f_date <= q1_cutoff");
    const std::string bucket =
        "((f_flag == 'A' ? 0 : (f_flag == 'N' ? 2 : 4)) + (f_status == 'F' ? 0 : 1))";
    auto agg = std::make_unique<HipKeyedAgg>(std::move(sel), "d_q1_aggs", bucket, 6);
    agg->addAggregate("sum_qty", AggKind::SumI64, "f_qty", 0);
    agg->addAggregate("sum_base_price", AggKind::SumI64, "f_price", 2);
    agg->addAggregate("sum_disc_price", AggKind::SumI64, "f_price * (i64)(100 - f_disc)", 4);
    agg->addAggregate("sum_charge", AggKind::SumI128, "f_price * (i64)(100 - f_disc) * (i64)(100 + f_tax)", 6);
    agg->addAggregate("sum_disc", AggKind::SumI64, "f_disc", 2);
    agg->addAggregate("count_order", AggKind::Count, "1", 0);
    agg->setBucketLabels({"A|F", "A|O", "N|F", "N|O", "R|F", "R|O"}, "f_flag,f_status");
    agg->addAverage("avg_qty", "sum_qty", "count_order");
    agg->addAverage("avg_price", "sum_base_price", "count_order");
    agg->addAverage("avg_disc", "sum_disc", "count_order");
    // ITEMS=4 measured best on MI355X: 16-byte lane-contiguous loads for the
    // int32 columns; at ITEMS=8 each lane's 64-byte int64 chunk halves the
    // achieved bandwidth
    auto& ph = appendPhase(plan, "Q1_reduce", std::move(agg), optInt(o, "block", 256), optInt(o, "items", 4),
                           {2, 4, 8});
    ph.launchBound = 256;
    ph.scalarParams = {"q1_cutoff"};
    ph.prefetch = optInt(o, "prefetch", 0) != 0;
    plan.requiredScalars = {"q1_cutoff"};
    plan.chunkable = true;
    return plan;
}

// Q6 — scalar revenue over a date range, discount band and quantity bound
// (metal_plan_builder.cpp:253-303, sql/q6.sql), exact: sum(price * disc).
HipQueryPlan buildQ6(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q6";
    auto scan = makeScan("F", {"f_qty", "f_price", "f_disc", "f_date"});
    auto sel = std::make_unique<HipSelection>(
        std::move(scan),
        "f_date >= q6_date_lo && f_date < q6_date_hi && f_disc >= q6_disc_lo && f_disc <= q6_disc_hi && f_qty < q6_qty_hi");
    auto agg = std::make_unique<HipKeyedAgg>(std::move(sel), "d_q6_sum", "0", 1);
    agg->addAggregate("revenue", AggKind::SumI64, "f_price * (i64)f_disc", 4);
    agg->addAggregate("count", AggKind::Count, "1", 0);
    auto& ph = appendPhase(plan, "Q6_reduce", std::move(agg), optInt(o, "block", 256), optInt(o, "items", 4), {2, 4, 8});
    ph.launchBound = 256;
    ph.scalarParams = {"q6_date_lo", "q6_date_hi", "q6_disc_lo", "q6_disc_hi", "q6_qty_hi"};
    plan.requiredScalars = ph.scalarParams;
    plan.chunkable = true;
    return plan;
}

// Q3 analog — A(seg) semi-join bitmap, hash build of qualifying B keys with
// b_date as payload, probe from F with the per-B-key revenue summed in the
// slot, then top-10 by (revenue desc, b_date asc, b_key asc).
// Reference: buildQ3Plan_byName (metal_plan_builder.cpp:1170-1239) and the
// CPU top-10 (codegen_main.cpp:1211-1248).
void q3BitmapPhase(HipQueryPlan& plan, const PlanOptions& o) {
    auto scan = makeScan("A", {"a_key", "a_seg"});
    auto sel = std::make_unique<HipSelection>(std::move(scan), "(i64)a_seg == q3_seg");
    auto bm = std::make_unique<HipBitmapBuild>(std::move(sel), "d_q3_a_bitmap", "a_key", "a_bits");
    auto& ph = appendPhase(plan, "Q3_build_a_bitmap", std::move(bm), 256, optInt(o, "items_a", 4));
    ph.scalarParams = {"q3_seg"};
}

// Exact count of the B rows the hash table will hold (the plain form's
// sizing pre-pass).
void q3CountPhase(HipQueryPlan& plan, const PlanOptions& o) {
    auto scan = makeScan("B", {"b_fkey", "b_date"});
    auto sel = std::make_unique<HipSelection>(std::move(scan), "(i64)b_date < q3_date");
    auto probe = std::make_unique<HipBitmapProbe>(std::move(sel), "d_q3_a_bitmap", "b_fkey", "a_bits");
    const int aWords = optInt(o, "a_lds_words", 0);
    probe->setLdsCache(aWords, true);
    auto cnt = std::make_unique<HipKeyedAgg>(std::move(probe), "d_q3_bcount", "0", 1);
    cnt->addAggregate("b_rows", AggKind::Count, "1", 0);
    auto& ph = appendPhase(plan, "Q3_count_b", std::move(cnt), optInt(o, "block_count", aWords > 0 ? 1024 : 256), 8);
    ph.scalarParams = {"q3_date"};
    ph.staged = optInt(o, "staged", 1) != 0;  // measured: staged bitmap loads help this phase
}

HipQueryPlan buildQ3Size(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q3_SIZE";
    q3BitmapPhase(plan, o);
    q3CountPhase(plan, o);
    plan.requiredScalars = {"q3_seg", "q3_date"};
    plan.requiredSymbols = {"a_bits"};
    return plan;
}

HipQueryPlan buildQ3(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q3";
    const bool probeBitmap = optInt(o, "probe_bitmap", 1) != 0;
    const bool bflags = optInt(o, "bflags", 0) != 0;
    const bool prepass = opt(o, "sizing", "prepass") == "prepass";
    q3BitmapPhase(plan, o);
    if (prepass) {
        q3CountPhase(plan, o);
        plan.sizingRules.push_back({"Q3_count_b", "b_slots", "d_q3_bcount", 0, 2, 64, true});
    }
    {
        auto scan = makeScan("B", {"b_key", "b_fkey", "b_date"});
        auto sel = std::make_unique<HipSelection>(std::move(scan), "(i64)b_date < q3_date");
        auto probe = std::make_unique<HipBitmapProbe>(std::move(sel), "d_q3_a_bitmap", "b_fkey", "a_bits");
        // a_lds_words=N: the A bitmap's first N words are tested in LDS (one
        // block of 1024 per CU then, unless block_b / block_count say otherwise)
        const int aWords = optInt(o, "a_lds_words", 0);
        probe->setLdsCache(aWords, true);
        std::unique_ptr<HipOperator> top =
            std::make_unique<HipHashBuild>(std::move(probe), "d_q3_b_ht", "b_key", "b_date", "b_slots", 2);
        // bflags=1: the B-key bitmap is built as byte flags (plain stores, no
        // atomic ORs) and packed into the bitmap by the next phase
        if (probeBitmap && bflags)
            top = std::make_unique<HipByteFlagBuild>(std::move(top), "d_q3_b_flags", "b_key", "b_bits");
        else if (probeBitmap)
            top = std::make_unique<HipBitmapBuild>(std::move(top), "d_q3_b_bitmap", "b_key", "b_bits");
        // block 512: measured 0.163 vs 0.166 ms at 256 (1024: 0.164, 128: 0.188)
        auto& ph = appendPhase(plan, "Q3_build_b_table", std::move(top), optInt(o, "block_b", aWords > 0 ? 1024 : 512),
                               optInt(o, "items_b", 4));
        ph.scalarParams = {"q3_date"};
        // staged=1 puts the home-slot CAS of a thread's ITEMS rows in flight
        // together; per-row emission measures faster here (0.166 vs 0.177 ms
        // at 15M rows: ~10% of the rows reach the CAS, so a batch rarely
        // holds more than one)
        ph.staged = optInt(o, "staged", 0) != 0;
        ph.prefetch = optInt(o, "prefetch", 0) != 0;
    }
    if (probeBitmap && bflags) {
        auto& pk = appendPhase(plan, "Q3_pack_b_bitmap",
                               std::make_unique<HipFlagsToBitmap>("d_q3_b_flags", "d_q3_b_bitmap", "b_bits"), 256, 1, {1});
        pk.launchBound = 256;
    }
    {
        // late=1 (default): price and discount are read only for the rows
        // that pass the date filter and the bitmap (about 5% at the headline
        // workload's filters): measured 0.404 ms vs 0.437 ms for the probe
        // with all four columns streamed (late=0)
        const bool late = optInt(o, "late", 1) != 0;
        auto scan = late ? makeScan("F", {"f_fkey", "f_date"}) : makeScan("F", {"f_fkey", "f_price", "f_disc", "f_date"});
        std::unique_ptr<HipOperator> cur = std::make_unique<HipSelection>(std::move(scan), "(i64)f_date > q3_date");
        if (probeBitmap)
            cur = std::make_unique<HipBitmapProbe>(std::move(cur), "d_q3_b_bitmap", "f_fkey", "b_bits");
        // compact=1: the rows that pass the date filter and the bitmap are
        // queued per wave, so the hash probe and slot add run lanes-dense
        const bool compact = optInt(o, "compact", late ? 1 : 0) != 0;
        std::string rev = "f_price * (i64)(100 - f_disc)";
        if (late) {
            if (compact)
                cur = std::make_unique<HipWaveCompact>(
                    std::move(cur), std::vector<HipWaveCompact::Var>{{"i32", "f_fkey"}, {"u64", "i"}});
            const auto& ft = catalogTable("F");
            cur = std::make_unique<HipBindColumns>(
                std::move(cur), "F",
                std::vector<std::pair<std::string, ColumnType>>{{"f_price", ft.type("f_price")}, {"f_disc", ft.type("f_disc")}});
            cur = std::make_unique<HipComputeExpr>(std::move(cur), "_rev", "i64", "p_f_price[i] * (i64)(100 - p_f_disc[i])");
            rev = "_rev";
        } else if (compact) {
            cur = std::make_unique<HipComputeExpr>(std::move(cur), "_rev", "i64", rev);
            cur = std::make_unique<HipWaveCompact>(std::move(cur),
                                                   std::vector<HipWaveCompact::Var>{{"i32", "f_fkey"}, {"i64", "_rev"}});
            rev = "_rev";
        }
        cur = std::make_unique<HipHashProbe>(std::move(cur), "d_q3_b_ht", "f_fkey", "b_slots", 2, "_bs");
        cur = std::make_unique<HipSlotAtomicAdd>(std::move(cur), "d_q3_b_ht", 2, "_bs", 1, rev);
        // ITEMS=4: the measured best on MI355X (ITEMS=8 and block 128 are
        // slower; per-row emission and item staging without the queue measure
        // the same as the queue with all columns streamed)
        auto& ph = appendPhase(plan, "Q3_probe_f", std::move(cur), optInt(o, "block", 256), optInt(o, "items", 4),
                               {4, 8});
        ph.scalarParams = {"q3_date"};
        ph.staged = compact || optInt(o, "staged", 0) != 0;
        ph.prefetch = optInt(o, "prefetch", 0) != 0;
        if (compact) ph.launchBound = 256;  // sizes the per-wave queues
    }
    {
        auto scan = std::make_unique<HipHashTableScan>("d_q3_b_ht", "b_slots", 2, "ht");
        auto sel = std::make_unique<HipSelection>(std::move(scan), "(i64)ht_w1 > 0");
        auto top = std::make_unique<HipTopK>(std::move(sel), "d_q3_top", 10, "ht_w1",
                                             "((u64)(0x7FFFFFFFu - ht_payload) << 32) | (u64)(0xFFFFFFFFu - ht_key)");
        top->addColumn("b_key", 1, 0, 32, 0xFFFFFFFFll);
        top->addColumn("revenue", 0, 0, 64, 0);
        top->addColumn("b_date", 1, 32, 32, 0x7FFFFFFFll);
        auto merge = top->makeMerge();
        // ITEMS slots per thread in flight (the grid is capped by the
        // candidate lists, so per-thread depth carries the bandwidth)
        auto& ph = appendPhase(plan, "Q3_topk", std::move(top), 256, optInt(o, "items_topk", 16), {4, 8, 16});
        ph.launchBound = 256;
        // one block merges every candidate list: 1024 threads keep each
        // thread's serial insertion chain short
        auto& pm = appendPhase(plan, "Q3_topk_merge", std::move(merge), optInt(o, "merge_block", 1024), 1, {1});
        pm.launchBound = 1024;
    }
    plan.requiredScalars = {"q3_seg", "q3_date"};
    plan.requiredSymbols = {"a_bits", "b_bits"};
    if (!prepass) plan.requiredSymbols.push_back("b_slots");
    plan.chunkable = true;
    return plan;
}

// GROUPBY — per-key sum/count/min/max of g_val (cfg 2; the reference's
// per-key aggregation is Q18, metal_plan_builder.cpp:1529-1545, direct
// addressed; here a hash table). Options: variant=lds|global|part, lds_slots,
// filter=1 adds g_val < gb_val_cut (selectivity control). variant=part is the
// partitioned form without global atomics (hot keys in LDS, cold rows packed
// by hash partition, one LDS-built table slice per partition; see
// HipPartitionCount); it needs gb_slots >= 4096, gb_slots / 4096 partitions
// within PartitionConfig::maxParts() and the symbol gbp_rows (rows of G).
void gbCountPhases(HipQueryPlan& plan, bool filter);
void gbCountFlagsPhase(HipQueryPlan& plan);

HipQueryPlan buildGroupBy(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "GROUPBY";
    const bool filter = optInt(o, "filter", 0) != 0;
    const bool prepass = opt(o, "sizing", "given") == "prepass";
    const std::string variant = opt(o, "variant", "lds");
    const bool part = variant == "part";
    // the partitioned form fuses the pre-pass's key flags into its first pass
    // (fused=0: the separate byte-flag pass first, as for the other variants)
    const bool fused = part && prepass && optInt(o, "fused", 1) != 0;
    const HipQueryPlan::SizingRule sizing{"GBC_count", "gb_slots", "d_gb_distinct", 0, 2,
                            part ? static_cast<uint64_t>(PartitionConfig::kPartSlots) : 64u, true};
    if (prepass && !fused) {
        gbCountPhases(plan, filter);
        plan.sizingRules.push_back(sizing);
    }
    auto scan = makeScan("G", {"g_key", "g_val"});
    std::unique_ptr<HipOperator> cur = std::move(scan);
    if (filter) cur = std::make_unique<HipSelection>(std::move(cur), "g_val < gb_val_cut");
    const int ldsSlots = optInt(o, "lds_slots", 2048);
    if (part) {
        PartitionConfig pc;
        pc.prefix = "d_gbp";
        pc.table = "d_gb_ht";
        pc.slots = "gb_slots";
        pc.rowsSym = "gbp_rows";
        pc.scanTable = "G";
        pc.ldsSlots = ldsSlots;
        pc.window = optInt(o, "window", 2);
        pc.narrow = optInt(o, "rec8", 0) != 0;  // 8-byte cold records, checked per value
        if (fused) {
            // the table size (and so its partition count) is known only after
            // the count: count at the most partitions the table could need —
            // 2 x min(keys in range, rows) slots — within the LDS counters
            pc.countParts = "min(" + std::to_string(pc.maxParts()) +
                            ", max(1, pow2(2 * min(gb_key_bits, gbp_rows)) / " +
                            std::to_string(PartitionConfig::kPartSlots) + "))";
            pc.flags = "d_gb_key_flags";
            pc.flagMin = "gb_key_min";
            pc.flagBits = "gb_key_bits";
        }
        const int block = optInt(o, "block", 1024);
        auto cnt = std::make_unique<HipPartitionCount>(std::move(cur), pc, "g_key", "g_val");
        auto& p1 = appendPhase(plan, fused ? "GBP_count_flags" : "GBP_count", std::move(cnt), block,
                               optInt(o, "items", 4), {4});
        p1.launchBound = 1024;
        if (filter) p1.scalarParams = {"gb_val_cut"};  // (phase references die at the next append)
        if (fused) {
            gbCountFlagsPhase(plan);
            plan.sizingRules.push_back(sizing);
        }
        auto& p2 = appendPhase(plan, "GBP_offsets", std::make_unique<HipPartitionOffsets>(pc), 256, 1, {1});
        p2.launchBound = 256;
        auto& p3 = appendPhase(plan, "GBP_bases", std::make_unique<HipPartitionBases>(pc), 1024, 1, {1});
        p3.launchBound = 1024;
        // pass 4 re-reads the rows (same scan, same filter)
        auto scan2 = makeScan("G", {"g_key", "g_val"});
        std::unique_ptr<HipOperator> cur2 = std::move(scan2);
        if (filter) cur2 = std::make_unique<HipSelection>(std::move(cur2), "g_val < gb_val_cut");
        auto sc = std::make_unique<HipPartitionScatter>(std::move(cur2), pc, "g_key", "g_val");
        auto& p4 = appendPhase(plan, "GBP_scatter", std::move(sc), block, optInt(o, "items", 4), {4});
        p4.launchBound = 1024;
        if (filter) p4.scalarParams = {"gb_val_cut"};
        auto& p5 = appendPhase(plan, "GBP_agg", std::make_unique<HipPartitionAgg>(pc), 1024, 4, {4});
        p5.launchBound = 1024;
        if (filter) plan.requiredScalars = {"gb_val_cut"};
        if (prepass) {
            plan.requiredScalars.push_back("gb_key_min");
            plan.requiredSymbols = {"gb_key_bits", "gbp_rows"};
        } else {
            plan.requiredSymbols = {"gb_slots", "gbp_rows"};
        }
        plan.chunkable = false;  // passes 1 and 4 both need every row of G
        return plan;
    }
    auto gb = std::make_unique<HipHashGroupBy>(std::move(cur), "d_gb_ht", "g_key", "g_val", "gb_slots",
                                               variant == "global" ? HipHashGroupBy::GLOBAL : HipHashGroupBy::LDS,
                                               ldsSlots);
    gb->setWindow(optInt(o, "window", 16));
    auto& ph = appendPhase(plan, "GB_agg", std::move(gb), optInt(o, "block", 512), optInt(o, "items", 4), {4});
    ph.launchBound = 1024;
    if (filter) {
        ph.scalarParams = {"gb_val_cut"};
        plan.requiredScalars = {"gb_val_cut"};
    }
    if (prepass) {
        plan.requiredScalars.push_back("gb_key_min");
        plan.requiredSymbols = {"gb_key_bits"};
    } else {
        plan.requiredSymbols = {"gb_slots"};
    }
    plan.chunkable = !prepass;  // the pre-pass needs every row before the table exists
    return plan;
}

// GB_COUNT — exact distinct-key count (the plain form's pre-pass): a byte
// flag per key of the range, set by (key - gb_key_min) with plain stores
// after a plain read (no atomics: every writer stores the same byte, and hot
// keys of a skewed column find their flag set in cache), then the set flags
// counted (the flags are scanned as table GBFL). On 120M Zipf(1.1) keys over
// 20M this measured 0.49 ms against 1.7-4.2 ms for bitmaps set by device-scope
// atomic ORs (one bitmap, or one per XCD). Counterpart of the exact max-key
// pre-pass (codegen/execution/max_key_symbols.cpp:188-235).
HipQueryPlan buildGbCount(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "GB_COUNT";
    gbCountPhases(plan, optInt(o, "filter", 0) != 0);
    plan.requiredScalars = {"gb_key_min"};
    plan.requiredSymbols = {"gb_key_bits"};
    return plan;
}

void gbCountPhases(HipQueryPlan& plan, bool filter) {
    {
        std::unique_ptr<HipOperator> cur = filter ? makeScan("G", {"g_key", "g_val"}) : makeScan("G", {"g_key"});
        if (filter) cur = std::make_unique<HipSelection>(std::move(cur), "g_val < gb_val_cut");
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(cur), "d_gb_key_flags", "(i64)g_key - gb_key_min",
                                                     "gb_key_bits", "1", true);
        auto& ph = appendPhase(plan, "GBC_flags", std::move(fl), 256, 8);
        ph.scalarParams = {"gb_key_min"};
        if (filter) ph.scalarParams.push_back("gb_val_cut");
    }
    gbCountFlagsPhase(plan);
}

// GBC_count: the set key flags (d_gb_key_flags, scanned as table GBFL) counted.
void gbCountFlagsPhase(HipQueryPlan& plan) {
    auto scan = makeScan("GBFL", {"gbfl_flag"});
    auto agg = std::make_unique<HipKeyedAgg>(std::move(scan), "d_gb_distinct", "0", 1);
    agg->addAggregate("distinct_keys", AggKind::SumI64, "(i64)(gbfl_flag != 0)", 0);
    appendPhase(plan, "GBC_count", std::move(agg), 256, 16, {16});
    plan.bufferTables.push_back({"GBFL", "gbfl_flag", "d_gb_key_flags", "gb_key_bits"});
}

// MAXKEY — min and max of one int32 column (table=..., column=...).
HipQueryPlan buildMaxKey(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "MAXKEY";
    const std::string table = opt(o, "table", "G");
    const std::string col = opt(o, "column", "g_key");
    auto scan = std::make_unique<HipGridStrideScan>(table, "i");
    scan->addColumn(col, catalogTable(table).type(col));
    auto agg = std::make_unique<HipKeyedAgg>(std::move(scan), "d_maxkey", "0", 1);
    agg->addAggregate("min_key", AggKind::MinI64, col, 0);
    agg->addAggregate("max_key", AggKind::MaxI64, col, 0);
    agg->addAggregate("rows", AggKind::Count, "1", 0);
    appendPhase(plan, "MAXKEY_reduce", std::move(agg), 256, 8);
    return plan;
}

// Q13L — byte-exact LIKE '%w1%w2%' over S.s_text (char[25]), count matching
// rows per A key, then the histogram of those counts over all A keys
// (buildQ13Plan_byName, metal_plan_builder.cpp:1244-1299, with LIKE in place
// of NOT LIKE and any two words; negate=1 restores NOT LIKE).
HipQueryPlan buildQ13L(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q13L";
    const bool negate = optInt(o, "negate", 0) != 0;
    const int histBins = optInt(o, "hist_bins", 1024);
    {
        auto scan = makeScan("S", {"s_fkey", "s_text"});
        auto sel = std::make_unique<HipLikeFilter>(std::move(scan), "s_text", 25, "d_like_words", "like_l1", "like_l2",
                                                   negate);
        auto cnt = std::make_unique<HipAtomicCount>(std::move(sel), "d_q13_counts", "s_fkey", "a_bits", 0);
        auto& ph = appendPhase(plan, "Q13_count", std::move(cnt), 256, optInt(o, "items", 4), {4});
        ph.launchBound = 256;
        ph.scalarParams = {"like_l1", "like_l2"};
        ph.externalBuffers = {{"d_like_words", "u8"}};
    }
    {
        auto scan = makeScan("A", {"a_key"});
        auto look = std::make_unique<HipArrayLookup>(std::move(scan), "d_q13_counts", "a_key", "a_bits", "_cnt",
                                                     -1ll);
        auto hist = std::make_unique<HipAtomicCount>(std::move(look), "d_q13_hist", "_cnt", std::to_string(histBins),
                                                     histBins);
        hist->setResult("c_count", "custdist", true);
        appendPhase(plan, "Q13_histogram", std::move(hist), 256, 4);
    }
    plan.requiredScalars = {"like_l1", "like_l2"};
    plan.requiredSymbols = {"a_bits"};
    plan.chunkable = true;
    return plan;
}

// Q22A — anti-join: A rows whose key is referenced by no B row with
// b_date in [q22_date_lo, q22_date_hi); count and sum(a_key) per segment
// (buildQ22Plan_byName, metal_plan_builder.cpp:1304-1357).
HipQueryPlan buildQ22A(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q22A";
    {
        auto scan = makeScan("B", {"b_fkey", "b_date"});
        auto sel = std::make_unique<HipSelection>(std::move(scan),
                                                  "(i64)b_date >= q22_date_lo && (i64)b_date < q22_date_hi");
        auto bm = std::make_unique<HipBitmapBuild>(std::move(sel), "d_q22_bitmap", "b_fkey", "a_bits");
        auto& ph = appendPhase(plan, "Q22_build_bitmap", std::move(bm), 256, 8);
        ph.scalarParams = {"q22_date_lo", "q22_date_hi"};
    }
    {
        auto scan = makeScan("A", {"a_key", "a_seg"});
        auto anti = std::make_unique<HipBitmapProbe>(std::move(scan), "d_q22_bitmap", "a_key", "a_bits", true);
        auto agg = std::make_unique<HipKeyedAgg>(std::move(anti), "d_q22_aggs", "(int)a_seg", 5);
        agg->addAggregate("num_keys", AggKind::Count, "1", 0);
        agg->addAggregate("sum_keys", AggKind::SumI64, "a_key", 0);
        agg->setBucketLabels({"0", "1", "2", "3", "4"}, "a_seg");
        auto& ph = appendPhase(plan, "Q22_anti_aggregate", std::move(agg), 256, 8);
        ph.staged = optInt(o, "staged", 1) != 0;
    }
    plan.requiredScalars = {"q22_date_lo", "q22_date_hi"};
    plan.requiredSymbols = {"a_bits"};
    plan.chunkable = true;
    return plan;
}

// PIPE_AGG — one filter + hash group-by pipeline of the batched workload:
// rows with p_date in [lo, hi) and p_flag == f, grouped by p_key with
// sum/count/min/max(p_val). The LDS table's capacity is compiled
// (lds_cap) and the slots in use come from the scalar pipe_lds_bits, so the
// autotuner can vary it per pipeline without recompiling.
HipQueryPlan buildPipeAgg(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "PIPE_AGG";
    auto scan = makeScan("P", {"p_key", "p_date", "p_val", "p_flag"});
    auto sel = std::make_unique<HipSelection>(
        std::move(scan), "(i64)p_date >= pipe_date_lo && (i64)p_date < pipe_date_hi && (i64)p_flag == pipe_flag");
    auto gb = std::make_unique<HipHashGroupBy>(std::move(sel), "d_pipe_ht", "p_key", "p_val", "pipe_slots",
                                               HipHashGroupBy::LDS, optInt(o, "lds_cap", 2048));
    gb->setRuntimeLdsBits("pipe_lds_bits");
    gb->setWindow(optInt(o, "window", 4));  // measured best for the batched pipelines
    // partials=N: each block (work item) stores its LDS table as one of at
    // most min(N, rows / 16384 + 1) partial tables and PIPE_merge folds them
    // into the table — one global merge per key and merge block instead of
    // one per key and scan block
    const int partials = optInt(o, "partials", 0);
    const int cap = optInt(o, "lds_cap", 2048);
    const std::string pmax = "min(" + std::to_string(partials) + ", n_P / 16384 + 1)";
    if (partials > 0) gb->setPartials(pmax);
    auto& ph = appendPhase(plan, "PIPE_agg", std::move(gb), optInt(o, "block", 256), optInt(o, "items", 4),
                           {1, 2, 4, 8, 16});
    ph.launchBound = 1024;
    ph.scalarParams = {"pipe_date_lo", "pipe_date_hi", "pipe_flag"};
    if (partials > 0) {
        auto mg = std::make_unique<HipGroupByMerge>("d_pipe_ht", "pipe_slots", cap, pmax);
        mg->setRuntimeLdsBits("pipe_lds_bits");
        // a block folds ITEMS x block partial slots per tile (one 4096-slot
        // partial at 256 x 16); at most 16 blocks, so a key is merged into
        // the global table at most 16 times
        auto& pm = appendPhase(plan, "PIPE_merge", std::move(mg), 256, 16, {16});
        pm.launchBound = 1024;
        pm.maxThreadgroups = 16;
    }
    plan.requiredScalars = {"pipe_date_lo", "pipe_date_hi", "pipe_flag", "pipe_lds_bits"};
    plan.requiredSymbols = {"pipe_slots"};
    // partial tables are counted from zero once per run, so a chunked run
    // (one scan + merge per chunk) would fold earlier chunks' partials again
    plan.chunkable = partials == 0;
    return plan;
}

// PIPE_SEMI / PIPE_ANTI — semi-/anti-join of P against a prebuilt key bitmap
// (registered buffer d_pipe_bitmap of pipe_bits bits): count and sum(p_val).
HipQueryPlan buildPipeJoin(const PlanOptions& o, bool anti) {
    HipQueryPlan plan;
    plan.name = anti ? "PIPE_ANTI" : "PIPE_SEMI";
    auto scan = makeScan("P", {"p_key", "p_val"});
    auto probe = std::make_unique<HipBitmapProbe>(std::move(scan), "d_pipe_bitmap", "p_key", "pipe_bits", anti);
    // lds_bitmap_words=N: batched bodies test keys below 32N bits in LDS
    probe->setLdsCache(optInt(o, "lds_bitmap_words", 0));
    auto agg = std::make_unique<HipKeyedAgg>(std::move(probe), anti ? "d_pipe_anti" : "d_pipe_semi", "0", 1);
    agg->addAggregate("rows", AggKind::Count, "1", 0);
    agg->addAggregate("sum_val", AggKind::SumI64, "p_val", 0);
    auto& ph = appendPhase(plan, anti ? "PIPE_anti" : "PIPE_semi", std::move(agg), optInt(o, "block", 256),
                           optInt(o, "items", 4), {1, 2, 4, 8, 16});
    ph.launchBound = 1024;
    ph.staged = optInt(o, "staged", 1) != 0;
    plan.requiredSymbols = {"pipe_bits"};
    plan.chunkable = true;
    return plan;
}

}  // namespace

const ColumnType& CatalogTable::type(const std::string& col) const {
    for (const auto& c : columns)
        if (c.first == col) return c.second;
    throw std::runtime_error("catalog: table " + name + " has no column " + col);
}

const CatalogTable& catalogTable(const std::string& name) {
    for (const auto& t : catalog())
        if (t.name == name) return t;
    throw std::runtime_error("catalog: unknown table " + name);
}

std::vector<std::string> catalogTableNames() {
    std::vector<std::string> n;
    for (const auto& t : catalog()) n.push_back(t.name);
    return n;
}

PlanOptions parsePlanOptions(const std::string& text) {
    PlanOptions o;
    std::stringstream ss(text);
    std::string item;
    while (std::getline(ss, item, ',')) {
        if (item.empty()) continue;
        auto eq = item.find('=');
        if (eq == std::string::npos) throw std::runtime_error("plan options: expected key=value, got '" + item + "'");
        o[item.substr(0, eq)] = item.substr(eq + 1);
    }
    return o;
}

std::vector<std::string> knownPlanNames() {
    std::vector<std::string> v = {"Q1",      "Q6",       "Q3",   "Q3_SIZE",  "GROUPBY",   "GB_COUNT", "MAXKEY",
                                  "Q13L",    "Q22A",     "PIPE_AGG", "PIPE_SEMI", "PIPE_ANTI"};
    for (const auto& n : tpchPlanNames()) v.push_back(n);
    return v;
}

std::optional<HipQueryPlan> buildHipPlan(const std::string& name, const PlanOptions& opts) {
    if (auto tp = buildTpchPlan(name, opts)) return tp;
    if (name == "Q1") return buildQ1(opts);
    if (name == "Q6") return buildQ6(opts);
    if (name == "Q3") return buildQ3(opts);
    if (name == "Q3_SIZE") return buildQ3Size(opts);
    if (name == "GROUPBY") return buildGroupBy(opts);
    if (name == "GB_COUNT") return buildGbCount(opts);
    if (name == "MAXKEY") return buildMaxKey(opts);
    if (name == "Q13L") return buildQ13L(opts);
    if (name == "Q22A") return buildQ22A(opts);
    if (name == "PIPE_AGG") return buildPipeAgg(opts);
    if (name == "PIPE_SEMI") return buildPipeJoin(opts, false);
    if (name == "PIPE_ANTI") return buildPipeJoin(opts, true);
    return std::nullopt;
}

std::string HipQueryPlan::describe() const {
    std::string out = "Plan " + name + " (" + std::to_string(phases.size()) + " phase(s))\n";
    for (const auto& p : phases) {
        out += "  Phase " + p.name + " [block " + std::to_string(p.threadgroupSize) + ", items " +
               std::to_string(p.items) + (p.singleThread ? ", single" : "") + "]\n";
        std::string tree;
        if (p.root) p.root->describeTree(tree, 2);
        out += tree;
    }
    return out;
}

HipCodegen generateFromPlan(const HipQueryPlan& plan, bool rangeMode) {
    HipCodegen cg;
    cg.setRangeMode(rangeMode);
    for (const auto& h : plan.helpers) cg.addHelper(h);
    for (const auto& phase : plan.phases) {
        cg.beginPhase(phase.name);
        cg.setPhaseThreadgroupSize(phase.threadgroupSize);
        cg.setPhaseSingleThread(phase.singleThread);
        cg.setPhaseItemsVariants(phase.itemsVariants);
        cg.setPhaseItems(phase.items);
        cg.setPhaseLaunchBound(phase.launchBound);
        cg.setPhaseStaged(phase.staged);
        cg.setPhasePrefetch(phase.prefetch);
        if (phase.maxThreadgroups) cg.setPhaseMaxThreadgroups(phase.maxThreadgroups);
        for (const auto& sp : phase.scalarParams) cg.addScalarParam(sp);
        for (const auto& eb : phase.externalBuffers)
            cg.addBufferParam(eb.first, eb.second, 1, "", InitKind::None, {}, true, true);
        if (phase.root) phase.root->produce(cg, []() {});
        cg.endPhase();
    }
    return cg;
}

}  // namespace codegen
}  // namespace gdb

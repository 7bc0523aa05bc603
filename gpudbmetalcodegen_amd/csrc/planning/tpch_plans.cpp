#include "tpch_plans.h"

#include "plan_util.h"

#include <stdexcept>

namespace gdb {
namespace codegen {

namespace {

// Money and rates are exact int64 hundredths (cents; 0.06 -> 6), so
// price * (100 - discount) carries 4 decimal digits and the charge 6.
constexpr const char* kRevenue = "l_extendedprice * (100 - l_discount)";

// ---------------------------------------------------------------------------
// Q1 — pricing summary (sql/q1.sql; the bucket encoding and slot list of
// metal_plan_builder.cpp:335-349, in exact integers).
// ---------------------------------------------------------------------------
HipQueryPlan tpchQ1(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q1";
    auto scan = makeScan("lineitem", {"l_quantity", "l_extendedprice", "l_discount", "l_tax", "l_returnflag",
                                      "l_linestatus", "l_shipdate"});
    auto sel = std::make_unique<HipSelection>(std::move(scan), "(i64)l_shipdate <= q1_cutoff");
    const std::string bucket =
        "((l_returnflag == 'A' ? 0 : (l_returnflag == 'N' ? 2 : 4)) + (l_linestatus == 'F' ? 0 : 1))";
    auto agg = std::make_unique<HipKeyedAgg>(std::move(sel), "d_q1_aggs", bucket, 6);
    agg->addAggregate("sum_qty", AggKind::SumI64, "l_quantity", 2);
    agg->addAggregate("sum_base_price", AggKind::SumI64, "l_extendedprice", 2);
    agg->addAggregate("sum_disc_price", AggKind::SumI64, kRevenue, 4);
    agg->addAggregate("sum_charge", AggKind::SumI128, std::string(kRevenue) + " * (100 + l_tax)", 6);
    agg->addAggregate("sum_disc", AggKind::SumI64, "l_discount", 2);
    agg->addAggregate("count_order", AggKind::Count, "1", 0);
    agg->setBucketLabels({"A|F", "A|O", "N|F", "N|O", "R|F", "R|O"}, "l_returnflag,l_linestatus");
    agg->addAverage("avg_qty", "sum_qty", "count_order");
    agg->addAverage("avg_price", "sum_base_price", "count_order");
    agg->addAverage("avg_disc", "sum_disc", "count_order");
    auto& ph = appendPhase(plan, "Q1_reduce", std::move(agg), optInt(o, "block", 256), optInt(o, "items", 2),
                           {2, 4});
    ph.launchBound = 256;
    ph.scalarParams = {"q1_cutoff"};
    plan.requiredScalars = {"q1_cutoff"};
    plan.chunkable = true;
    return plan;
}

// Q6 — forecasting revenue change (sql/q6.sql; metal_plan_builder.cpp:253-303).
HipQueryPlan tpchQ6(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q6";
    auto scan = makeScan("lineitem", {"l_quantity", "l_extendedprice", "l_discount", "l_shipdate"});
    auto sel = std::make_unique<HipSelection>(
        std::move(scan),
        "(i64)l_shipdate >= q6_date_lo && (i64)l_shipdate < q6_date_hi && l_discount >= q6_disc_lo && "
        "l_discount <= q6_disc_hi && l_quantity < q6_qty_hi");
    auto agg = std::make_unique<HipKeyedAgg>(std::move(sel), "d_q6_sum", "0", 1);
    agg->addAggregate("revenue", AggKind::SumI64, "l_extendedprice * l_discount", 4);
    agg->addAggregate("count", AggKind::Count, "1", 0);
    auto& ph = appendPhase(plan, "Q6_reduce", std::move(agg), optInt(o, "block", 256), optInt(o, "items", 2), {2, 4});
    ph.launchBound = 256;
    ph.scalarParams = {"q6_date_lo", "q6_date_hi", "q6_disc_lo", "q6_disc_hi", "q6_qty_hi"};
    plan.requiredScalars = ph.scalarParams;
    plan.chunkable = true;
    return plan;
}

// Q3 — shipping priority (sql/q3.sql; buildQ3Plan_byName,
// metal_plan_builder.cpp:1170-1239, with a device hash join): customer
// segment bitmap, hash table of qualifying orders keyed by o_orderkey with
// o_orderdate as payload, lineitem probe summing revenue in the slot, top-10
// by (revenue desc, o_orderdate asc, l_orderkey asc). o_shippriority is
// attached by the caller for the 10 result keys.
HipQueryPlan tpchQ3(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q3";
    {
        auto scan = makeScan("customer", {"c_custkey", "c_mktsegment"});
        auto sel = std::make_unique<HipSelection>(std::move(scan), "(i64)c_mktsegment == q3_seg");
        auto bm = std::make_unique<HipBitmapBuild>(std::move(sel), "d_q3_c_bitmap", "c_custkey", "c_bits");
        auto& ph = appendPhase(plan, "Q3_build_c_bitmap", std::move(bm), 256, 4);
        ph.scalarParams = {"q3_seg"};
    }
    {
        auto scan = makeScan("orders", {"o_orderkey", "o_custkey", "o_orderdate"});
        auto sel = std::make_unique<HipSelection>(std::move(scan), "(i64)o_orderdate < q3_date");
        auto probe = std::make_unique<HipBitmapProbe>(std::move(sel), "d_q3_c_bitmap", "o_custkey", "c_bits");
        auto ht = std::make_unique<HipHashBuild>(std::move(probe), "d_q3_o_ht", "o_orderkey", "o_orderdate", "o_slots", 2);
        auto bm = std::make_unique<HipBitmapBuild>(std::move(ht), "d_q3_o_bitmap", "o_orderkey", "o_bits");
        auto& ph = appendPhase(plan, "Q3_build_o_table", std::move(bm), 256, 4);
        ph.scalarParams = {"q3_date"};
        ph.staged = optInt(o, "staged", 1) != 0;
    }
    {
        auto scan = makeScan("lineitem", {"l_orderkey", "l_extendedprice", "l_discount", "l_shipdate"});
        std::unique_ptr<HipOperator> cur = std::make_unique<HipSelection>(std::move(scan), "(i64)l_shipdate > q3_date");
        cur = std::make_unique<HipBitmapProbe>(std::move(cur), "d_q3_o_bitmap", "l_orderkey", "o_bits");
        cur = std::make_unique<HipHashProbe>(std::move(cur), "d_q3_o_ht", "l_orderkey", "o_slots", 2, "_os");
        cur = std::make_unique<HipSlotAtomicAdd>(std::move(cur), "d_q3_o_ht", 2, "_os", 1, kRevenue);
        auto& ph = appendPhase(plan, "Q3_probe_l", std::move(cur), optInt(o, "block", 256), 2, {2});
        ph.scalarParams = {"q3_date"};
    }
    {
        auto scan = std::make_unique<HipHashTableScan>("d_q3_o_ht", "o_slots", 2, "ht");
        auto sel = std::make_unique<HipSelection>(std::move(scan), "(i64)ht_w1 > 0");
        auto top = std::make_unique<HipTopK>(std::move(sel), "d_q3_top", 10, "ht_w1",
                                             "((u64)(0x7FFFFFFFu - ht_payload) << 32) | (u64)(0xFFFFFFFFu - ht_key)");
        top->addColumn("l_orderkey", 1, 0, 32, 0xFFFFFFFFll);
        top->addColumn("revenue", 0, 0, 64, 0, 4);
        top->addColumn("o_orderdate", 1, 32, 32, 0x7FFFFFFFll, 0, true);
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
    plan.requiredSymbols = {"c_bits", "o_bits", "o_slots"};
    plan.chunkable = true;
    return plan;
}

// Q4 — order priority checking (sql/q4.sql): flags of the order keys with a
// late lineitem (commit < receipt; byte flags, or a bitmap with bytes=0),
// then orders of the quarter probing them, counted per priority ('1'..'5').
HipQueryPlan tpchQ4(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q4";
    {
        auto scan = makeScan("lineitem", {"l_orderkey", "l_commitdate", "l_receiptdate"});
        auto sel = std::make_unique<HipSelection>(std::move(scan), "l_commitdate < l_receiptdate");
        std::unique_ptr<HipOperator> bm;
        if (optInt(o, "bytes", 1))
            bm = std::make_unique<HipByteFlagBuild>(std::move(sel), "d_q4_late", "l_orderkey", "o_bits");
        else
            bm = std::make_unique<HipBitmapBuild>(std::move(sel), "d_q4_late", "l_orderkey", "o_bits");
        auto& ph = appendPhase(plan, "Q4_late_orders", std::move(bm), 256, optInt(o, "items", 8), {4, 8});
        ph.staged = optInt(o, "staged", 0) != 0;
    }
    {
        auto scan = makeScan("orders", {"o_orderkey", "o_orderdate", "o_orderpriority"});
        auto sel = std::make_unique<HipSelection>(std::move(scan),
                                                  "(i64)o_orderdate >= q4_date_lo && (i64)o_orderdate < q4_date_hi");
        std::unique_ptr<HipOperator> probe;
        if (optInt(o, "bytes", 1))
            probe = std::make_unique<HipByteFlagProbe>(std::move(sel), "d_q4_late", "o_orderkey", "o_bits", false);
        else
            probe = std::make_unique<HipBitmapProbe>(std::move(sel), "d_q4_late", "o_orderkey", "o_bits");
        auto agg = std::make_unique<HipKeyedAgg>(std::move(probe), "d_q4_counts", "(int)o_orderpriority - '1'", 5);
        agg->addAggregate("order_count", AggKind::Count, "1", 0);
        agg->setBucketLabels({"1-URGENT", "2-HIGH", "3-MEDIUM", "4-NOT SPECIFIED", "5-LOW"}, "o_orderpriority");
        auto& ph = appendPhase(plan, "Q4_count", std::move(agg), 256, optInt(o, "items", 8), {4, 8});
        ph.scalarParams = {"q4_date_lo", "q4_date_hi"};
    }
    plan.requiredScalars = {"q4_date_lo", "q4_date_hi"};
    plan.requiredSymbols = {"o_bits"};
    plan.chunkable = true;
    return plan;
}

// Q12 — shipping modes and order priority (sql/q12.sql): a byte per order key
// with the order's priority class (high = '1-URGENT' or '2-HIGH'), then the
// qualifying lineitems (ship mode code in {q12_mode1, q12_mode2}, the three
// date conditions) joined to orders and counted per mode and priority class.
// A mode code is the first two bytes of l_shipmode (char[2] in the catalog).
HipQueryPlan tpchQ12(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q12";
    {
        // one byte per order key: 0 no order, 1 other priority, 2 high priority
        auto scan = makeScan("orders", {"o_orderkey", "o_orderpriority"});
        auto cls = std::make_unique<HipByteFlagBuild>(
            std::move(scan), "d_q12_orders", "o_orderkey", "o_bits",
            "(o_orderpriority == '1' || o_orderpriority == '2') ? 2 : 1");
        appendPhase(plan, "Q12_orders", std::move(cls), 256, optInt(o, "items", 8), {4, 8});
    }
    {
        auto scan = makeScanLazy("lineitem", {"l_orderkey", "l_shipmode", "l_shipdate", "l_commitdate", "l_receiptdate"},
                                 {"l_shipmode"});
        auto mc = std::make_unique<HipComputeExpr>(std::move(scan), "_mc", "i64",
                                                   "gdb_code2(l_shipmode)");
        auto sel = std::make_unique<HipSelection>(
            std::move(mc),
            "(_mc == q12_mode1 || _mc == q12_mode2) && l_commitdate < l_receiptdate && l_shipdate < l_commitdate && "
            "(i64)l_receiptdate >= q12_date_lo && (i64)l_receiptdate < q12_date_hi");
        auto join = std::make_unique<HipByteFlagProbe>(std::move(sel), "d_q12_orders", "l_orderkey", "o_bits", false,
                                                       "_pcls");
        auto agg = std::make_unique<HipKeyedAgg>(std::move(join), "d_q12_counts", "(_mc == q12_mode1 ? 0 : 1)", 2);
        agg->addAggregate("high_line_count", AggKind::SumI64, "(i64)(_pcls == 2u)", 0);
        agg->addAggregate("low_line_count", AggKind::SumI64, "(i64)(_pcls != 2u)", 0);
        agg->setBucketLabels({opt(o, "mode1", "MAIL"), opt(o, "mode2", "SHIP")}, "l_shipmode");
        auto& ph = appendPhase(plan, "Q12_count", std::move(agg), 256, 4);
        ph.launchBound = 256;
        ph.scalarParams = {"q12_mode1", "q12_mode2", "q12_date_lo", "q12_date_hi"};
    }
    plan.requiredScalars = {"q12_mode1", "q12_mode2", "q12_date_lo", "q12_date_hi"};
    plan.requiredSymbols = {"o_bits"};
    plan.chunkable = true;
    return plan;
}

// Q13 — customer distribution (sql/q13.sql; buildQ13Plan_byName,
// metal_plan_builder.cpp:1244-1299): orders whose o_comment is NOT LIKE
// '%w1%w2%' counted per customer, then the histogram of those counts over all
// customers (customers without such orders count 0), ordered custdist desc,
// c_count desc. The words are the registered buffer d_like_words (w1 at 0,
// w2 at 32) with lengths like_l1, like_l2.
HipQueryPlan tpchQ13(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q13";
    const int histBins = optInt(o, "hist_bins", 1024);
    {
        auto scan = makeScan("orders", {"o_custkey", "o_comment"});
        auto sel = std::make_unique<HipLikeFilter>(std::move(scan), "o_comment", 79, "d_like_words", "like_l1", "like_l2",
                                                   true);
        auto cnt = std::make_unique<HipAtomicCount>(std::move(sel), "d_q13_counts", "o_custkey", "c_bits", 0);
        auto& ph = appendPhase(plan, "Q13_count", std::move(cnt), 256, 2, {2});
        ph.launchBound = 256;
        ph.scalarParams = {"like_l1", "like_l2"};
        ph.externalBuffers = {{"d_like_words", "u8"}};
    }
    {
        auto scan = makeScan("customer", {"c_custkey"});
        auto look = std::make_unique<HipArrayLookup>(std::move(scan), "d_q13_counts", "c_custkey", "c_bits", "_cnt",
                                                     -1ll);
        auto hist = std::make_unique<HipAtomicCount>(std::move(look), "d_q13_hist", "_cnt", std::to_string(histBins),
                                                     histBins);
        hist->setResult("c_count", "custdist", true);
        appendPhase(plan, "Q13_histogram", std::move(hist), 256, 4);
    }
    plan.requiredScalars = {"like_l1", "like_l2"};
    plan.requiredSymbols = {"c_bits"};
    plan.chunkable = true;
    return plan;
}

// Q14 — promotion effect (sql/q14.sql): a byte per part key telling whether
// p_type starts with the registered prefix (d_like_words, like_l1 bytes),
// then the month's lineitems joined to part:
// promo_revenue = 100 * sum(promo revenue) / sum(revenue), exact to 6 digits.
HipQueryPlan tpchQ14(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q14";
    {
        // one byte per part key: 0 no part, 1 other type, 2 promotional type
        auto scan = makeScanLazy("part", {"p_partkey", "p_type"}, {"p_type"});
        auto cls = std::make_unique<HipByteFlagBuild>(
            std::move(scan), "d_q14_parts", "p_partkey", "p_bits",
            "gdb_like_prefix(p_type, 25, (const char*)d_like_words, (int)like_l1) ? 2 : 1");
        auto& ph = appendPhase(plan, "Q14_parts", std::move(cls), 256, 4);
        ph.launchBound = 256;
        ph.scalarParams = {"like_l1"};
        ph.externalBuffers = {{"d_like_words", "u8"}};
    }
    {
        auto scan = makeScan("lineitem", {"l_partkey", "l_extendedprice", "l_discount", "l_shipdate"});
        auto sel = std::make_unique<HipSelection>(std::move(scan),
                                                  "(i64)l_shipdate >= q14_date_lo && (i64)l_shipdate < q14_date_hi");
        auto join = std::make_unique<HipByteFlagProbe>(std::move(sel), "d_q14_parts", "l_partkey", "p_bits", false,
                                                       "_pcls");
        auto agg = std::make_unique<HipKeyedAgg>(std::move(join), "d_q14_sums", "0", 1);
        agg->addAggregate("promo_revenue_x100", AggKind::SumI64, std::string(kRevenue) + " * 100 * (i64)(_pcls == 2u)",
                          4);
        agg->addAggregate("revenue", AggKind::SumI64, kRevenue, 4);
        agg->addAggregate("count", AggKind::Count, "1", 0);
        agg->addAverage("promo_revenue", "promo_revenue_x100", "revenue");
        auto& ph = appendPhase(plan, "Q14_sum", std::move(agg), 256, optInt(o, "items", 4));
        ph.scalarParams = {"q14_date_lo", "q14_date_hi"};
    }
    plan.requiredScalars = {"like_l1", "q14_date_lo", "q14_date_hi"};
    plan.requiredSymbols = {"p_bits"};
    plan.chunkable = true;
    return plan;
}

// Q18 — large volume customer (sql/q18.sql; buildQ18Plan_byName,
// metal_plan_builder.cpp:1529-1545): sum(l_quantity) per order key into a
// key-indexed array, then the orders whose sum exceeds q18_qty appended as
// (o_custkey, o_orderkey, o_orderdate, o_totalprice, sum) records, ordered
// o_totalprice desc, o_orderdate asc, cut to 100. c_name is attached by the
// caller for the result's customer keys.
HipQueryPlan tpchQ18(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q18";
    {
        auto scan = makeScan("lineitem", {"l_orderkey", "l_quantity"});
        auto add = std::make_unique<HipDenseAtomicAdd>(std::move(scan), "d_q18_qty", "l_orderkey", "o_bits",
                                                       "l_quantity");
        appendPhase(plan, "Q18_order_qty", std::move(add), 256, optInt(o, "items", 4), {4, 8});
    }
    {
        auto scan = makeScan("orders", {"o_orderkey", "o_custkey", "o_orderdate", "o_totalprice"});
        auto look = std::make_unique<HipArrayLookup>(std::move(scan), "d_q18_qty", "o_orderkey", "o_bits", "_qty", -1ll,
                                                     "i64");
        auto sel = std::make_unique<HipSelection>(std::move(look), "_qty > q18_qty");
        auto rows = std::make_unique<HipAppendRows>(std::move(sel), "d_q18_rows", "n_orders");
        rows->addField("c_custkey", "o_custkey");
        rows->addField("o_orderkey", "o_orderkey");
        rows->addField("o_orderdate", "o_orderdate", 0, true);
        rows->addField("o_totalprice", "o_totalprice", 2);
        rows->addField("sum(l_quantity)", "_qty", 2);
        rows->orderBy("o_totalprice", true);
        rows->orderBy("o_orderdate", false);
        rows->setLimit(optInt(o, "limit", 100));
        auto& ph = appendPhase(plan, "Q18_orders", std::move(rows), 256, 4);
        ph.scalarParams = {"q18_qty"};
    }
    plan.requiredScalars = {"q18_qty"};
    plan.requiredSymbols = {"o_bits"};
    plan.chunkable = true;  // the streamed table's phase only accumulates (associative atomics, byte flags)
    return plan;
}

// Q22 — global sales opportunity (sql/q22.sql; buildQ22Plan_byName,
// metal_plan_builder.cpp:1304-1357): the average positive balance of
// customers whose country code (first two phone digits) is in the set
// (bit c of q22_mask_lo / q22_mask_hi for codes 0-63 / 64-127) is left in a
// buffer as (sum, count); the customers with a code in the set, a balance
// above that average (balance * count > sum, exact) and no order are counted
// and their balances summed per code.
HipQueryPlan tpchQ22(const PlanOptions& o) {
    (void)o;
    HipQueryPlan plan;
    plan.name = "Q22";
    const std::string code = "(i64)(((i64)(u8)c_phone[0] - 48) * 10 + ((i64)(u8)c_phone[1] - 48))";
    const std::string inSet =
        "(_cc >= 0 && _cc < 128 && (((_cc < 64 ? (u64)q22_mask_lo >> _cc : (u64)q22_mask_hi >> (_cc - 64)) & 1ull) != 0))";
    {
        auto scan = makeScanLazy("customer", {"c_phone", "c_acctbal"}, {"c_phone"});
        auto cc = std::make_unique<HipComputeExpr>(std::move(scan), "_cc", "i64", code);
        auto sel = std::make_unique<HipSelection>(std::move(cc), "c_acctbal > 0 && " + inSet);
        auto agg = std::make_unique<HipKeyedAgg>(std::move(sel), "d_q22_avg", "0", 1);
        agg->addAggregate("sum_acctbal", AggKind::SumI64, "c_acctbal", 2);
        agg->addAggregate("count", AggKind::Count, "1", 0);
        auto& ph = appendPhase(plan, "Q22_avg_balance", std::move(agg), 256, 4);
        ph.launchBound = 256;
        ph.scalarParams = {"q22_mask_lo", "q22_mask_hi"};
    }
    {
        auto scan = makeScan("orders", {"o_custkey"});
        // o_custkey repeats (about 10 orders per customer) from every block
        // over a small key domain: byte flags written with plain stores
        auto bm = std::make_unique<HipByteFlagBuild>(std::move(scan), "d_q22_has_orders", "o_custkey", "c_bits");
        appendPhase(plan, "Q22_orders_bitmap", std::move(bm), 256, 4);
    }
    {
        auto scan = makeScanLazy("customer", {"c_custkey", "c_phone", "c_acctbal"}, {"c_phone"});
        std::vector<std::pair<std::string, int>> vars = {{"_avg_sum", 0}, {"_avg_cnt", 1}};
        auto avg = std::make_unique<HipBufferScalars>(std::move(scan), "d_q22_avg", "16", vars);
        auto cc = std::make_unique<HipComputeExpr>(std::move(avg), "_cc", "i64", code);
        auto sel = std::make_unique<HipSelection>(std::move(cc), inSet + " && c_acctbal * _avg_cnt > _avg_sum");
        auto anti = std::make_unique<HipByteFlagProbe>(std::move(sel), "d_q22_has_orders", "c_custkey", "c_bits", true);
        auto agg = std::make_unique<HipKeyedAgg>(std::move(anti), "d_q22_out", "(int)_cc", 100);
        agg->addAggregate("numcust", AggKind::Count, "1", 0);
        agg->addAggregate("totacctbal", AggKind::SumI64, "c_acctbal", 2);
        std::vector<std::string> labels;
        for (int c = 0; c < 100; ++c) labels.push_back((c < 10 ? "0" : "") + std::to_string(c));
        agg->setBucketLabels(labels, "cntrycode");
        auto& ph = appendPhase(plan, "Q22_anti", std::move(agg), 256, 4);
        ph.launchBound = 256;
        ph.scalarParams = {"q22_mask_lo", "q22_mask_hi"};
    }
    plan.requiredScalars = {"q22_mask_lo", "q22_mask_hi"};
    plan.requiredSymbols = {"c_bits"};
    plan.chunkable = true;  // the streamed table's phase only accumulates (associative atomics, byte flags)
    return plan;
}

// Q5 — local supplier volume (sql/q5.sql): byte maps carry nation keys
// (+1, 0 = absent) down the joins: region named like string constant 0 ->
// its nations -> suppliers and customers of those nations -> orders of the
// year (the customer's nation) -> lineitems whose order's and supplier's
// nations agree, revenue per nation. The caller attaches n_name and orders by
// revenue desc. String constants: registered buffer d_str_consts of 32-byte
// slots (length byte, then the bytes).
HipQueryPlan tpchQ5(const PlanOptions& o) {
    (void)o;
    HipQueryPlan plan;
    plan.name = "Q5";
    {
        auto scan = makeScan("region", {"r_regionkey", "r_name"});
        auto sel = std::make_unique<HipSelection>(std::move(scan),
                                                  "gdb_str_eq(r_name, 25, (const unsigned char*)d_str_consts)");
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(sel), "d_q5_region", "r_regionkey", "r_bits");
        auto& ph = appendPhase(plan, "Q5_region", std::move(fl), 64, 1, {1});
        ph.launchBound = 64;
        ph.externalBuffers = {{"d_str_consts", "u8"}};
    }
    {
        auto scan = makeScan("nation", {"n_nationkey", "n_regionkey"});
        auto in = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q5_region", "n_regionkey", "r_bits", false);
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(in), "d_q5_nation", "n_nationkey", "n_bits",
                                                     "n_nationkey + 1");
        auto& ph = appendPhase(plan, "Q5_nations", std::move(fl), 64, 1, {1});
        ph.launchBound = 64;
    }
    {
        auto scan = makeScan("supplier", {"s_suppkey", "s_nationkey"});
        auto in = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q5_nation", "s_nationkey", "n_bits", false,
                                                     "_sn");
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(in), "d_q5_supp", "s_suppkey", "s_bits", "_sn");
        appendPhase(plan, "Q5_suppliers", std::move(fl), 256, 4);
    }
    {
        auto scan = makeScan("customer", {"c_custkey", "c_nationkey"});
        auto in = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q5_nation", "c_nationkey", "n_bits", false,
                                                     "_cn");
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(in), "d_q5_cust", "c_custkey", "c_bits", "_cn");
        appendPhase(plan, "Q5_customers", std::move(fl), 256, 4);
    }
    {
        auto scan = makeScan("orders", {"o_orderkey", "o_custkey", "o_orderdate"});
        auto sel = std::make_unique<HipSelection>(std::move(scan),
                                                  "(i64)o_orderdate >= q5_date_lo && (i64)o_orderdate < q5_date_hi");
        auto in = std::make_unique<HipByteFlagProbe>(std::move(sel), "d_q5_cust", "o_custkey", "c_bits", false, "_on");
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(in), "d_q5_order", "o_orderkey", "o_bits", "_on");
        auto& ph = appendPhase(plan, "Q5_orders", std::move(fl), 256, 4);
        ph.scalarParams = {"q5_date_lo", "q5_date_hi"};
    }
    {
        auto scan = makeScan("lineitem", {"l_orderkey", "l_suppkey", "l_extendedprice", "l_discount"});
        auto ord = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q5_order", "l_orderkey", "o_bits", false,
                                                      "_ln");
        auto sup = std::make_unique<HipByteFlagProbe>(std::move(ord), "d_q5_supp", "l_suppkey", "s_bits", false, "_ls");
        auto same = std::make_unique<HipSelection>(std::move(sup), "_ln == _ls");
        auto agg = std::make_unique<HipKeyedAgg>(std::move(same), "d_q5_rev", "(int)_ln - 1", 256);
        agg->addAggregate("revenue", AggKind::SumI64, kRevenue, 4);
        agg->addAggregate("count", AggKind::Count, "1", 0);
        auto& ph = appendPhase(plan, "Q5_lineitem", std::move(agg), 256, optInt(o, "items", 4), {4, 8});
        ph.launchBound = 256;
    }
    plan.requiredSymbols = {"r_bits", "n_bits", "s_bits", "c_bits", "o_bits"};
    plan.requiredScalars = {"q5_date_lo", "q5_date_hi"};
    plan.chunkable = true;  // the streamed table's phase only accumulates (associative atomics, byte flags)
    return plan;
}

// Q10 — returned item reporting (sql/q10.sql): orders of the quarter store
// their customer key (+1) in an order-key-indexed array; lineitems with
// return flag R add their revenue to a customer-key-indexed array (runs of
// one order's lines combine before the atomic); customers with revenue go to
// a top-20 by (revenue desc, c_custkey asc). The caller attaches the
// customer's name, balance, nation name, address, phone and comment.
HipQueryPlan tpchQ10(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q10";
    {
        auto scan = makeScan("orders", {"o_orderkey", "o_custkey", "o_orderdate"});
        auto sel = std::make_unique<HipSelection>(std::move(scan),
                                                  "(i64)o_orderdate >= q10_date_lo && (i64)o_orderdate < q10_date_hi");
        auto st = std::make_unique<HipDenseStore>(std::move(sel), "d_q10_ocust", "o_orderkey", "o_bits", "o_custkey + 1");
        auto& ph = appendPhase(plan, "Q10_orders", std::move(st), 256, 4);
        ph.scalarParams = {"q10_date_lo", "q10_date_hi"};
    }
    {
        auto scan = makeScan("lineitem", {"l_orderkey", "l_returnflag", "l_extendedprice", "l_discount"});
        auto sel = std::make_unique<HipSelection>(std::move(scan), "(i64)l_returnflag == q10_flag");
        auto look = std::make_unique<HipArrayLookup>(std::move(sel), "d_q10_ocust", "l_orderkey", "o_bits", "_ck1", 0ll);
        auto add = std::make_unique<HipDenseAtomicAdd>(std::move(look), "d_q10_rev", "_ck1 - 1", "c_bits", kRevenue);
        auto& ph = appendPhase(plan, "Q10_revenue", std::move(add), 256, optInt(o, "items", 4), {4, 8});
        ph.scalarParams = {"q10_flag"};
    }
    {
        auto scan = makeScan("customer", {"c_custkey"});
        auto look = std::make_unique<HipArrayLookup>(std::move(scan), "d_q10_rev", "c_custkey", "c_bits", "_rev", 0ll,
                                                     "i64");
        auto top = std::make_unique<HipTopK>(std::move(look), "d_q10_top", 20, "(u64)_rev",
                                             "(u64)(0xFFFFFFFFu - (u32)c_custkey)");
        top->addColumn("c_custkey", 1, 0, 32, 0xFFFFFFFFll);
        top->addColumn("revenue", 0, 0, 64, 0, 4);
        auto merge = top->makeMerge();
        auto& ph = appendPhase(plan, "Q10_topk", std::move(top), 256, 4);
        ph.launchBound = 256;
        // K = 20 lists need more than the 128 VGPRs a 1024-thread block allows
        auto& pm = appendPhase(plan, "Q10_topk_merge", std::move(merge), 256, 1, {1});
        pm.launchBound = 256;
    }
    plan.requiredScalars = {"q10_date_lo", "q10_date_hi", "q10_flag"};
    plan.requiredSymbols = {"o_bits", "c_bits"};
    plan.chunkable = true;  // the streamed table's phase only accumulates (associative atomics, byte flags)
    return plan;
}

// Q19 — discounted revenue (sql/q19.sql): a byte per part key with the
// brand/container/size class the part satisfies (1, 2, 3 or 0), then the
// lineitems shipped by AIR / REG AIR with DELIVER IN PERSON whose quantity
// lies in their part's class range. String constants in d_str_consts slots:
// 0-2 brands, 3-6 / 7-10 / 11-14 the classes' containers, 15 the ship
// instruction; ship modes are 2-byte codes in scalars.
HipQueryPlan tpchQ19(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q19";
    auto slot = [](int k) { return "(const unsigned char*)d_str_consts + " + std::to_string(32 * k); };
    auto anyOf = [&](const std::string& col, int width, int first) {
        std::string e;
        for (int k = first; k < first + 4; ++k)
            e += std::string(e.empty() ? "" : " || ") + "gdb_str_eq(" + col + ", " + std::to_string(width) + ", " + slot(k) + ")";
        return "(" + e + ")";
    };
    {
        std::string cls;
        const char* sizes[3] = {"5", "10", "15"};
        for (int c = 0; c < 3; ++c) {
            cls += "(gdb_str_eq(p_brand, 10, " + slot(c) + ") && " + anyOf("p_container", 10, 3 + 4 * c) +
                   " && p_size >= 1 && p_size <= " + sizes[c] + ") ? " + std::to_string(c + 1) + " : ";
        }
        cls += "0";
        auto scan = makeScan("part", {"p_partkey", "p_brand", "p_container", "p_size"});
        auto cc = std::make_unique<HipComputeExpr>(std::move(scan), "_pcls", "u32", cls);
        auto sel = std::make_unique<HipSelection>(std::move(cc), "_pcls != 0u");
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(sel), "d_q19_parts", "p_partkey", "p_bits", "_pcls");
        auto& ph = appendPhase(plan, "Q19_parts", std::move(fl), 256, 4);
        ph.launchBound = 256;
        ph.externalBuffers = {{"d_str_consts", "u8"}};
    }
    {
        auto scan = makeScanLazy("lineitem", {"l_partkey", "l_quantity", "l_extendedprice", "l_discount", "l_shipmode",
                                              "l_shipinstruct"}, {"l_shipmode", "l_shipinstruct"});
        auto mc = std::make_unique<HipComputeExpr>(std::move(scan), "_mc", "i64",
                                                   "gdb_code2(l_shipmode)");
        auto sel = std::make_unique<HipSelection>(
            std::move(mc), "(_mc == q19_mode1 || _mc == q19_mode2) && gdb_str_eq(l_shipinstruct, 25, " + slot(15) + ")");
        auto cls = std::make_unique<HipByteFlagProbe>(std::move(sel), "d_q19_parts", "l_partkey", "p_bits", false,
                                                      "_pc");
        auto qty = std::make_unique<HipSelection>(
            std::move(cls),
            "(_pc == 1u && l_quantity >= 100 && l_quantity <= 1100) || (_pc == 2u && l_quantity >= 1000 && "
            "l_quantity <= 2000) || (_pc == 3u && l_quantity >= 2000 && l_quantity <= 3000)");
        auto agg = std::make_unique<HipKeyedAgg>(std::move(qty), "d_q19_sum", "0", 1);
        agg->addAggregate("revenue", AggKind::SumI64, kRevenue, 4);
        agg->addAggregate("count", AggKind::Count, "1", 0);
        auto& ph = appendPhase(plan, "Q19_lineitem", std::move(agg), 256, optInt(o, "items", 4), {4});
        ph.launchBound = 256;
        ph.scalarParams = {"q19_mode1", "q19_mode2"};
        ph.externalBuffers = {{"d_str_consts", "u8"}};
    }
    plan.requiredScalars = {"q19_mode1", "q19_mode2"};
    plan.requiredSymbols = {"p_bits"};
    plan.chunkable = true;  // the streamed table's phase only accumulates (associative atomics, byte flags)
    return plan;
}

// Byte flags of the suppliers of the nation named like string constant 0
// (two phases: nation name -> nation key flag -> supplier flag).
void nationSupplierFlags(HipQueryPlan& plan, const std::string& prefix, const std::string& flags) {
    {
        auto scan = makeScan("nation", {"n_nationkey", "n_name"});
        auto sel = std::make_unique<HipSelection>(std::move(scan),
                                                  "gdb_str_eq(n_name, 25, (const unsigned char*)d_str_consts)");
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(sel), "d_" + prefix + "_nation", "n_nationkey", "n_bits");
        auto& ph = appendPhase(plan, prefix + "_nation", std::move(fl), 64, 1, {1});
        ph.launchBound = 64;
        ph.externalBuffers = {{"d_str_consts", "u8"}};
    }
    {
        auto scan = makeScan("supplier", {"s_suppkey", "s_nationkey"});
        auto in = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_" + prefix + "_nation", "s_nationkey", "n_bits",
                                                     false);
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(in), flags, "s_suppkey", "s_bits");
        appendPhase(plan, prefix + "_suppliers", std::move(fl), 256, 4);
    }
}

// Q11 — important stock identification (sql/q11.sql): supply value per part
// over the suppliers of one nation (key-indexed sums) and its total in the
// same pass, then the parts whose value exceeds total * fraction
// (value * q11_frac_den > total, exact), ordered by value desc.
HipQueryPlan tpchQ11(const PlanOptions& o) {
    (void)o;
    HipQueryPlan plan;
    plan.name = "Q11";
    nationSupplierFlags(plan, "Q11", "d_q11_supp");
    {
        auto scan = makeScan("partsupp", {"ps_partkey", "ps_suppkey", "ps_supplycost", "ps_availqty"});
        auto in = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q11_supp", "ps_suppkey", "s_bits", false);
        auto add = std::make_unique<HipDenseAtomicAdd>(std::move(in), "d_q11_value", "ps_partkey", "p_bits",
                                                       "ps_supplycost * (i64)ps_availqty");
        auto tot = std::make_unique<HipKeyedAgg>(std::move(add), "d_q11_total", "0", 1);
        tot->addAggregate("total", AggKind::SumI64, "ps_supplycost * (i64)ps_availqty", 2);
        appendPhase(plan, "Q11_value", std::move(tot), 256, 4);
    }
    {
        auto scan = makeScan("part", {"p_partkey"});
        auto tot = std::make_unique<HipBufferScalars>(std::move(scan), "d_q11_total", "16",
                                                      std::vector<std::pair<std::string, int>>{{"_tot", 0}});
        auto look = std::make_unique<HipArrayLookup>(std::move(tot), "d_q11_value", "p_partkey", "p_bits", "_val", 0ll,
                                                     "i64");
        auto sel = std::make_unique<HipSelection>(std::move(look), "_val * q11_frac_den > _tot");
        auto rows = std::make_unique<HipAppendRows>(std::move(sel), "d_q11_rows", "n_part");
        rows->addField("ps_partkey", "p_partkey");
        rows->addField("value", "_val", 2);
        rows->orderBy("value", true);
        auto& ph = appendPhase(plan, "Q11_parts", std::move(rows), 256, 4);
        ph.scalarParams = {"q11_frac_den"};
    }
    plan.requiredScalars = {"q11_frac_den"};
    plan.requiredSymbols = {"n_bits", "s_bits", "p_bits"};
    // partsupp streams (Q11_value only accumulates: key-indexed atomic sums
    // and the total), as the reference's Tier C streams partsupp for Q11
    // (metal_plan_builder.cpp:2089-2102)
    plan.chunkable = true;
    return plan;
}

// Q15 — top supplier (sql/q15.sql): revenue per supplier over the quarter
// (key-indexed sums), its maximum, then the suppliers whose revenue equals
// it, ordered by s_suppkey. The caller attaches name, address and phone.
HipQueryPlan tpchQ15(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q15";
    {
        auto scan = makeScan("lineitem", {"l_suppkey", "l_extendedprice", "l_discount", "l_shipdate"});
        auto sel = std::make_unique<HipSelection>(std::move(scan),
                                                  "(i64)l_shipdate >= q15_date_lo && (i64)l_shipdate < q15_date_hi");
        auto add = std::make_unique<HipDenseAtomicAdd>(std::move(sel), "d_q15_rev", "l_suppkey", "s_bits", kRevenue);
        auto& ph = appendPhase(plan, "Q15_revenue", std::move(add), 256, optInt(o, "items", 4));
        ph.scalarParams = {"q15_date_lo", "q15_date_hi"};
    }
    {
        auto scan = makeScan("supplier", {"s_suppkey"});
        auto look = std::make_unique<HipArrayLookup>(std::move(scan), "d_q15_rev", "s_suppkey", "s_bits", "_rev", 0ll,
                                                     "i64");
        auto mx = std::make_unique<HipKeyedAgg>(std::move(look), "d_q15_max", "0", 1);
        mx->addAggregate("max_revenue", AggKind::MaxI64, "_rev", 4);
        appendPhase(plan, "Q15_max", std::move(mx), 256, 4);
    }
    {
        auto scan = makeScan("supplier", {"s_suppkey"});
        auto mx = std::make_unique<HipBufferScalars>(std::move(scan), "d_q15_max", "16",
                                                     std::vector<std::pair<std::string, int>>{{"_max", 0}});
        auto look = std::make_unique<HipArrayLookup>(std::move(mx), "d_q15_rev", "s_suppkey", "s_bits", "_rev", 0ll,
                                                     "i64");
        auto sel = std::make_unique<HipSelection>(std::move(look), "_rev == _max");
        auto rows = std::make_unique<HipAppendRows>(std::move(sel), "d_q15_rows", "n_supplier");
        rows->addField("s_suppkey", "s_suppkey");
        rows->addField("total_revenue", "_rev", 4);
        rows->orderBy("s_suppkey", false);
        appendPhase(plan, "Q15_top", std::move(rows), 256, 4);
    }
    plan.requiredScalars = {"q15_date_lo", "q15_date_hi"};
    plan.requiredSymbols = {"s_bits"};
    plan.chunkable = true;  // the streamed table's phase only accumulates (associative atomics, byte flags)
    return plan;
}

// Q17 — small-quantity-order revenue (sql/q17.sql): parts of one brand and
// container (string constants 0 and 1) flagged; per flagged part the sum and
// count of l_quantity (key-indexed); then the lineitems whose quantity is
// below 0.2 * the part's average (5 * qty * count < sum, exact) summed. The
// caller divides by 7 (avg_yearly).
HipQueryPlan tpchQ17(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q17";
    {
        auto scan = makeScan("part", {"p_partkey", "p_brand", "p_container"});
        auto sel = std::make_unique<HipSelection>(
            std::move(scan),
            "gdb_str_eq(p_brand, 10, (const unsigned char*)d_str_consts) && "
            "gdb_str_eq(p_container, 10, (const unsigned char*)d_str_consts + 32)");
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(sel), "d_q17_parts", "p_partkey", "p_bits");
        auto& ph = appendPhase(plan, "Q17_parts", std::move(fl), 256, 4);
        ph.launchBound = 256;
        ph.externalBuffers = {{"d_str_consts", "u8"}};
    }
    {
        auto scan = makeScan("lineitem", {"l_partkey", "l_quantity"});
        auto in = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q17_parts", "l_partkey", "p_bits", false);
        auto sum = std::make_unique<HipDenseAtomicAdd>(std::move(in), "d_q17_qsum", "l_partkey", "p_bits", "l_quantity");
        auto cnt = std::make_unique<HipDenseAtomicAdd>(std::move(sum), "d_q17_qcnt", "l_partkey", "p_bits", "1");
        appendPhase(plan, "Q17_part_qty", std::move(cnt), 256, optInt(o, "items", 4));
    }
    {
        auto scan = makeScan("lineitem", {"l_partkey", "l_quantity", "l_extendedprice"});
        auto in = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q17_parts", "l_partkey", "p_bits", false);
        auto qs = std::make_unique<HipArrayLookup>(std::move(in), "d_q17_qsum", "l_partkey", "p_bits", "_qs", -1ll, "i64");
        auto qc = std::make_unique<HipArrayLookup>(std::move(qs), "d_q17_qcnt", "l_partkey", "p_bits", "_qc", -1ll, "i64");
        auto sel = std::make_unique<HipSelection>(std::move(qc), "5 * l_quantity * _qc < _qs");
        auto agg = std::make_unique<HipKeyedAgg>(std::move(sel), "d_q17_sum", "0", 1);
        agg->addAggregate("sum_price", AggKind::SumI64, "l_extendedprice", 2);
        agg->addAggregate("count", AggKind::Count, "1", 0);
        auto& ph = appendPhase(plan, "Q17_small_qty", std::move(agg), 256, optInt(o, "items", 4));
        ph.newStreamPass = true;  // reads every part's quantity sum and count over the whole of lineitem
    }
    plan.requiredSymbols = {"p_bits"};
    // streamed in two passes over lineitem: per-part sums and counts, then
    // the rows below a fifth of their part's average
    plan.chunkable = true;
    return plan;
}

std::string strSlot(int k) { return "(const unsigned char*)d_str_consts + " + std::to_string(32 * k); }

// Byte flags of the nations of the region named like string constant k.
void regionNationFlags(HipQueryPlan& plan, const std::string& prefix, int k, const std::string& flags) {
    {
        auto scan = makeScan("region", {"r_regionkey", "r_name"});
        auto sel = std::make_unique<HipSelection>(std::move(scan), "gdb_str_eq(r_name, 25, " + strSlot(k) + ")");
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(sel), "d_" + prefix + "_region", "r_regionkey", "r_bits");
        auto& ph = appendPhase(plan, prefix + "_region", std::move(fl), 64, 1, {1});
        ph.launchBound = 64;
        ph.externalBuffers = {{"d_str_consts", "u8"}};
    }
    {
        auto scan = makeScan("nation", {"n_nationkey", "n_regionkey"});
        auto in = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_" + prefix + "_region", "n_regionkey", "r_bits",
                                                     false);
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(in), flags, "n_nationkey", "n_bits");
        auto& ph = appendPhase(plan, prefix + "_nations", std::move(fl), 64, 1, {1});
        ph.launchBound = 64;
    }
}

// Q2 — minimum cost supplier (sql/q2.sql): suppliers of the region (string
// constant 0), parts of size q2_size whose type ends with string constant 1;
// the minimum supply cost per part over those suppliers (key-indexed atomic
// minimum), then the partsupp rows at that minimum appended as (supplier,
// part). The caller attaches supplier, nation and part columns, orders by
// s_acctbal desc, n_name, s_name, p_partkey and keeps 100.
HipQueryPlan tpchQ2(const PlanOptions& o) {
    (void)o;
    HipQueryPlan plan;
    plan.name = "Q2";
    regionNationFlags(plan, "Q2", 0, "d_q2_nation");
    {
        auto scan = makeScan("supplier", {"s_suppkey", "s_nationkey"});
        auto in = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q2_nation", "s_nationkey", "n_bits", false);
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(in), "d_q2_supp", "s_suppkey", "s_bits");
        appendPhase(plan, "Q2_suppliers", std::move(fl), 256, 4);
    }
    {
        auto scan = makeScanLazy("part", {"p_partkey", "p_size", "p_type"}, {"p_type"});
        auto sel = std::make_unique<HipSelection>(
            std::move(scan), "(i64)p_size == q2_size && gdb_like_suffix(p_type, 25, (const char*)d_str_consts + 33, "
                             "(int)((const unsigned char*)d_str_consts)[32])");
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(sel), "d_q2_parts", "p_partkey", "p_bits");
        auto& ph = appendPhase(plan, "Q2_parts", std::move(fl), 256, 4);
        ph.scalarParams = {"q2_size"};
        ph.externalBuffers = {{"d_str_consts", "u8"}};
    }
    auto probes = [](std::unique_ptr<HipOperator> scan) {
        auto p = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q2_parts", "ps_partkey", "p_bits", false);
        return std::make_unique<HipByteFlagProbe>(std::move(p), "d_q2_supp", "ps_suppkey", "s_bits", false);
    };
    {
        auto scan = makeScan("partsupp", {"ps_partkey", "ps_suppkey", "ps_supplycost"});
        auto mn = std::make_unique<HipDenseAtomicMinMax>(probes(std::move(scan)), "d_q2_min", "ps_partkey", "p_bits",
                                                         "ps_supplycost", false);
        appendPhase(plan, "Q2_min_cost", std::move(mn), 256, 4);
    }
    {
        auto scan = makeScan("partsupp", {"ps_partkey", "ps_suppkey", "ps_supplycost"});
        auto look = std::make_unique<HipArrayLookup>(probes(std::move(scan)), "d_q2_min", "ps_partkey", "p_bits", "_mn",
                                                     -1ll, "i64");
        auto sel = std::make_unique<HipSelection>(std::move(look), "ps_supplycost == _mn");
        auto rows = std::make_unique<HipAppendRows>(std::move(sel), "d_q2_rows", "n_partsupp");
        rows->addField("s_suppkey", "ps_suppkey");
        rows->addField("p_partkey", "ps_partkey");
        appendPhase(plan, "Q2_rows", std::move(rows), 256, 4);
    }
    plan.requiredScalars = {"q2_size"};
    plan.requiredSymbols = {"r_bits", "n_bits", "s_bits", "p_bits"};
    return plan;
}

// Q7 — volume shipping (sql/q7.sql): nations named like string constants 0
// and 1 get codes 1 and 2; suppliers, customers and (through the customer)
// orders carry their nation's code; lineitems shipped in the two years
// (q7_date_lo .. q7_date_hi inclusive) whose supplier and order codes differ
// are summed per (supplier code, customer code, year - q7_year0). The caller
// names the groups and orders them.
HipQueryPlan tpchQ7(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q7";
    {
        auto scan = makeScan("nation", {"n_nationkey", "n_name"});
        auto cc = std::make_unique<HipComputeExpr>(std::move(scan), "_nc", "u32",
                                                   "gdb_str_eq(n_name, 25, " + strSlot(0) + ") ? 1u : (gdb_str_eq(n_name, 25, " +
                                                       strSlot(1) + ") ? 2u : 0u)");
        auto sel = std::make_unique<HipSelection>(std::move(cc), "_nc != 0u");
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(sel), "d_q7_nation", "n_nationkey", "n_bits", "_nc");
        auto& ph = appendPhase(plan, "Q7_nations", std::move(fl), 64, 1, {1});
        ph.launchBound = 64;
        ph.externalBuffers = {{"d_str_consts", "u8"}};
    }
    {
        auto scan = makeScan("supplier", {"s_suppkey", "s_nationkey"});
        auto in = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q7_nation", "s_nationkey", "n_bits", false, "_sc");
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(in), "d_q7_supp", "s_suppkey", "s_bits", "_sc");
        appendPhase(plan, "Q7_suppliers", std::move(fl), 256, 4);
    }
    {
        auto scan = makeScan("customer", {"c_custkey", "c_nationkey"});
        auto in = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q7_nation", "c_nationkey", "n_bits", false, "_cc");
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(in), "d_q7_cust", "c_custkey", "c_bits", "_cc");
        appendPhase(plan, "Q7_customers", std::move(fl), 256, 4);
    }
    {
        auto scan = makeScan("orders", {"o_orderkey", "o_custkey"});
        auto in = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q7_cust", "o_custkey", "c_bits", false, "_oc");
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(in), "d_q7_order", "o_orderkey", "o_bits", "_oc");
        appendPhase(plan, "Q7_orders", std::move(fl), 256, optInt(o, "items", 8), {4, 8});
    }
    {
        auto scan = makeScan("lineitem", {"l_orderkey", "l_suppkey", "l_shipdate", "l_extendedprice", "l_discount"});
        auto sel = std::make_unique<HipSelection>(std::move(scan),
                                                  "(i64)l_shipdate >= q7_date_lo && (i64)l_shipdate <= q7_date_hi");
        auto sup = std::make_unique<HipByteFlagProbe>(std::move(sel), "d_q7_supp", "l_suppkey", "s_bits", false, "_ls");
        auto ord = std::make_unique<HipByteFlagProbe>(std::move(sup), "d_q7_order", "l_orderkey", "o_bits", false, "_lo");
        auto diff = std::make_unique<HipSelection>(std::move(ord), "_ls != _lo");
        auto agg = std::make_unique<HipKeyedAgg>(
            std::move(diff), "d_q7_rev", "((int)_ls - 1) * 4 + ((int)_lo - 1) * 2 + (int)((i64)l_shipdate / 10000 - q7_year0)",
            8);
        agg->addAggregate("revenue", AggKind::SumI64, kRevenue, 4);
        agg->addAggregate("count", AggKind::Count, "1", 0);
        auto& ph = appendPhase(plan, "Q7_lineitem", std::move(agg), 256, optInt(o, "items", 4), {4, 8});
        ph.scalarParams = {"q7_date_lo", "q7_date_hi", "q7_year0"};
    }
    plan.requiredScalars = {"q7_date_lo", "q7_date_hi", "q7_year0"};
    plan.requiredSymbols = {"n_bits", "s_bits", "c_bits", "o_bits"};
    plan.chunkable = true;  // the streamed table's phase only accumulates (associative atomics, byte flags)
    return plan;
}

// Q8 — national market share (sql/q8.sql): customers of the region (string
// constant 0) through their orders of the two years (year code 1/2),
// suppliers marked when their nation is string constant 1, parts of type
// string constant 2; per year the volume and the marked nation's volume,
// mkt_share = marked / volume (exact to 6 digits).
HipQueryPlan tpchQ8(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q8";
    regionNationFlags(plan, "Q8", 0, "d_q8_amer");
    {
        auto scan = makeScan("nation", {"n_nationkey", "n_name"});
        auto sel = std::make_unique<HipSelection>(std::move(scan), "gdb_str_eq(n_name, 25, " + strSlot(1) + ")");
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(sel), "d_q8_mark", "n_nationkey", "n_bits");
        auto& ph = appendPhase(plan, "Q8_marked_nation", std::move(fl), 64, 1, {1});
        ph.launchBound = 64;
        ph.externalBuffers = {{"d_str_consts", "u8"}};
    }
    {
        auto scan = makeScan("supplier", {"s_suppkey", "s_nationkey"});
        auto mk = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q8_mark", "s_nationkey", "n_bits", false, "_mk",
                                                     true);
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(mk), "d_q8_supp", "s_suppkey", "s_bits", "1u + _mk");
        appendPhase(plan, "Q8_suppliers", std::move(fl), 256, 4);
    }
    {
        auto scan = makeScanLazy("part", {"p_partkey", "p_type"}, {"p_type"});
        auto sel = std::make_unique<HipSelection>(std::move(scan), "gdb_str_eq(p_type, 25, " + strSlot(2) + ")");
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(sel), "d_q8_parts", "p_partkey", "p_bits");
        auto& ph = appendPhase(plan, "Q8_parts", std::move(fl), 256, 4);
        ph.externalBuffers = {{"d_str_consts", "u8"}};
    }
    {
        auto scan = makeScan("customer", {"c_custkey", "c_nationkey"});
        auto in = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q8_amer", "c_nationkey", "n_bits", false);
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(in), "d_q8_cust", "c_custkey", "c_bits");
        appendPhase(plan, "Q8_customers", std::move(fl), 256, 4);
    }
    {
        auto scan = makeScan("orders", {"o_orderkey", "o_custkey", "o_orderdate"});
        auto sel = std::make_unique<HipSelection>(std::move(scan),
                                                  "(i64)o_orderdate >= q8_date_lo && (i64)o_orderdate <= q8_date_hi");
        auto in = std::make_unique<HipByteFlagProbe>(std::move(sel), "d_q8_cust", "o_custkey", "c_bits", false);
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(in), "d_q8_order", "o_orderkey", "o_bits",
                                                     "(i64)o_orderdate / 10000 - q8_year0 + 1");
        auto& ph = appendPhase(plan, "Q8_orders", std::move(fl), 256, 4);
        ph.scalarParams = {"q8_date_lo", "q8_date_hi", "q8_year0"};
    }
    {
        auto scan = makeScan("lineitem", {"l_partkey", "l_suppkey", "l_orderkey", "l_extendedprice", "l_discount"});
        auto part = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q8_parts", "l_partkey", "p_bits", false);
        auto ord = std::make_unique<HipByteFlagProbe>(std::move(part), "d_q8_order", "l_orderkey", "o_bits", false, "_oy");
        auto sup = std::make_unique<HipByteFlagProbe>(std::move(ord), "d_q8_supp", "l_suppkey", "s_bits", false, "_sv");
        auto agg = std::make_unique<HipKeyedAgg>(std::move(sup), "d_q8_vol", "(int)_oy - 1", 2);
        agg->addAggregate("volume", AggKind::SumI64, kRevenue, 4);
        agg->addAggregate("marked_volume", AggKind::SumI64, std::string(kRevenue) + " * (i64)(_sv == 2u)", 4);
        agg->addAggregate("count", AggKind::Count, "1", 0);
        agg->addAverage("mkt_share", "marked_volume", "volume");
        appendPhase(plan, "Q8_lineitem", std::move(agg), 256, optInt(o, "items", 4), {4, 8});
    }
    plan.requiredScalars = {"q8_date_lo", "q8_date_hi", "q8_year0"};
    plan.requiredSymbols = {"r_bits", "n_bits", "s_bits", "c_bits", "o_bits", "p_bits"};
    plan.chunkable = true;  // the streamed table's phase only accumulates (associative atomics, byte flags)
    return plan;
}

// Q21 — suppliers who kept orders waiting (sql/q21.sql). Per order the
// smallest and largest l_suppkey of all its lines and of its late lines
// (receipt > commit) reduce the EXISTS / NOT EXISTS subqueries to
// comparisons: another supplier exists when min != s or max != s; no other
// late supplier when late min == late max == s. A late line of a supplier of
// the nation (string constant 0) on an order with status q21_status passing
// both counts for its supplier; the caller attaches s_name, orders by
// numwait desc, s_name and keeps 100.
HipQueryPlan tpchQ21(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q21";
    nationSupplierFlags(plan, "Q21", "d_q21_supp");
    {
        auto scan = makeScan("orders", {"o_orderkey", "o_orderstatus"});
        auto sel = std::make_unique<HipSelection>(std::move(scan), "(i64)o_orderstatus == q21_status");
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(sel), "d_q21_orders", "o_orderkey", "o_bits");
        auto& ph = appendPhase(plan, "Q21_orders", std::move(fl), 256, 8, {8});
        ph.scalarParams = {"q21_status"};
    }
    {
        auto scan = makeScan("lineitem", {"l_orderkey", "l_suppkey", "l_receiptdate", "l_commitdate"});
        std::unique_ptr<HipOperator> cur = std::make_unique<HipDenseAtomicMinMax>(
            std::move(scan), "d_q21_min", "l_orderkey", "o_bits", "l_suppkey", false);
        cur = std::make_unique<HipDenseAtomicMinMax>(std::move(cur), "d_q21_max", "l_orderkey", "o_bits", "l_suppkey", true);
        cur = std::make_unique<HipSelection>(std::move(cur), "l_receiptdate > l_commitdate");
        cur = std::make_unique<HipDenseAtomicMinMax>(std::move(cur), "d_q21_lmin", "l_orderkey", "o_bits", "l_suppkey",
                                                     false);
        cur = std::make_unique<HipDenseAtomicMinMax>(std::move(cur), "d_q21_lmax", "l_orderkey", "o_bits", "l_suppkey",
                                                     true);
        appendPhase(plan, "Q21_order_suppliers", std::move(cur), 256, optInt(o, "items", 8), {4, 8});
    }
    {
        auto scan = makeScan("lineitem", {"l_orderkey", "l_suppkey", "l_receiptdate", "l_commitdate"});
        std::unique_ptr<HipOperator> cur = std::make_unique<HipSelection>(std::move(scan), "l_receiptdate > l_commitdate");
        cur = std::make_unique<HipByteFlagProbe>(std::move(cur), "d_q21_supp", "l_suppkey", "s_bits", false);
        cur = std::make_unique<HipByteFlagProbe>(std::move(cur), "d_q21_orders", "l_orderkey", "o_bits", false);
        for (const char* a : {"min", "max", "lmin", "lmax"})
            cur = std::make_unique<HipArrayLookup>(std::move(cur), std::string("d_q21_") + a, "l_orderkey", "o_bits",
                                                   std::string("_") + a, -1ll, "i64");
        cur = std::make_unique<HipSelection>(
            std::move(cur), "(_min != (i64)l_suppkey || _max != (i64)l_suppkey) && _lmin == (i64)l_suppkey && "
                            "_lmax == (i64)l_suppkey");
        cur = std::make_unique<HipDenseAtomicAdd>(std::move(cur), "d_q21_wait", "l_suppkey", "s_bits", "1");
        appendPhase(plan, "Q21_waiting", std::move(cur), 256, optInt(o, "items", 8), {4, 8});
    }
    {
        auto scan = makeScan("supplier", {"s_suppkey"});
        auto look = std::make_unique<HipArrayLookup>(std::move(scan), "d_q21_wait", "s_suppkey", "s_bits", "_nw", 0ll,
                                                     "i64");
        auto rows = std::make_unique<HipAppendRows>(std::move(look), "d_q21_rows", "n_supplier");
        rows->addField("s_suppkey", "s_suppkey");
        rows->addField("numwait", "_nw");
        appendPhase(plan, "Q21_suppliers_out", std::move(rows), 256, 4);
    }
    plan.requiredScalars = {"q21_status"};
    plan.requiredSymbols = {"n_bits", "s_bits", "o_bits"};
    return plan;
}

// The (partkey, suppkey) pair of a partsupp or lineitem row as one 64-bit key.
std::string pairKey(const std::string& part, const std::string& supp) {
    return "(((u64)(u32)" + part + " << 32) | (u64)(u32)" + supp + ")";
}

// Q9 — product type profit (sql/q9.sql): parts whose name contains word 1 of
// the LIKE words (LIKE '%w%': w2 empty); their partsupp rows in a 64-bit-key
// hash table (partkey, suppkey) -> supply cost; suppliers carry nation key +
// 1, orders their year - q9_year0 + 1; lineitems of those parts sum
// price * (1 - discount) - cost * quantity per (nation, year). The caller
// names the groups and orders them nation asc, year desc.
HipQueryPlan tpchQ9(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q9";
    {
        auto scan = makeScan("part", {"p_partkey", "p_name"});
        auto like = std::make_unique<HipLikeFilter>(std::move(scan), "p_name", 55, "d_like_words", "like_l1", "like_l2",
                                                    false);
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(like), "d_q9_parts", "p_partkey", "p_bits");
        auto& ph = appendPhase(plan, "Q9_parts", std::move(fl), 256, 2, {2});
        ph.launchBound = 256;
        ph.scalarParams = {"like_l1", "like_l2"};
        ph.externalBuffers = {{"d_like_words", "u8"}};
    }
    {
        auto scan = makeScan("partsupp", {"ps_partkey", "ps_suppkey", "ps_supplycost"});
        auto in = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q9_parts", "ps_partkey", "p_bits", false);
        auto ht = std::make_unique<HipHash64Build>(std::move(in), "d_q9_ps", pairKey("ps_partkey", "ps_suppkey"),
                                                   "ps_supplycost", "q9_slots");
        appendPhase(plan, "Q9_partsupp", std::move(ht), 256, 4);
    }
    {
        auto scan = makeScan("supplier", {"s_suppkey", "s_nationkey"});
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(scan), "d_q9_supp", "s_suppkey", "s_bits", "s_nationkey + 1");
        appendPhase(plan, "Q9_suppliers", std::move(fl), 256, 4);
    }
    {
        auto scan = makeScan("orders", {"o_orderkey", "o_orderdate"});
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(scan), "d_q9_year", "o_orderkey", "o_bits",
                                                     "(i64)o_orderdate / 10000 - q9_year0 + 1");
        auto& ph = appendPhase(plan, "Q9_orders", std::move(fl), 256, optInt(o, "items", 8), {4, 8});
        ph.scalarParams = {"q9_year0"};
    }
    {
        auto scan = makeScan("lineitem", {"l_partkey", "l_suppkey", "l_orderkey", "l_quantity", "l_extendedprice",
                                          "l_discount"});
        std::unique_ptr<HipOperator> cur =
            std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q9_parts", "l_partkey", "p_bits", false);
        cur = std::make_unique<HipHash64Probe>(std::move(cur), "d_q9_ps", pairKey("l_partkey", "l_suppkey"), "q9_slots",
                                               "_ps");
        cur = std::make_unique<HipByteFlagProbe>(std::move(cur), "d_q9_supp", "l_suppkey", "s_bits", false, "_sn");
        cur = std::make_unique<HipByteFlagProbe>(std::move(cur), "d_q9_year", "l_orderkey", "o_bits", false, "_oy");
        auto agg = std::make_unique<HipKeyedAgg>(std::move(cur), "d_q9_profit", "((int)_sn - 1) * 8 + ((int)_oy - 1)",
                                                 256 * 8);
        agg->addAggregate("sum_profit", AggKind::SumI64,
                          std::string(kRevenue) + " - (i64)_ps_payload * l_quantity", 4);
        agg->addAggregate("count", AggKind::Count, "1", 0);
        appendPhase(plan, "Q9_lineitem", std::move(agg), 256, optInt(o, "items", 4), {4, 8});
    }
    plan.requiredScalars = {"like_l1", "like_l2", "q9_year0"};
    plan.requiredSymbols = {"p_bits", "s_bits", "o_bits", "q9_slots"};
    plan.chunkable = true;  // the streamed table's phase only accumulates (associative atomics, byte flags)
    return plan;
}

// Q20 — potential part promotion (sql/q20.sql): parts whose name starts with
// the LIKE word; their partsupp pairs in a 64-bit-key table whose payload
// sums the year's l_quantity of the pair; pairs with lines and
// ps_availqty > 0.5 * sum (200 * availqty > sum of hundredths, exact) flag
// their supplier; the flagged suppliers of the nation (string constant 0)
// are returned. The caller attaches name and address and orders by name.
HipQueryPlan tpchQ20(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q20";
    {
        auto scan = makeScanLazy("part", {"p_partkey", "p_name"}, {"p_name"});
        auto sel = std::make_unique<HipSelection>(std::move(scan),
                                                  "gdb_like_prefix(p_name, 55, (const char*)d_like_words, (int)like_l1)");
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(sel), "d_q20_parts", "p_partkey", "p_bits");
        auto& ph = appendPhase(plan, "Q20_parts", std::move(fl), 256, 4);
        ph.scalarParams = {"like_l1"};
        ph.externalBuffers = {{"d_like_words", "u8"}};
    }
    {
        auto scan = makeScan("partsupp", {"ps_partkey", "ps_suppkey"});
        auto in = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q20_parts", "ps_partkey", "p_bits", false);
        auto ht = std::make_unique<HipHash64Build>(std::move(in), "d_q20_ps", pairKey("ps_partkey", "ps_suppkey"), "0",
                                                   "q20_slots");
        appendPhase(plan, "Q20_partsupp", std::move(ht), 256, 4);
    }
    {
        auto scan = makeScan("lineitem", {"l_partkey", "l_suppkey", "l_quantity", "l_shipdate"});
        std::unique_ptr<HipOperator> cur = std::make_unique<HipSelection>(
            std::move(scan), "(i64)l_shipdate >= q20_date_lo && (i64)l_shipdate < q20_date_hi");
        cur = std::make_unique<HipByteFlagProbe>(std::move(cur), "d_q20_parts", "l_partkey", "p_bits", false);
        cur = std::make_unique<HipHash64Probe>(std::move(cur), "d_q20_ps", pairKey("l_partkey", "l_suppkey"), "q20_slots",
                                               "_ps");
        cur = std::make_unique<HipSlotAtomicAdd>(std::move(cur), "d_q20_ps", 2, "_ps", 1, "l_quantity");
        auto& ph = appendPhase(plan, "Q20_lineitem", std::move(cur), 256, optInt(o, "items", 4), {4, 8});
        ph.scalarParams = {"q20_date_lo", "q20_date_hi"};
    }
    {
        auto scan = makeScan("partsupp", {"ps_partkey", "ps_suppkey", "ps_availqty"});
        std::unique_ptr<HipOperator> cur =
            std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q20_parts", "ps_partkey", "p_bits", false);
        cur = std::make_unique<HipHash64Probe>(std::move(cur), "d_q20_ps", pairKey("ps_partkey", "ps_suppkey"),
                                               "q20_slots", "_ps");
        cur = std::make_unique<HipSelection>(std::move(cur),
                                             "(i64)_ps_payload > 0 && 200 * (i64)ps_availqty > (i64)_ps_payload");
        cur = std::make_unique<HipByteFlagBuild>(std::move(cur), "d_q20_supp_ok", "ps_suppkey", "s_bits");
        appendPhase(plan, "Q20_excess", std::move(cur), 256, 4);
    }
    nationSupplierFlags(plan, "Q20", "d_q20_nation_supp");
    {
        auto scan = makeScan("supplier", {"s_suppkey"});
        auto a = std::make_unique<HipByteFlagProbe>(std::move(scan), "d_q20_nation_supp", "s_suppkey", "s_bits", false);
        auto b = std::make_unique<HipByteFlagProbe>(std::move(a), "d_q20_supp_ok", "s_suppkey", "s_bits", false);
        auto rows = std::make_unique<HipAppendRows>(std::move(b), "d_q20_rows", "n_supplier");
        rows->addField("s_suppkey", "s_suppkey");
        appendPhase(plan, "Q20_out", std::move(rows), 256, 4);
    }
    plan.requiredScalars = {"like_l1", "q20_date_lo", "q20_date_hi"};
    plan.requiredSymbols = {"p_bits", "s_bits", "n_bits", "q20_slots"};
    plan.chunkable = true;  // the streamed table's phase only accumulates (associative atomics, byte flags)
    return plan;
}

// Q16 — parts/supplier relationship (sql/q16.sql). Suppliers whose comment
// is LIKE '%w1%w2%' (the LIKE words) are flagged. Parts whose brand is not
// string constant 0, whose type does not start with string constant 1 and
// whose size is a set bit of q16_sizes are grouped by a 63-bit FNV-1a
// fingerprint of (p_brand, p_type, p_size) in a 64-bit-key hash table whose
// payload is the first part row to arrive (the group's representative); each
// part key stores its group's slot + 1. Partsupp rows of grouped parts whose
// supplier is not flagged insert (group, supplier) into a second table; first
// arrivals count per group (COUNT(DISTINCT ps_suppkey)). The last pass checks
// every grouped part's key bytes against its representative's (a fingerprint
// collision fails the run) and returns the representatives with their
// counts; the caller attaches brand, type and size and orders the groups.
HipQueryPlan tpchQ16(const PlanOptions& o) {
    HipQueryPlan plan;
    plan.name = "Q16";
    {
        auto scan = makeScan("supplier", {"s_suppkey", "s_comment"});
        auto like = std::make_unique<HipLikeFilter>(std::move(scan), "s_comment", 101, "d_like_words", "like_l1",
                                                    "like_l2", false);
        auto fl = std::make_unique<HipByteFlagBuild>(std::move(like), "d_q16_complaints", "s_suppkey", "s_bits");
        const int block = optInt(o, "supp_block", 64);
        auto& ph = appendPhase(plan, "Q16_complaints", std::move(fl), block, 1, {1});
        ph.launchBound = block;
        ph.scalarParams = {"like_l1", "like_l2"};
        ph.externalBuffers = {{"d_like_words", "u8"}};
    }
    {
        auto scan = optInt(o, "lazy", 0)
                        ? makeScanLazy("part", {"p_partkey", "p_brand", "p_type", "p_size"}, {"p_brand", "p_type"})
                        : makeScan("part", {"p_partkey", "p_brand", "p_type", "p_size"});
        auto sel = std::make_unique<HipSelection>(
            std::move(scan),
            "(u64)p_size < 64ull && ((q16_sizes >> p_size) & 1) && !gdb_str_eq(p_brand, 10, " + strSlot(0) +
                ") && !gdb_like_prefix(p_type, 25, (const char*)d_str_consts + 33, "
                "(int)((const unsigned char*)d_str_consts)[32])");
        auto ht = std::make_unique<HipHash64Build>(
            std::move(sel), "d_q16_groups",
            "(gdb_fnv64_u32(gdb_fnv64(gdb_fnv64(0xcbf29ce484222325ull, p_brand, 10), p_type, 25), (u32)p_size) & "
            "0x7fffffffffffffffull)",
            "i", "q16_slots", false, "_g");
        auto st = std::make_unique<HipDenseStore>(std::move(ht), "d_q16_part_group", "p_partkey", "p_bits", "_g + 1");
        auto& ph = appendPhase(plan, "Q16_groups", std::move(st), 256, optInt(o, "part_items", 2), {1, 2, 4});
        ph.launchBound = 256;
        ph.scalarParams = {"q16_sizes"};
        ph.externalBuffers = {{"d_str_consts", "u8"}};
    }
    {
        auto scan = makeScan("partsupp", {"ps_partkey", "ps_suppkey"});
        std::unique_ptr<HipOperator> cur = std::make_unique<HipArrayLookup>(
            std::move(scan), "d_q16_part_group", "ps_partkey", "p_bits", "_pg", 0ll, "u32");
        cur = std::make_unique<HipByteFlagProbe>(std::move(cur), "d_q16_complaints", "ps_suppkey", "s_bits", true);
        cur = std::make_unique<HipHash64Build>(std::move(cur), "d_q16_pairs",
                                               "(((u64)_pg << 32) | (u64)(u32)ps_suppkey)", "0", "q16_pair_slots", true);
        cur = std::make_unique<HipDenseAtomicAdd>(std::move(cur), "d_q16_count", "_pg - 1", "q16_slots", "1");
        appendPhase(plan, "Q16_supplier_pairs", std::move(cur), 256, optInt(o, "items", 4), {4, 8});
    }
    {
        auto scan = makeScanLazy("part", {"p_partkey", "p_brand", "p_type", "p_size"}, {"p_brand", "p_type"});
        std::unique_ptr<HipOperator> cur = std::make_unique<HipArrayLookup>(
            std::move(scan), "d_q16_part_group", "p_partkey", "p_bits", "_pg", 0ll, "u32");
        cur = std::make_unique<HipArrayLookup>(std::move(cur), "d_q16_groups", "(_pg - 1) * 2 + 1", "q16_slots * 2",
                                               "_rep", -1ll, "i64");
        const std::string R = "(u64)_rep";
        cur = std::make_unique<HipRowCheck>(
            std::move(cur), "q16_fp",
            "gdb_bytes_eq(p_brand, (const char*)" + tableDataName("part", "p_brand") + " + " + R +
                " * 10, 10) && gdb_bytes_eq(p_type, (const char*)" + tableDataName("part", "p_type") + " + " + R +
                " * 25, 25) && " + tableDataName("part", "p_size") + "[" + R + "] == p_size",
            "two (p_brand, p_type, p_size) groups share a fingerprint");
        cur = std::make_unique<HipSelection>(std::move(cur), "(u64)_rep == i");
        cur = std::make_unique<HipArrayLookup>(std::move(cur), "d_q16_count", "_pg - 1", "q16_slots", "_cnt", 0ll,
                                               "i64");
        auto rows = std::make_unique<HipAppendRows>(std::move(cur), "d_q16_rows", "n_part");
        rows->addField("p_partkey", "p_partkey");
        rows->addField("supplier_cnt", "_cnt");
        auto& ph = appendPhase(plan, "Q16_out", std::move(rows), 256, optInt(o, "part_items", 2), {1, 2, 4});
        ph.launchBound = 256;
    }
    plan.requiredScalars = {"like_l1", "like_l2", "q16_sizes"};
    plan.requiredSymbols = {"s_bits", "p_bits", "q16_slots", "q16_pair_slots"};
    return plan;
}

bool tpchSchema(const PlanOptions& o) { return opt(o, "schema", "") == "tpch"; }

}  // namespace

std::vector<std::string> tpchPlanNames() {
    return {"Q2",  "Q4",  "Q5",  "Q7",  "Q8",  "Q9",  "Q10", "Q11", "Q12",
            "Q13", "Q14", "Q15", "Q16", "Q17", "Q18", "Q19", "Q20", "Q21", "Q22"};
}

std::optional<HipQueryPlan> buildTpchPlan(const std::string& name, const PlanOptions& opts) {
    const std::string schema = opt(opts, "schema", "");
    if (!schema.empty() && schema != "tpch") throw std::runtime_error("unknown schema " + schema);
    if (name == "Q1" && tpchSchema(opts)) return tpchQ1(opts);
    if (name == "Q3" && tpchSchema(opts)) return tpchQ3(opts);
    if (name == "Q6" && tpchSchema(opts)) return tpchQ6(opts);
    if (name == "Q4") return tpchQ4(opts);
    if (name == "Q5") return tpchQ5(opts);
    if (name == "Q10") return tpchQ10(opts);
    if (name == "Q19") return tpchQ19(opts);
    if (name == "Q11") return tpchQ11(opts);
    if (name == "Q2") return tpchQ2(opts);
    if (name == "Q21") return tpchQ21(opts);
    if (name == "Q9") return tpchQ9(opts);
    if (name == "Q20") return tpchQ20(opts);
    if (name == "Q7") return tpchQ7(opts);
    if (name == "Q8") return tpchQ8(opts);
    if (name == "Q15") return tpchQ15(opts);
    if (name == "Q16") return tpchQ16(opts);
    if (name == "Q17") return tpchQ17(opts);
    if (name == "Q12") return tpchQ12(opts);
    if (name == "Q13") return tpchQ13(opts);
    if (name == "Q14") return tpchQ14(opts);
    if (name == "Q18") return tpchQ18(opts);
    if (name == "Q22") return tpchQ22(opts);
    return std::nullopt;
}

}  // namespace codegen
}  // namespace gdb

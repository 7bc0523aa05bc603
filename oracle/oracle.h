/*
 * oracle.h — CPU restatement of the reference's plan semantics, in exact
 * integers. TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it as the checker and as the timed CPU
 * baseline. The engine never links, loads or calls it.
 *
 * Each function cites the reference code whose semantics it restates (paths
 * relative to the reference repository root). The reference has no CPU
 * executor; its arithmetic runs only as Metal kernels. Where this build
 * deliberately differs (exact fixed-point instead of FP32-then-truncate,
 * int128 charge, hash join instead of direct addressing, signed 64-bit
 * min/max, total-order top-k) the oracle follows the build's stated
 * semantics.
 *
 * Every function has a single-threaded form (the bit-exact checker) and, where
 * it is timed as the CPU baseline, a multi-threaded form that splits rows over
 * `threads` threads with per-thread partials merged at the end (the host
 * pattern of codegen/execution/max_key_symbols.cpp:43-70).
 */
#ifndef GDB_ORACLE_H
#define GDB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Q1: filter + group-by (flag, status) (metal_plan_builder.cpp:309-354) ----
 * Row qualifies when date <= cutoff. Bucket = (flag 'A'->0, 'N'->2, other->4)
 * + (status 'F'->0, other->1) (metal_plan_builder.cpp:335).
 * Output per bucket b (6 buckets), 8 int64 words:
 *   [sum_qty, sum_base_price, sum_disc_price, charge_lo, charge_hi, sum_disc, count, 0]
 * where charge = sum(price*(100-disc)*(100+tax)) as a two's-complement int128
 * (charge_lo = low 64 bits, charge_hi = high 64 bits). */
void oracle_q1(uint64_t n, const int32_t* qty, const int64_t* price, const int32_t* disc, const int32_t* tax,
               const uint8_t* flag, const uint8_t* status, const int32_t* date, int64_t cutoff, int64_t* out48);
void oracle_q1_mt(uint64_t n, const int32_t* qty, const int64_t* price, const int32_t* disc, const int32_t* tax,
                  const uint8_t* flag, const uint8_t* status, const int32_t* date, int64_t cutoff, int64_t* out48,
                  int threads);

/* ---- Q6: scalar revenue (metal_plan_builder.cpp:253-303) ----
 * Rows with date_lo <= date < date_hi, disc_lo <= disc <= disc_hi, qty < qty_hi:
 * out[0] = sum(price*disc), out[1] = count. */
void oracle_q6(uint64_t n, const int32_t* qty, const int64_t* price, const int32_t* disc, const int32_t* date,
               int64_t date_lo, int64_t date_hi, int64_t disc_lo, int64_t disc_hi, int64_t qty_hi, int64_t* out2);

/* ---- Q3 analog: A-semi-join, B hash join, per-B-key revenue, top-k ----
 * (buildQ3Plan_byName, metal_plan_builder.cpp:1170-1239; top-10 ordering
 * codegen_main.cpp:1211-1248 extended to a total order.)
 *   Aset = { a_key : a_seg == seg }
 *   Bq   = { b_key -> b_date : b_date < d and b_fkey in Aset } (b_key unique)
 *   rev(k) = sum over F rows with f_date > d and f_fkey == k in Bq of
 *            f_price * (100 - f_disc)
 *   groups = { k in Bq : rev(k) > 0 }
 * oracle_q3 writes every group in ascending key order (key, rev, b_date) up
 * to `cap` entries and returns the number of groups; oracle_q3_topk writes
 * the first k groups under (rev desc, b_date asc, b_key asc). Keys must be
 * in [0, max_key]. `threads` > 1 parallelises the F probe. */
int64_t oracle_q3(uint64_t n_a, const int32_t* a_key, const uint8_t* a_seg, uint64_t n_b, const int32_t* b_key,
                  const int32_t* b_fkey, const int32_t* b_date, uint64_t n_f, const int32_t* f_fkey,
                  const int64_t* f_price, const int32_t* f_disc, const int32_t* f_date, int64_t seg, int64_t d,
                  int64_t* out_key, int64_t* out_rev, int64_t* out_date, int64_t cap, int threads);
int64_t oracle_q3_topk(uint64_t n_a, const int32_t* a_key, const uint8_t* a_seg, uint64_t n_b, const int32_t* b_key,
                       const int32_t* b_fkey, const int32_t* b_date, uint64_t n_f, const int32_t* f_fkey,
                       const int64_t* f_price, const int32_t* f_disc, const int32_t* f_date, int64_t seg, int64_t d,
                       int k, int64_t* out_key, int64_t* out_rev, int64_t* out_date, int threads);

/* ---- Hash group-by: per key count, sum, min, max of an int64 value ----
 * (per-key aggregation of Q18, metal_plan_builder.cpp:1529-1545, and the
 * KeyedAgg min/max fallback, metal_operators.cpp:466-494, with signed 64-bit
 * min/max). Rows qualify when val < val_cut (pass INT64_MAX for no filter).
 * Writes groups in ascending key order, up to cap; returns the group count. */
int64_t oracle_groupby(uint64_t n, const int32_t* key, const int64_t* val, int64_t val_cut, int64_t* out_key,
                       int64_t* out_cnt, int64_t* out_sum, int64_t* out_min, int64_t* out_max, int64_t cap,
                       int threads);
/* Exact number of distinct qualifying keys (the plain form's pre-pass). */
int64_t oracle_distinct(uint64_t n, const int32_t* key, const int64_t* val, int64_t val_cut);

/* ---- SQL LIKE '%w1%w2%' on a fixed-width NUL-padded field ----
 * (q13_comment_match, metal_plan_builder.cpp:1249-1266, for any two words):
 * the value ends at the first NUL or at width; it matches when w1 occurs and
 * w2 occurs at or after the end of the leftmost w1. */
int oracle_like2(const char* s, int width, const char* w1, int l1, const char* w2, int l2);

/* ---- Q13 analog: count matching S rows per A key, histogram of counts ----
 * (buildQ13Plan_byName, metal_plan_builder.cpp:1244-1299; ordering
 * codegen_main.cpp:1250-1277). counts(k) = #{S rows: s_fkey == k and LIKE}
 * (negated when negate != 0); hist[c] = #{A rows with counts(a_key) == c}
 * for c < bins. Writes (c_count, custdist) rows ordered custdist desc,
 * c_count desc; returns the row count. */
int64_t oracle_q13(uint64_t n_s, const int32_t* s_fkey, const char* s_text, int width, uint64_t n_a,
                   const int32_t* a_key, int64_t max_key, const char* w1, int l1, const char* w2, int l2, int negate,
                   int bins, int64_t* out_count, int64_t* out_dist, int threads);

/* ---- Q22 analog: anti-join (buildQ22Plan_byName, metal_plan_builder.cpp:1304-1357) ----
 * Bset = { b_fkey : date_lo <= b_date < date_hi }; for A rows with a_key not
 * in Bset: per a_seg (0..4) count and sum(a_key). out10 = [cnt0..cnt4, sum0..sum4]. */
void oracle_q22(uint64_t n_b, const int32_t* b_fkey, const int32_t* b_date, uint64_t n_a, const int32_t* a_key,
                const uint8_t* a_seg, int64_t max_key, int64_t date_lo, int64_t date_hi, int64_t* out10);

/* ---- Group-count estimator MLP (16 -> 64 -> 64 -> 1, ReLU), double precision ----
 * No counterpart in the reference. Parameter layout (5313 floats):
 * W1[64][16] @0, b1 @1024, W2[64][64] @1088, b2 @5184, W3[64] @5248, b3 @5312.
 * oracle_mlp_forward: y[s] for s < n.
 * oracle_mlp_grad: gradient of L = scale * sum_s (y_s - t_s)^2 w.r.t. every
 * parameter (grad[5313]); returns L. */
void oracle_mlp_forward(const float* params, const float* X, int n, double* y);
double oracle_mlp_grad(const float* params, const float* X, const float* T, int n, double scale, double* grad);

#ifdef __cplusplus
}
#endif

#endif

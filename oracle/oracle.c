/*
 * oracle.c — CPU oracle (test infrastructure; see oracle.h). Plain C11 +
 * pthreads, exact integer arithmetic.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef __int128 i128;

/* ------------------------------------------------------------------------ */
/* Q1 — metal_plan_builder.cpp:309-354 (bucket :335, slots :342-349)         */
/* ------------------------------------------------------------------------ */

typedef struct {
    int64_t sum_qty, sum_base, sum_disc_price, sum_disc, count;
    i128 charge;
} q1_acc;

static int q1_bucket(uint8_t flag, uint8_t status) {
    const int f = flag == 'A' ? 0 : (flag == 'N' ? 2 : 4);
    return f + (status == 'F' ? 0 : 1);
}

static void q1_range(uint64_t b, uint64_t e, const int32_t* qty, const int64_t* price, const int32_t* disc,
                     const int32_t* tax, const uint8_t* flag, const uint8_t* status, const int32_t* date,
                     int64_t cutoff, q1_acc* acc /* [6] */) {
    for (uint64_t i = b; i < e; ++i) {
        if ((int64_t)date[i] > cutoff) continue;
        q1_acc* a = &acc[q1_bucket(flag[i], status[i])];
        const int64_t dp = price[i] * (int64_t)(100 - disc[i]);
        a->sum_qty += qty[i];
        a->sum_base += price[i];
        a->sum_disc_price += dp;
        a->charge += (i128)dp * (i128)(100 + tax[i]);
        a->sum_disc += disc[i];
        a->count += 1;
    }
}

static void q1_emit(const q1_acc* acc, int64_t* out48) {
    for (int b = 0; b < 6; ++b) {
        const q1_acc* a = &acc[b];
        int64_t* o = out48 + 8 * b;
        o[0] = a->sum_qty;
        o[1] = a->sum_base;
        o[2] = a->sum_disc_price;
        o[3] = (int64_t)(uint64_t)(unsigned __int128)a->charge;
        o[4] = (int64_t)(uint64_t)((unsigned __int128)a->charge >> 64);
        o[5] = a->sum_disc;
        o[6] = a->count;
        o[7] = 0;
    }
}

void oracle_q1(uint64_t n, const int32_t* qty, const int64_t* price, const int32_t* disc, const int32_t* tax,
               const uint8_t* flag, const uint8_t* status, const int32_t* date, int64_t cutoff, int64_t* out48) {
    q1_acc acc[6];
    memset(acc, 0, sizeof(acc));
    q1_range(0, n, qty, price, disc, tax, flag, status, date, cutoff, acc);
    q1_emit(acc, out48);
}

typedef struct {
    uint64_t b, e;
    const int32_t *qty, *disc, *tax, *date;
    const int64_t* price;
    const uint8_t *flag, *status;
    int64_t cutoff;
    q1_acc acc[6];
} q1_task;

static void* q1_worker(void* arg) {
    q1_task* t = (q1_task*)arg;
    memset(t->acc, 0, sizeof(t->acc));
    q1_range(t->b, t->e, t->qty, t->price, t->disc, t->tax, t->flag, t->status, t->date, t->cutoff, t->acc);
    return NULL;
}

void oracle_q1_mt(uint64_t n, const int32_t* qty, const int64_t* price, const int32_t* disc, const int32_t* tax,
                  const uint8_t* flag, const uint8_t* status, const int32_t* date, int64_t cutoff, int64_t* out48,
                  int threads) {
    if (threads < 1) threads = 1;
    q1_task* tasks = (q1_task*)calloc((size_t)threads, sizeof(q1_task));
    pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    const uint64_t chunk = (n + (uint64_t)threads - 1) / (uint64_t)threads;
    for (int t = 0; t < threads; ++t) {
        q1_task* k = &tasks[t];
        k->b = (uint64_t)t * chunk < n ? (uint64_t)t * chunk : n;
        k->e = k->b + chunk < n ? k->b + chunk : n;
        k->qty = qty, k->price = price, k->disc = disc, k->tax = tax, k->flag = flag, k->status = status;
        k->date = date, k->cutoff = cutoff;
        pthread_create(&th[t], NULL, q1_worker, k);
    }
    q1_acc acc[6];
    memset(acc, 0, sizeof(acc));
    for (int t = 0; t < threads; ++t) {
        pthread_join(th[t], NULL);
        for (int b = 0; b < 6; ++b) {
            acc[b].sum_qty += tasks[t].acc[b].sum_qty;
            acc[b].sum_base += tasks[t].acc[b].sum_base;
            acc[b].sum_disc_price += tasks[t].acc[b].sum_disc_price;
            acc[b].charge += tasks[t].acc[b].charge;
            acc[b].sum_disc += tasks[t].acc[b].sum_disc;
            acc[b].count += tasks[t].acc[b].count;
        }
    }
    q1_emit(acc, out48);
    free(tasks);
    free(th);
}

/* ------------------------------------------------------------------------ */
/* Q6 — metal_plan_builder.cpp:253-303, sql/q6.sql                           */
/* ------------------------------------------------------------------------ */

void oracle_q6(uint64_t n, const int32_t* qty, const int64_t* price, const int32_t* disc, const int32_t* date,
               int64_t date_lo, int64_t date_hi, int64_t disc_lo, int64_t disc_hi, int64_t qty_hi, int64_t* out2) {
    int64_t rev = 0, cnt = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (date[i] >= date_lo && date[i] < date_hi && disc[i] >= disc_lo && disc[i] <= disc_hi && qty[i] < qty_hi) {
            rev += price[i] * (int64_t)disc[i];
            cnt += 1;
        }
    }
    out2[0] = rev;
    out2[1] = cnt;
}

/* ------------------------------------------------------------------------ */
/* Generic pthread range splitter                                            */
/* ------------------------------------------------------------------------ */

typedef void (*range_fn)(void* ctx, uint64_t b, uint64_t e, int tid);
typedef struct {
    range_fn fn;
    void* ctx;
    uint64_t b, e;
    int tid;
} range_task;

static void* range_worker(void* arg) {
    range_task* t = (range_task*)arg;
    t->fn(t->ctx, t->b, t->e, t->tid);
    return NULL;
}

static void parallel_ranges(uint64_t n, int threads, range_fn fn, void* ctx) {
    if (threads <= 1 || n < 4096) {
        fn(ctx, 0, n, 0);
        return;
    }
    range_task* tasks = (range_task*)calloc((size_t)threads, sizeof(range_task));
    pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    const uint64_t chunk = (n + (uint64_t)threads - 1) / (uint64_t)threads;
    for (int t = 0; t < threads; ++t) {
        uint64_t b = (uint64_t)t * chunk;
        if (b > n) b = n;
        uint64_t e = b + chunk < n ? b + chunk : n;
        tasks[t] = (range_task){fn, ctx, b, e, t};
        pthread_create(&th[t], NULL, range_worker, &tasks[t]);
    }
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    free(tasks);
    free(th);
}

static int64_t max_i32(uint64_t n, const int32_t* a) {
    int64_t m = -1;
    for (uint64_t i = 0; i < n; ++i)
        if (a[i] > m) m = a[i];
    return m;
}

/* ------------------------------------------------------------------------ */
/* Q3 analog — metal_plan_builder.cpp:1170-1239, codegen_main.cpp:1211-1248  */
/* ------------------------------------------------------------------------ */

typedef struct {
    const int32_t *f_fkey, *f_disc, *f_date;
    const int64_t* f_price;
    const int32_t* bdate; /* per key, -1 = not in Bq */
    int64_t maxb;
    int64_t d;
    int64_t* rev;
    int atomic;
} q3_probe_ctx;

static void q3_probe(void* vctx, uint64_t b, uint64_t e, int tid) {
    (void)tid;
    q3_probe_ctx* c = (q3_probe_ctx*)vctx;
    for (uint64_t i = b; i < e; ++i) {
        if ((int64_t)c->f_date[i] <= c->d) continue;
        const int64_t k = c->f_fkey[i];
        if (k < 0 || k > c->maxb || c->bdate[k] < 0) continue;
        const int64_t v = c->f_price[i] * (int64_t)(100 - c->f_disc[i]);
        if (c->atomic)
            __atomic_fetch_add(&c->rev[k], v, __ATOMIC_RELAXED);
        else
            c->rev[k] += v;
    }
}

/* Returns per-key arrays (bdate, rev) sized maxb+1; caller frees. */
static void q3_core(uint64_t n_a, const int32_t* a_key, const uint8_t* a_seg, uint64_t n_b, const int32_t* b_key,
                    const int32_t* b_fkey, const int32_t* b_date, uint64_t n_f, const int32_t* f_fkey,
                    const int64_t* f_price, const int32_t* f_disc, const int32_t* f_date, int64_t seg, int64_t d,
                    int threads, int32_t** bdate_out, int64_t** rev_out, int64_t* maxb_out) {
    const int64_t maxa = max_i32(n_a, a_key);
    const int64_t maxb = max_i32(n_b, b_key);
    unsigned char* aset = (unsigned char*)calloc((size_t)(maxa + 2), 1);
    for (uint64_t i = 0; i < n_a; ++i)
        if ((int64_t)a_seg[i] == seg && a_key[i] >= 0) aset[a_key[i]] = 1;
    int32_t* bdate = (int32_t*)malloc(sizeof(int32_t) * (size_t)(maxb + 2));
    for (int64_t k = 0; k <= maxb + 1; ++k) bdate[k] = -1;
    for (uint64_t i = 0; i < n_b; ++i) {
        const int64_t fk = b_fkey[i];
        if ((int64_t)b_date[i] < d && fk >= 0 && fk <= maxa && aset[fk] && b_key[i] >= 0) bdate[b_key[i]] = b_date[i];
    }
    int64_t* rev = (int64_t*)calloc((size_t)(maxb + 2), sizeof(int64_t));
    q3_probe_ctx pc = {f_fkey, f_disc, f_date, f_price, bdate, maxb, d, rev, threads > 1};
    parallel_ranges(n_f, threads, q3_probe, &pc);
    free(aset);
    *bdate_out = bdate;
    *rev_out = rev;
    *maxb_out = maxb;
}

int64_t oracle_q3(uint64_t n_a, const int32_t* a_key, const uint8_t* a_seg, uint64_t n_b, const int32_t* b_key,
                  const int32_t* b_fkey, const int32_t* b_date, uint64_t n_f, const int32_t* f_fkey,
                  const int64_t* f_price, const int32_t* f_disc, const int32_t* f_date, int64_t seg, int64_t d,
                  int64_t* out_key, int64_t* out_rev, int64_t* out_date, int64_t cap, int threads) {
    int32_t* bdate;
    int64_t* rev;
    int64_t maxb;
    q3_core(n_a, a_key, a_seg, n_b, b_key, b_fkey, b_date, n_f, f_fkey, f_price, f_disc, f_date, seg, d, threads,
            &bdate, &rev, &maxb);
    int64_t g = 0;
    for (int64_t k = 0; k <= maxb; ++k) {
        if (bdate[k] < 0 || rev[k] <= 0) continue;
        if (g < cap) {
            out_key[g] = k;
            out_rev[g] = rev[k];
            out_date[g] = bdate[k];
        }
        ++g;
    }
    free(bdate);
    free(rev);
    return g;
}

/* rank order: rev desc, date asc, key asc */
static int q3_better(int64_t r1, int64_t d1, int64_t k1, int64_t r2, int64_t d2, int64_t k2) {
    if (r1 != r2) return r1 > r2;
    if (d1 != d2) return d1 < d2;
    return k1 < k2;
}

int64_t oracle_q3_topk(uint64_t n_a, const int32_t* a_key, const uint8_t* a_seg, uint64_t n_b, const int32_t* b_key,
                       const int32_t* b_fkey, const int32_t* b_date, uint64_t n_f, const int32_t* f_fkey,
                       const int64_t* f_price, const int32_t* f_disc, const int32_t* f_date, int64_t seg, int64_t d,
                       int k, int64_t* out_key, int64_t* out_rev, int64_t* out_date, int threads) {
    int32_t* bdate;
    int64_t* rev;
    int64_t maxb;
    q3_core(n_a, a_key, a_seg, n_b, b_key, b_fkey, b_date, n_f, f_fkey, f_price, f_disc, f_date, seg, d, threads,
            &bdate, &rev, &maxb);
    int64_t m = 0; /* current list length, sorted best-first */
    for (int64_t key = 0; key <= maxb; ++key) {
        if (bdate[key] < 0 || rev[key] <= 0) continue;
        const int64_t r = rev[key], dt = bdate[key];
        if (m == k && !q3_better(r, dt, key, out_rev[k - 1], out_date[k - 1], out_key[k - 1])) continue;
        int64_t pos = m < k ? m : k - 1;
        while (pos > 0 && q3_better(r, dt, key, out_rev[pos - 1], out_date[pos - 1], out_key[pos - 1])) {
            out_rev[pos] = out_rev[pos - 1];
            out_date[pos] = out_date[pos - 1];
            out_key[pos] = out_key[pos - 1];
            --pos;
        }
        out_rev[pos] = r;
        out_date[pos] = dt;
        out_key[pos] = key;
        if (m < k) ++m;
    }
    free(bdate);
    free(rev);
    return m;
}

/* ------------------------------------------------------------------------ */
/* Hash group-by — Q18 per-key aggregation (metal_plan_builder.cpp:1529-1545) */
/* ------------------------------------------------------------------------ */

typedef struct {
    const int32_t* key;
    const int64_t* val;
    int64_t val_cut;
    int64_t kmin, range;
    int64_t *cnt, *sum, *mn, *mx; /* shared arrays (atomic updates) */
} gb_ctx;

static void gb_range(void* vctx, uint64_t b, uint64_t e, int tid) {
    (void)tid;
    gb_ctx* c = (gb_ctx*)vctx;
    for (uint64_t i = b; i < e; ++i) {
        const int64_t v = c->val[i];
        if (v >= c->val_cut) continue;
        const int64_t k = (int64_t)c->key[i] - c->kmin;
        __atomic_fetch_add(&c->cnt[k], 1, __ATOMIC_RELAXED);
        __atomic_fetch_add(&c->sum[k], v, __ATOMIC_RELAXED);
        int64_t cur = __atomic_load_n(&c->mn[k], __ATOMIC_RELAXED);
        while (v < cur && !__atomic_compare_exchange_n(&c->mn[k], &cur, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
        }
        cur = __atomic_load_n(&c->mx[k], __ATOMIC_RELAXED);
        while (v > cur && !__atomic_compare_exchange_n(&c->mx[k], &cur, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
        }
    }
}

int64_t oracle_groupby(uint64_t n, const int32_t* key, const int64_t* val, int64_t val_cut, int64_t* out_key,
                       int64_t* out_cnt, int64_t* out_sum, int64_t* out_min, int64_t* out_max, int64_t cap,
                       int threads) {
    if (n == 0) return 0;
    int64_t kmin = key[0], kmax = key[0];
    for (uint64_t i = 1; i < n; ++i) {
        if (key[i] < kmin) kmin = key[i];
        if (key[i] > kmax) kmax = key[i];
    }
    const int64_t range = kmax - kmin + 1;
    int64_t* cnt = (int64_t*)calloc((size_t)range, 8);
    int64_t* sum = (int64_t*)calloc((size_t)range, 8);
    int64_t* mn = (int64_t*)malloc((size_t)range * 8);
    int64_t* mx = (int64_t*)malloc((size_t)range * 8);
    for (int64_t k = 0; k < range; ++k) {
        mn[k] = INT64_MAX;
        mx[k] = INT64_MIN;
    }
    gb_ctx c = {key, val, val_cut, kmin, range, cnt, sum, mn, mx};
    parallel_ranges(n, threads, gb_range, &c);
    int64_t g = 0;
    for (int64_t k = 0; k < range; ++k) {
        if (!cnt[k]) continue;
        if (g < cap) {
            out_key[g] = k + kmin;
            out_cnt[g] = cnt[k];
            out_sum[g] = sum[k];
            out_min[g] = mn[k];
            out_max[g] = mx[k];
        }
        ++g;
    }
    free(cnt);
    free(sum);
    free(mn);
    free(mx);
    return g;
}

int64_t oracle_distinct(uint64_t n, const int32_t* key, const int64_t* val, int64_t val_cut) {
    if (n == 0) return 0;
    int64_t kmin = key[0], kmax = key[0];
    for (uint64_t i = 1; i < n; ++i) {
        if (key[i] < kmin) kmin = key[i];
        if (key[i] > kmax) kmax = key[i];
    }
    const int64_t range = kmax - kmin + 1;
    unsigned char* seen = (unsigned char*)calloc((size_t)range, 1);
    int64_t g = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (val && val[i] >= val_cut) continue;
        unsigned char* s = &seen[key[i] - kmin];
        if (!*s) {
            *s = 1;
            ++g;
        }
    }
    free(seen);
    return g;
}

/* ------------------------------------------------------------------------ */
/* LIKE '%w1%w2%' — q13_comment_match, metal_plan_builder.cpp:1249-1266      */
/* ------------------------------------------------------------------------ */

static int match_at(const char* s, int at, const char* w, int lw) {
    for (int j = 0; j < lw; ++j)
        if (s[at + j] != w[j]) return 0;
    return 1;
}

int oracle_like2(const char* s, int width, const char* w1, int l1, const char* w2, int l2) {
    int len = width;
    for (int i = 0; i < width; ++i)
        if (s[i] == 0) {
            len = i;
            break;
        }
    int p = -1;
    for (int i = 0; i + l1 <= len; ++i)
        if (match_at(s, i, w1, l1)) {
            p = i;
            break;
        }
    if (p < 0) return 0;
    for (int i = p + l1; i + l2 <= len; ++i)
        if (match_at(s, i, w2, l2)) return 1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Q13 analog — metal_plan_builder.cpp:1244-1299, codegen_main.cpp:1250-1277 */
/* ------------------------------------------------------------------------ */

typedef struct {
    const int32_t* s_fkey;
    const char* s_text;
    int width;
    int64_t max_key;
    const char *w1, *w2;
    int l1, l2, negate;
    uint32_t* counts;
} q13_ctx;

static void q13_range(void* vctx, uint64_t b, uint64_t e, int tid) {
    (void)tid;
    q13_ctx* c = (q13_ctx*)vctx;
    for (uint64_t i = b; i < e; ++i) {
        int m = oracle_like2(c->s_text + i * (uint64_t)c->width, c->width, c->w1, c->l1, c->w2, c->l2);
        if (c->negate) m = !m;
        const int64_t k = c->s_fkey[i];
        if (m && k >= 0 && k <= c->max_key) __atomic_fetch_add(&c->counts[k], 1u, __ATOMIC_RELAXED);
    }
}

int64_t oracle_q13(uint64_t n_s, const int32_t* s_fkey, const char* s_text, int width, uint64_t n_a,
                   const int32_t* a_key, int64_t max_key, const char* w1, int l1, const char* w2, int l2, int negate,
                   int bins, int64_t* out_count, int64_t* out_dist, int threads) {
    uint32_t* counts = (uint32_t*)calloc((size_t)(max_key + 1), 4);
    q13_ctx c = {s_fkey, s_text, width, max_key, w1, w2, l1, l2, negate, counts};
    parallel_ranges(n_s, threads, q13_range, &c);
    int64_t* hist = (int64_t*)calloc((size_t)bins, 8);
    for (uint64_t i = 0; i < n_a; ++i) {
        const int64_t k = a_key[i];
        const int64_t cnt = (k >= 0 && k <= max_key) ? counts[k] : 0;
        if (cnt < bins) hist[cnt] += 1;
    }
    int64_t m = 0;
    for (int c2 = 0; c2 < bins; ++c2)
        if (hist[c2]) {
            out_count[m] = c2;
            out_dist[m] = hist[c2];
            ++m;
        }
    /* order: custdist desc, then c_count desc (insertion sort; few rows) */
    for (int64_t i = 1; i < m; ++i) {
        const int64_t cc = out_count[i], dd = out_dist[i];
        int64_t j = i;
        while (j > 0 && (out_dist[j - 1] < dd || (out_dist[j - 1] == dd && out_count[j - 1] < cc))) {
            out_count[j] = out_count[j - 1];
            out_dist[j] = out_dist[j - 1];
            --j;
        }
        out_count[j] = cc;
        out_dist[j] = dd;
    }
    free(counts);
    free(hist);
    return m;
}

/* ------------------------------------------------------------------------ */
/* Q22 analog — metal_plan_builder.cpp:1304-1357                              */
/* ------------------------------------------------------------------------ */

void oracle_q22(uint64_t n_b, const int32_t* b_fkey, const int32_t* b_date, uint64_t n_a, const int32_t* a_key,
                const uint8_t* a_seg, int64_t max_key, int64_t date_lo, int64_t date_hi, int64_t* out10) {
    unsigned char* bset = (unsigned char*)calloc((size_t)(max_key + 1), 1);
    for (uint64_t i = 0; i < n_b; ++i) {
        const int64_t k = b_fkey[i];
        if (b_date[i] >= date_lo && b_date[i] < date_hi && k >= 0 && k <= max_key) bset[k] = 1;
    }
    for (int j = 0; j < 10; ++j) out10[j] = 0;
    for (uint64_t i = 0; i < n_a; ++i) {
        const int64_t k = a_key[i];
        const int in = (k >= 0 && k <= max_key) ? bset[k] : 0;
        if (in || a_seg[i] >= 5) continue;
        out10[a_seg[i]] += 1;
        out10[5 + a_seg[i]] += k;
    }
    free(bset);
}

/* ------------------------------------------------------------------------ */
/* Estimator MLP 16 -> 64 -> 64 -> 1 (ReLU), double precision                */
/* ------------------------------------------------------------------------ */

enum { MI = 16, MH = 64, OW1 = 0, OB1 = 1024, OW2 = 1088, OB2 = 5184, OW3 = 5248, OB3 = 5312, MNP = 5313 };

static double mlp_one(const float* P, const float* x, double* a1, double* h1, double* a2, double* h2) {
    for (int j = 0; j < MH; ++j) {
        double a = P[OB1 + j];
        for (int i = 0; i < MI; ++i) a += (double)P[OW1 + j * MI + i] * x[i];
        a1[j] = a;
        h1[j] = a > 0 ? a : 0;
    }
    double y = P[OB3];
    for (int j = 0; j < MH; ++j) {
        double a = P[OB2 + j];
        for (int k = 0; k < MH; ++k) a += (double)P[OW2 + j * MH + k] * h1[k];
        a2[j] = a;
        h2[j] = a > 0 ? a : 0;
        y += (double)P[OW3 + j] * h2[j];
    }
    return y;
}

void oracle_mlp_forward(const float* params, const float* X, int n, double* y) {
    double a1[MH], h1[MH], a2[MH], h2[MH];
    for (int s = 0; s < n; ++s) y[s] = mlp_one(params, X + (size_t)s * MI, a1, h1, a2, h2);
}

double oracle_mlp_grad(const float* params, const float* X, const float* T, int n, double scale, double* grad) {
    double a1[MH], h1[MH], a2[MH], h2[MH], d2[MH];
    double loss = 0;
    for (int p = 0; p < MNP; ++p) grad[p] = 0;
    for (int s = 0; s < n; ++s) {
        const float* x = X + (size_t)s * MI;
        const double y = mlp_one(params, x, a1, h1, a2, h2);
        const double diff = y - T[s];
        loss += scale * diff * diff;
        const double d3 = 2.0 * scale * diff;
        grad[OB3] += d3;
        for (int j = 0; j < MH; ++j) {
            grad[OW3 + j] += d3 * h2[j];
            d2[j] = a2[j] > 0 ? d3 * params[OW3 + j] : 0;
            grad[OB2 + j] += d2[j];
            for (int k = 0; k < MH; ++k) grad[OW2 + j * MH + k] += d2[j] * h1[k];
        }
        for (int k = 0; k < MH; ++k) {
            double acc = 0;
            for (int j = 0; j < MH; ++j) acc += (double)params[OW2 + j * MH + k] * d2[j];
            const double d1 = a1[k] > 0 ? acc : 0;
            grad[OB1 + k] += d1;
            for (int i = 0; i < MI; ++i) grad[OW1 + k * MI + i] += d1 * x[i];
        }
    }
    return loss;
}

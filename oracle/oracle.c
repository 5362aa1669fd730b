/*
 * oracle.c — CPU restatement of the reference pipelines (see oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker / timed CPU baseline.  The
 * product library never links or calls anything here.
 *
 * Reference paths are relative to the DataManagementLab/BespokeOLAP_Artifacts
 * root.  The reference's TPC-H pipelines (bespoke_tpch/query_q*.cpp) compile
 * from their own sources and are built, where the sources exist, into
 * oracle/_ref/ behind the shim oracle/ref_tpch.cpp; its loader, builder and
 * query() driver need the Arrow C++ libraries and two harness headers that
 * are not in the tree and are not built.  This restatement is pinned by
 * independent brute-force checks, invariants and closed forms in
 * tests/test_oracle.py, and, for the TPC-H pipelines, by CSV equality with the
 * compiled reference (tests/test_ref_tpch.py).
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

static int32_t g_mutation = 0;
void oracle_set_mutation(int32_t m) { g_mutation = m; }
int32_t oracle_mutation(void) { return g_mutation; }

/* ------------------------------------------------------------------------ */
/* Seeded generators.                                                         */
/* SplitMix64 output function (Steele, Lea & Flood, OOPSLA 2014), used as a   */
/* counter-based hash of (seed, stream, row): every value depends only on     */
/* its coordinates, so tables can be generated in any order or in parallel.   */
/* ------------------------------------------------------------------------ */
uint64_t og_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

uint64_t og_hash(uint64_t seed, uint32_t stream, uint64_t row) {
    const uint64_t k = og_mix64(seed + 0x9e3779b97f4a7c15ULL * ((uint64_t)stream + 1));
    return og_mix64(k + 0x9e3779b97f4a7c15ULL * (row + 1));
}

/* Uniform integer in [lo, hi] (hi - lo < 2^32) from the top 32 hash bits. */
int64_t og_uniform(uint64_t h, int64_t lo, int64_t hi) {
    const uint64_t range = (uint64_t)(hi - lo) + 1;
    return lo + (int64_t)(((h >> 32) * range) >> 32);
}

#define STREAM(t, c) ((uint32_t)(((t) << 8) | (c)))

/* Fact/lineitem money columns: qty U{1..50}; price = qty * U{90000..200000}
   / 100 in cents (integer division); disc U{0..10}; tax U{0..8}. */
void og_lineitem(uint64_t seed, int64_t rows, int64_t row0, int32_t* ship, int64_t* qty,
                 int64_t* price, int64_t* disc, int64_t* tax, uint8_t* flag_a,
                 uint8_t* flag_b) {
    for (int64_t i = 0; i < rows; ++i) {
        const uint64_t r = (uint64_t)(row0 + i);
        const int32_t s = (int32_t)og_uniform(og_hash(seed, STREAM(1, 0), r), 0, 2525);
        const int64_t q = og_uniform(og_hash(seed, STREAM(1, 1), r), 1, 50);
        const int64_t unit = og_uniform(og_hash(seed, STREAM(1, 2), r), 90000, 200000);
        const int32_t receipt =
            s + (int32_t)og_uniform(og_hash(seed, STREAM(1, 5), r), 1, 30);
        ship[i] = s;
        qty[i] = q;
        price[i] = q * unit / 100;
        disc[i] = og_uniform(og_hash(seed, STREAM(1, 3), r), 0, 10);
        tax[i] = og_uniform(og_hash(seed, STREAM(1, 4), r), 0, 8);
        /* TPC-H linestatus: 'O' when shipped after 1995-06-17 (day 1263);
           returnflag: 'R'/'A' when received by then, else 'N'.  Dictionaries
           flag_a = {A, N, R}, flag_b = {F, O}. */
        flag_b[i] = (uint8_t)(s > 1263 ? 1 : 0);
        flag_a[i] = (uint8_t)(receipt <= 1263
                                  ? (og_uniform(og_hash(seed, STREAM(1, 6), r), 0, 1) ? 2 : 0)
                                  : 1);
    }
}

static uint64_t gcd_u64(uint64_t a, uint64_t b) {
    while (b) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

/* Unique keys 1..rows in a shuffled order: an affine bijection of 0..rows-1. */
static uint64_t perm_multiplier(uint64_t n) {
    uint64_t a = 1000003ULL;
    while (gcd_u64(a, n) != 1) a += 2;
    return a;
}

void og_dim(uint64_t seed, int64_t rows, int32_t* key, uint8_t* cat, uint8_t* cont) {
    if (rows <= 0) return;
    const uint64_t n = (uint64_t)rows;
    const uint64_t a = perm_multiplier(n) % n;
    const uint64_t off = seed % n;
    for (int64_t i = 0; i < rows; ++i) {
        const uint64_t r = (uint64_t)i;
        key[i] = (int32_t)(((u128)r * a + off) % n) + 1;
        cat[i] = (uint8_t)og_uniform(og_hash(seed, STREAM(2, 1), r), 0, 149);
        cont[i] = (uint8_t)og_uniform(og_hash(seed, STREAM(2, 2), r), 0, 39);
    }
}

void og_fact(uint64_t seed, int64_t rows, int64_t row0, int64_t dim_rows, int32_t* fk,
             int64_t* price, int64_t* disc) {
    for (int64_t i = 0; i < rows; ++i) {
        const uint64_t r = (uint64_t)(row0 + i);
        fk[i] = (int32_t)og_uniform(og_hash(seed, STREAM(3, 0), r), 1, dim_rows);
        const int64_t q = og_uniform(og_hash(seed, STREAM(3, 1), r), 1, 50);
        const int64_t unit = og_uniform(og_hash(seed, STREAM(3, 2), r), 90000, 200000);
        price[i] = q * unit / 100;
        disc[i] = og_uniform(og_hash(seed, STREAM(3, 3), r), 0, 10);
    }
}

void og_customers(uint64_t seed, int64_t rows, int32_t* key, uint8_t* segment) {
    for (int64_t i = 0; i < rows; ++i) {
        key[i] = (int32_t)(i + 1);
        segment[i] = (uint8_t)og_uniform(og_hash(seed, STREAM(4, 1), (uint64_t)i), 0, 4);
    }
}

int64_t og_orders(uint64_t seed, int64_t rows, int64_t cust_rows, int32_t* key, int32_t* day,
                  int32_t* cust, int64_t* fact_offsets) {
    int64_t off = 0;
    for (int64_t i = 0; i < rows; ++i) {
        const uint64_t r = (uint64_t)i;
        key[i] = (int32_t)(i + 1);
        day[i] = (int32_t)og_uniform(og_hash(seed, STREAM(5, 1), r), 0, 2405);
        cust[i] = (int32_t)og_uniform(og_hash(seed, STREAM(5, 2), r), 1, cust_rows);
        fact_offsets[i] = off;
        off += og_uniform(og_hash(seed, STREAM(5, 3), r), 1, 7);
    }
    fact_offsets[rows] = off;
    return off;
}

void og_order_facts(uint64_t seed, int64_t order_rows, const int64_t* fact_offsets,
                    int32_t* fk_order, int64_t* price, int64_t* disc) {
    for (int64_t o = 0; o < order_rows; ++o) {
        for (int64_t j = fact_offsets[o]; j < fact_offsets[o + 1]; ++j) {
            const uint64_t r = (uint64_t)j;
            fk_order[j] = (int32_t)(o + 1);
            const int64_t q = og_uniform(og_hash(seed, STREAM(6, 1), r), 1, 50);
            const int64_t unit = og_uniform(og_hash(seed, STREAM(6, 2), r), 90000, 200000);
            price[j] = q * unit / 100;
            disc[j] = og_uniform(og_hash(seed, STREAM(6, 3), r), 0, 10);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Q1.  Restates the scan of run_q1 (bespoke_tpch/query_q1.cpp:143-228):      */
/* rows with ship <= cutoff add quantity, extendedprice, discounted price,    */
/* discounted price * (100 + tax) into a 128-bit charge, discount and a       */
/* count into aggregates[group_code] (Q1Agg, query_q1.cpp:27-34).  The range  */
/* lower bound generalises the single <= of the reference.  Rows whose codes  */
/* fall outside the dictionaries are ignored.                                 */
/* ------------------------------------------------------------------------ */
static void q1_add(oracle_q1_group* g, int64_t q, int64_t p, int64_t dp, int64_t d,
                   int64_t t) {
    const int64_t tax_mul = (g_mutation == 1) ? (100 - t) : (100 + t);
    const int64_t ch = (int64_t)((uint64_t)dp * (uint64_t)tax_mul);
    u128 acc = ((u128)(uint64_t)g->sum_charge_hi << 64) | g->sum_charge_lo;
    acc += (u128)(__int128)ch;
    g->sum_charge_lo = (uint64_t)acc;
    g->sum_charge_hi = (int64_t)(uint64_t)(acc >> 64);
    g->sum_qty = (int64_t)((uint64_t)g->sum_qty + (uint64_t)q);
    g->sum_base_price = (int64_t)((uint64_t)g->sum_base_price + (uint64_t)p);
    g->sum_disc_price = (int64_t)((uint64_t)g->sum_disc_price + (uint64_t)dp);
    g->sum_disc = (int64_t)((uint64_t)g->sum_disc + (uint64_t)d);
    g->count += 1;
}

void oracle_q1_wide(int64_t rows, const int32_t* ship, int64_t ship_lo, int64_t ship_hi,
                    const int64_t* qty, const int64_t* price, const int64_t* disc,
                    const int64_t* tax, const uint8_t* flag_a, const uint8_t* flag_b,
                    int32_t n_a, int32_t n_b, oracle_q1_group* out) {
    memset(out, 0, sizeof(oracle_q1_group) * (size_t)(n_a * n_b));
    for (int64_t i = 0; i < rows; ++i) {
        if (ship[i] < ship_lo || ship[i] > ship_hi) continue;
        if (flag_a[i] >= n_a || flag_b[i] >= n_b) continue;
        const int64_t dp = (int64_t)((uint64_t)price[i] * (uint64_t)(100 - disc[i]));
        q1_add(&out[flag_a[i] * n_b + flag_b[i]], qty[i], price[i], dp, disc[i], tax[i]);
    }
}

void oracle_q1_ref(int64_t rows, const int16_t* ship, int64_t ship_lo, int64_t ship_hi,
                   const int16_t* qty, const int32_t* price, const int32_t* disc_price,
                   const uint8_t* disc, const uint8_t* tax, const uint8_t* code,
                   int32_t n_groups, oracle_q1_group* out) {
    memset(out, 0, sizeof(oracle_q1_group) * (size_t)n_groups);
    for (int64_t i = 0; i < rows; ++i) {
        if (ship[i] < ship_lo || ship[i] > ship_hi) continue;
        if (code[i] >= n_groups) continue;
        q1_add(&out[code[i]], qty[i], price[i], disc_price[i], disc[i], tax[i]);
    }
}

/* ------------------------------------------------------------------------ */
/* Predicates: conjunction of range / set tests on one row.                   */
/* ------------------------------------------------------------------------ */
static int64_t load_col(const void* col, int32_t dtype, int64_t row) {
    switch (dtype) {
        case 0: return ((const uint8_t*)col)[row];
        case 1: return ((const int16_t*)col)[row];
        case 2: return ((const int32_t*)col)[row];
        default: return ((const int64_t*)col)[row];
    }
}

int oracle_pred_row(const oracle_pred* p, int64_t row) {
    const int64_t v = load_col(p->col, p->dtype, row);
    if (p->op == 0) return v >= p->lo && v <= p->hi;
    if (p->set) {
        if (v < 0 || v >= p->set_bits) return 0;
        return (p->set[v >> 5] >> (v & 31)) & 1u;
    }
    if (v < 0 || v > 255) return 0;
    return (int)((p->inline_set[v >> 6] >> (v & 63)) & 1u);
}

/* ------------------------------------------------------------------------ */
/* Join + scalar sum.  Restates run_q19 / run_q17: the dimension scan marks  */
/* passing keys in a byte mask over the key domain                            */
/* (bespoke_tpch/query_q19.cpp:210-250, query_q17.cpp:161-177), then every    */
/* fact row whose key is marked adds price * (100 - disc) into one int64      */
/* (query_q19.cpp:271-345).  A key repeated among passing dimension rows     */
/* still matches each fact row once (the mask semantics).                     */
/* ------------------------------------------------------------------------ */
void oracle_join_sum(int64_t dim_rows, const int32_t* dim_key, int32_t n_preds,
                     const oracle_pred* preds, int64_t fact_rows, const int32_t* fact_fk,
                     const void* fact_price, int32_t price_dtype, const void* fact_disc,
                     int32_t disc_dtype, oracle_join_sum_out* out) {
    /* The reference indexes its mask by the key itself; an offset by the
       smallest key extends that to negative keys. */
    int64_t min_key = 0, max_key = -1;
    for (int64_t i = 0; i < dim_rows; ++i) {
        if (i == 0 || dim_key[i] < min_key) min_key = dim_key[i];
        if (i == 0 || dim_key[i] > max_key) max_key = dim_key[i];
    }
    const int64_t span = max_key - min_key + 1;
    uint8_t* mask = (uint8_t*)calloc((size_t)(span > 0 ? span : 1), 1);
    int64_t build = 0;
    for (int64_t i = 0; i < dim_rows; ++i) {
        int pass = 1;
        for (int32_t p = 0; p < n_preds && pass; ++p) pass = oracle_pred_row(&preds[p], i);
        if (pass) {
            mask[dim_key[i] - min_key] = 1;
            ++build;
        }
    }
    uint64_t revenue = 0;
    int64_t matches = 0;
    for (int64_t i = 0; i < fact_rows; ++i) {
        const int64_t k = fact_fk[i];
        if (k < min_key || k > max_key || !mask[k - min_key]) continue;
        const int64_t p = load_col(fact_price, price_dtype, i);
        const int64_t d = load_col(fact_disc, disc_dtype, i);
        const int64_t factor = (g_mutation == 2) ? d : (100 - d);
        revenue += (uint64_t)p * (uint64_t)factor;
        ++matches;
    }
    free(mask);
    out->revenue = (int64_t)revenue;
    out->build_rows = build;
    out->matches = matches;
}

/* ------------------------------------------------------------------------ */
/* Chain join + group-by.  Restates run_q3 (bespoke_tpch/query_q3.cpp:       */
/* 182-318): a byte mask of customers in the segment (:182-205), orders with  */
/* day < limit whose customer is marked (:207-210, :306-318), and per order   */
/* the sum of its fact rows' revenue (:219-260).  The reference walks the     */
/* orders' contiguous fact ranges; for fact tables in any order this uses the */
/* dense per-key array of run_q18 (query_q18.cpp:221-255).  Groups come out   */
/* in order-row order; orders without fact rows are dropped (:244-246).       */
/* Order keys are unique; customer keys may repeat (set semantics).          */
/* ------------------------------------------------------------------------ */
int64_t oracle_chain_groups(int64_t cust_rows, const int32_t* cust_key,
                            const uint8_t* cust_segment, int32_t segment, int64_t order_rows,
                            const int32_t* order_key, const int32_t* order_day,
                            const int32_t* order_cust, int32_t day_lt, int64_t fact_rows,
                            const int32_t* fact_order, const void* fact_price,
                            int32_t price_dtype, const void* fact_disc, int32_t disc_dtype,
                            const int32_t* fact_disc_price, const int16_t* fact_ship,
                            int32_t ship_gt, oracle_group_row* out, int64_t max_out,
                            oracle_chain_stats* stats) {
    /* Key-indexed arrays as in the reference, offset by the smallest key. */
    int64_t cmin = 0, cmax = -1;
    for (int64_t i = 0; i < cust_rows; ++i) {
        if (i == 0 || cust_key[i] < cmin) cmin = cust_key[i];
        if (i == 0 || cust_key[i] > cmax) cmax = cust_key[i];
    }
    uint8_t* in_seg = (uint8_t*)calloc((size_t)(cmax >= cmin ? cmax - cmin + 1 : 1), 1);
    int64_t cust_pass = 0;
    for (int64_t i = 0; i < cust_rows; ++i) {
        if (cust_segment[i] == segment) {
            in_seg[cust_key[i] - cmin] = 1;
            ++cust_pass;
        }
    }
    int64_t omin = 0, omax = -1;
    for (int64_t i = 0; i < order_rows; ++i) {
        if (i == 0 || order_key[i] < omin) omin = order_key[i];
        if (i == 0 || order_key[i] > omax) omax = order_key[i];
    }
    const int64_t ospan = omax >= omin ? omax - omin + 1 : 1;
    int32_t* row_of = (int32_t*)malloc(sizeof(int32_t) * (size_t)ospan);
    for (int64_t k = 0; k < ospan; ++k) row_of[k] = -1;
    int64_t order_pass = 0;
    const int32_t day_limit = (g_mutation == 3) ? day_lt + 1 : day_lt;
    for (int64_t o = 0; o < order_rows; ++o) {
        const int64_t c = order_cust[o];
        if (order_day[o] >= day_limit) continue;
        if (c < cmin || c > cmax || !in_seg[c - cmin]) continue;
        row_of[order_key[o] - omin] = (int32_t)o;
        ++order_pass;
    }
    uint64_t* rev = (uint64_t*)calloc((size_t)order_rows + 1, sizeof(uint64_t));
    int32_t* cnt = (int32_t*)calloc((size_t)order_rows + 1, sizeof(int32_t));
    int64_t matches = 0;
    for (int64_t i = 0; i < fact_rows; ++i) {
        const int64_t k = fact_order[i];
        if (k < omin || k > omax) continue;
        const int32_t o = row_of[k - omin];
        if (o < 0) continue;
        /* run_q3 sums only rows shipped after the date (query_q3.cpp:236-241) */
        if (fact_ship && fact_ship[i] <= ship_gt) continue;
        int64_t v;
        if (fact_disc_price) {
            v = fact_disc_price[i];
        } else {
            v = (int64_t)((uint64_t)load_col(fact_price, price_dtype, i) *
                          (uint64_t)(100 - load_col(fact_disc, disc_dtype, i)));
        }
        rev[o] += (uint64_t)v;
        cnt[o] += 1;
        ++matches;
    }
    int64_t groups = 0;
    for (int64_t o = 0; o < order_rows; ++o) {
        if (cnt[o] == 0) continue;
        if (groups < max_out) {
            out[groups].key = order_key[o];
            out[groups].day = order_day[o];
            out[groups].revenue = (int64_t)rev[o];
        }
        ++groups;
    }
    free(in_seg);
    free(row_of);
    free(rev);
    free(cnt);
    if (stats) {
        stats->cust_pass = cust_pass;
        stats->order_pass = order_pass;
        stats->fact_matches = matches;
        stats->groups = groups;
    }
    return groups;
}

/* Ordering of run_q3 (query_q3.cpp:348-356): revenue descending, then day
   ascending; ties beyond that are broken by key ascending. */
static int cmp_groups(const void* pa, const void* pb) {
    const oracle_group_row* a = (const oracle_group_row*)pa;
    const oracle_group_row* b = (const oracle_group_row*)pb;
    if (a->revenue != b->revenue) {
        const int desc = (a->revenue > b->revenue) ? -1 : 1;
        return (g_mutation == 4) ? -desc : desc;
    }
    if (a->day != b->day) return a->day < b->day ? -1 : 1;
    if (a->key != b->key) return a->key < b->key ? -1 : 1;
    return 0;
}

void oracle_sort_groups(oracle_group_row* rows, int64_t n) {
    qsort(rows, (size_t)n, sizeof(oracle_group_row), cmp_groups);
}

/* ------------------------------------------------------------------------ */
/* run_q3's own algorithm over the reference layout (query_q3.cpp:182-356):   */
/* a byte mask over customer keys for the segment (:182-205); orders sorted   */
/* by date at build (builder_impl.cpp:597-608), so the passing orders are the */
/* prefix day < day_lt found by lower_bound (:207-210); for every prefix      */
/* order whose customer is marked, the sum of its fact range                  */
/* [range_start, range_end) (orders.lineitem_ranges, an empty range has end   */
/* 0; :219-260, :306-318); then the full sort of the group rows (:348-356).   */
/* Revenue is price * (100 - disc) of i64 columns.  Used as the CPU baseline  */
/* of the chain join; returns the number of groups (rows sorted in out).      */
/* ------------------------------------------------------------------------ */
int64_t oracle_q3_ranges(int64_t cust_rows, const int32_t* cust_key, const uint8_t* cust_segment,
                         int32_t segment, int64_t order_rows, const int32_t* order_key,
                         const int32_t* order_day, const int32_t* order_cust,
                         const uint32_t* range_start, const uint32_t* range_end, int32_t day_lt,
                         const int64_t* fact_price, const int64_t* fact_disc,
                         oracle_group_row* out, int64_t max_out) {
    int32_t max_cust = -1;
    for (int64_t i = 0; i < cust_rows; ++i)
        if (cust_key[i] > max_cust) max_cust = cust_key[i];
    const size_t mask_n = (size_t)(max_cust + 1);
    uint8_t* in_seg = (uint8_t*)calloc(mask_n ? mask_n : 1, 1);
    for (int64_t i = 0; i < cust_rows; ++i)
        if (cust_key[i] >= 0 && cust_segment[i] == segment) in_seg[cust_key[i]] = 1;
    int64_t lo = 0, hi = order_rows;  /* lower_bound(order_day, day_lt) */
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (order_day[mid] < day_lt) lo = mid + 1;
        else hi = mid;
    }
    const int64_t limit = lo;
    int64_t groups = 0;
    for (int64_t o = 0; o < limit; ++o) {
        const int32_t c = order_cust[o];
        if (c < 0 || (size_t)c >= mask_n || !in_seg[c]) continue;
        const uint32_t s = range_start[o], e = range_end[o];
        if (e == 0) continue;
        int64_t rev = 0;
        int has = 0;
        for (uint32_t r = s; r < e; ++r) {
            rev += fact_price[r] * (100 - fact_disc[r]);
            has = 1;
        }
        if (!has) continue;
        if (groups < max_out) {
            out[groups].key = order_key[o];
            out[groups].day = order_day[o];
            out[groups].revenue = rev;
        }
        ++groups;
    }
    free(in_seg);
    const int64_t kept = groups < max_out ? groups : max_out;
    oracle_sort_groups(out, kept);
    return groups;
}


/* ------------------------------------------------------------------------ */
/* LSD radix sort with 16-bit digits, stable counting-sort passes, as         */
/* radix_sort_rows_by_orderkey / radix_sort_groups                            */
/* (bespoke_tpch/query_q10.cpp:134-166, 318-373).  Only the low key_bits bits */
/* of each key take part.                                                     */
/* ------------------------------------------------------------------------ */
void oracle_radix_sort_u64(uint64_t* keys, uint32_t* vals, int64_t n, int32_t key_bits) {
    if (n < 2 || key_bits <= 0) return;
    const uint64_t kmask = key_bits >= 64 ? ~0ULL : ((1ULL << key_bits) - 1);
    uint64_t* tk = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
    uint32_t* tv = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    uint32_t* counts = (uint32_t*)malloc(sizeof(uint32_t) * 65536);
    uint64_t* sk = keys;
    uint32_t* sv = vals;
    uint64_t* dk = tk;
    uint32_t* dv = tv;
    for (int32_t shift = 0; shift < key_bits; shift += 16) {
        const int32_t dshift = (g_mutation == 5 && shift > 0) ? shift + 1 : shift;
        memset(counts, 0, sizeof(uint32_t) * 65536);
        for (int64_t i = 0; i < n; ++i) counts[((sk[i] & kmask) >> dshift) & 0xffff] += 1;
        uint32_t sum = 0;
        for (int b = 0; b < 65536; ++b) {
            const uint32_t c = counts[b];
            counts[b] = sum;
            sum += c;
        }
        for (int64_t i = 0; i < n; ++i) {
            const uint32_t pos = counts[((sk[i] & kmask) >> dshift) & 0xffff]++;
            dk[pos] = sk[i];
            dv[pos] = sv[i];
        }
        uint64_t* t1 = sk; sk = dk; dk = t1;
        uint32_t* t2 = sv; sv = dv; dv = t2;
    }
    if (sk != keys) {
        memcpy(keys, sk, sizeof(uint64_t) * (size_t)n);
        memcpy(vals, sv, sizeof(uint32_t) * (size_t)n);
    }
    free(tk);
    free(tv);
    free(counts);
}

/* k smallest keys ascending; equal keys keep input order (a stable sort's
   first k entries). */
void oracle_topk_u64(const uint64_t* keys, const uint32_t* vals, int64_t n, int32_t k,
                     uint64_t* keys_out, uint32_t* vals_out) {
    uint64_t* kk = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(n > 0 ? n : 1));
    uint32_t* vv = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    memcpy(kk, keys, sizeof(uint64_t) * (size_t)n);
    memcpy(vv, vals, sizeof(uint32_t) * (size_t)n);
    oracle_radix_sort_u64(kk, vv, n, 64);
    const int64_t m = n < k ? n : k;
    memcpy(keys_out, kk, sizeof(uint64_t) * (size_t)m);
    memcpy(vals_out, vv, sizeof(uint32_t) * (size_t)m);
    free(kk);
    free(vv);
}

/*
 * tpch.c — CPU restatement of the reference's Q6 and Q18 pipelines and of the
 * build-time derivations the device mirror needs (see oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's cpu_baseline leg
 * as the checker / timed CPU baseline.  The product library never links or
 * calls anything here.  Reference paths are relative to the
 * DataManagementLab/BespokeOLAP_Artifacts root.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Civil dates.  Howard Hinnant's days-from-civil / civil-from-days algorithms */
/* ("chrono-Compatible Low-Level Date Algorithms"), which the reference uses  */
/* in bespoke_tpch/builder_impl.cpp:22-40 and query_q6.cpp:33-40.             */
/* ------------------------------------------------------------------------ */
int64_t oracle_days_from_civil(int32_t y, int32_t m, int32_t d) {
    y -= m <= 2;
    const int32_t era = (y >= 0 ? y : y - 399) / 400;
    const uint32_t yoe = (uint32_t)(y - era * 400);
    const uint32_t doy = (153u * (uint32_t)(m + (m > 2 ? -3 : 9)) + 2u) / 5u + (uint32_t)d - 1u;
    const uint32_t doe = yoe * 365u + yoe / 4u - yoe / 100u + doy;
    return (int64_t)era * 146097 + (int64_t)doe - 719468;
}

void oracle_civil_from_days(int32_t z, int32_t* y, int32_t* m, int32_t* d) {
    z += 719468;
    const int32_t era = (z >= 0 ? z : z - 146096) / 146097;
    const uint32_t doe = (uint32_t)(z - era * 146097);
    const uint32_t yoe = (doe - doe / 1460u + doe / 36524u - doe / 146096u) / 365u;
    int32_t yy = (int32_t)yoe + era * 400;
    const uint32_t doy = doe - (365u * yoe + yoe / 4u - yoe / 100u);
    const uint32_t mp = (5u * doy + 2u) / 153u;
    *d = (int32_t)(doy - (153u * mp + 2u) / 5u + 1u);
    *m = (int32_t)(mp < 10u ? mp + 3u : mp - 9u);
    *y = yy + (*m <= 2);
}

/* Days of 1992-01-01, the reference's base date (builder_impl.cpp:454). */
static int32_t base_days(void) { return (int32_t)oracle_days_from_civil(1992, 1, 1); }

/* ------------------------------------------------------------------------ */
/* Q6, flat loop: bespoke_tpch/query_q6.cpp:197-223.  A row passes when        */
/* (u32)(ship - lo) < (u32)(hi - lo), (u32)(disc - dlo) <= (u32)(dhi - dlo)    */
/* and qty < qty_lt; revenue sums price * disc in int64.                      */
/* ------------------------------------------------------------------------ */
void oracle_q6_flat(int64_t rows, const int16_t* ship, const uint8_t* disc, const int16_t* qty,
                    const int32_t* price, int32_t ship_lo, int32_t ship_hi, int32_t disc_lo,
                    int32_t disc_hi, int32_t qty_lt, oracle_q6_out* out) {
    const uint32_t date_span = (uint32_t)(ship_hi - ship_lo);
    const uint32_t disc_span = (uint32_t)(disc_hi - disc_lo);
    uint64_t rev = 0;
    int64_t emitted = 0;
    for (int64_t i = 0; i < rows; ++i) {
        if ((uint32_t)((int32_t)ship[i] - ship_lo) >= date_span) continue;
        const int32_t d = (int32_t)disc[i];
        if ((uint32_t)(d - disc_lo) > disc_span) continue;
        if ((int32_t)qty[i] >= qty_lt) continue;
        rev += (uint64_t)((int64_t)price[i] * (int64_t)d);
        emitted += 1;
    }
    out->revenue = (int64_t)rev;
    out->rows_emitted = emitted;
    out->rows_scanned = rows;
}

/* ------------------------------------------------------------------------ */
/* Q6 over shards.  Shards as built by builder_impl.cpp:855-911: rows grouped  */
/* by (ship year*12 + month - 1, discount, quantity / 100, supplier nation);   */
/* here every row has the "no nation" bucket (255).  Each shard keeps min/max  */
/* ship date, discount and quantity.  The scan follows query_q6.cpp:224-379:   */
/* month-index pruning, zone-map pruning, then per-row tests only for the      */
/* predicates a shard does not satisfy wholly, with the shard's single         */
/* discount value as the multiplier.                                          */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t key;
    int64_t row;
} shard_pair;

static int cmp_shard_pair(const void* a, const void* b) {
    const shard_pair* x = (const shard_pair*)a;
    const shard_pair* y = (const shard_pair*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->row < y->row ? -1 : x->row > y->row;
}

void oracle_q6_sharded(int64_t rows, const int16_t* ship, const uint8_t* disc,
                       const int16_t* qty, const int32_t* price, int32_t ship_lo,
                       int32_t ship_hi, int32_t disc_lo, int32_t disc_hi, int32_t qty_lt,
                       int32_t month_lo, int32_t month_hi, oracle_q6_out* out) {
    const int32_t base = base_days();
    shard_pair* p = (shard_pair*)malloc(sizeof(shard_pair) * (size_t)(rows > 0 ? rows : 1));
    for (int64_t i = 0; i < rows; ++i) {
        int32_t y, m, d;
        oracle_civil_from_days(base + ship[i], &y, &m, &d);
        const int64_t month_index = (int64_t)y * 12 + (m - 1);
        p[i].key = (month_index << 24) | ((int64_t)disc[i] << 16) |
                   ((int64_t)(qty[i] / 100) << 8) | 255;
        p[i].row = i;
    }
    /* Rows of a shard keep their table order (row_indices are pushed in row
       order); shards are visited here in key order, which only changes the
       order of an exact integer sum. */
    qsort(p, (size_t)rows, sizeof(shard_pair), cmp_shard_pair);
    const uint32_t date_span = (uint32_t)(ship_hi - ship_lo);
    uint64_t rev = 0;
    int64_t emitted = 0, scanned = 0;
    for (int64_t s = 0; s < rows;) {
        int64_t e = s + 1;
        while (e < rows && p[e].key == p[s].key) ++e;
        int16_t smin = ship[p[s].row], smax = smin, qmin = qty[p[s].row], qmax = qmin;
        int32_t dmin = disc[p[s].row], dmax = dmin;
        for (int64_t j = s; j < e; ++j) {
            const int64_t r = p[j].row;
            if (ship[r] < smin) smin = ship[r];
            if (ship[r] > smax) smax = ship[r];
            if (qty[r] < qmin) qmin = qty[r];
            if (qty[r] > qmax) qmax = qty[r];
            if (disc[r] < dmin) dmin = disc[r];
            if (disc[r] > dmax) dmax = disc[r];
        }
        const int32_t shard_index = (int32_t)(p[s].key >> 24);
        const int skip = shard_index < month_lo || shard_index > month_hi ||
                         smax < ship_lo || smin >= ship_hi || dmax < disc_lo ||
                         dmin > disc_hi || (int32_t)qmin >= qty_lt;
        if (!skip) {
            const int full_date = smin >= ship_lo && smax < ship_hi;
            /* planted mistake 9: an off-by-one "whole shard below the limit" */
            const int full_qty = oracle_mutation() == 9 ? qmax <= qty_lt : qmax < qty_lt;
            const int64_t dv = dmin;
            scanned += e - s;
            for (int64_t j = s; j < e; ++j) {
                const int64_t r = p[j].row;
                if (!full_date && (uint32_t)((int32_t)ship[r] - ship_lo) >= date_span) continue;
                if (!full_qty && (int32_t)qty[r] >= qty_lt) continue;
                rev += (uint64_t)((int64_t)price[r] * dv);
                emitted += 1;
            }
        }
        s = e;
    }
    free(p);
    out->revenue = (int64_t)rev;
    out->rows_emitted = emitted;
    out->rows_scanned = scanned;
}

/* ------------------------------------------------------------------------ */
/* Q18: bespoke_tpch/query_q18.cpp:183-455.                                   */
/* Dense path (:204-345) when orders and customers carry keys 1..N (checked    */
/* on the first and last key only, as the reference does): sum_qty indexed by  */
/* order key, an order qualifies the first time its running sum exceeds the    */
/* threshold, order row = key - 1, customer row = custkey - 1.  Otherwise      */
/* (:346-440): sum_qty over 0..max key, every order row tested in table order, */
/* customers through custkey -> row (last row wins).  Rows are then ordered by */
/* totalprice desc, orderdate asc, orderkey asc (:442-455); the order row      */
/* breaks the remaining ties (duplicate order keys), which std::sort leaves    */
/* open.  Lineitem keys outside the sum array are skipped (the reference would */
/* index out of bounds there).                                                 */
/* ------------------------------------------------------------------------ */
static int cmp_q18(const void* a, const void* b) {
    const oracle_q18_row* x = (const oracle_q18_row*)a;
    const oracle_q18_row* y = (const oracle_q18_row*)b;
    if (x->totalprice != y->totalprice) return x->totalprice > y->totalprice ? -1 : 1;
    if (x->orderdate != y->orderdate) return x->orderdate < y->orderdate ? -1 : 1;
    if (x->orderkey != y->orderkey) return x->orderkey < y->orderkey ? -1 : 1;
    return x->order_row < y->order_row ? -1 : x->order_row > y->order_row;
}

static int32_t max_i32(const int32_t* v, int64_t n) {
    int32_t m = -1;
    for (int64_t i = 0; i < n; ++i)
        if (v[i] > m) m = v[i];
    return m;
}

int64_t oracle_q18(int64_t li_rows, const int32_t* l_orderkey, const int16_t* l_qty,
                   int64_t o_rows, const int32_t* o_orderkey, const int32_t* o_custkey,
                   const int16_t* o_orderdate, const int32_t* o_totalprice, int64_t c_rows,
                   const int32_t* c_custkey, int32_t threshold, oracle_q18_row* out,
                   int64_t max_out, int32_t* dense_path) {
    const int dense_orders = o_rows > 0 && o_orderkey[0] == 1 && o_orderkey[o_rows - 1] == o_rows;
    const int dense_cust = c_rows > 0 && c_custkey[0] == 1 && c_custkey[c_rows - 1] == c_rows;
    *dense_path = dense_orders;
    int64_t n = 0;
    int32_t* cust_row_of = NULL;
    int64_t cust_map = 0;
    if (!dense_cust || !dense_orders) {
        const int32_t mc = max_i32(c_custkey, c_rows);
        if (mc < 0) return 0;
        cust_map = (int64_t)mc + 1;
        cust_row_of = (int32_t*)malloc(sizeof(int32_t) * (size_t)cust_map);
        for (int64_t i = 0; i < cust_map; ++i) cust_row_of[i] = -1;
        for (int64_t r = 0; r < c_rows; ++r)
            if (c_custkey[r] >= 0) cust_row_of[c_custkey[r]] = (int32_t)r;
    }
    if (dense_orders) {
        if (c_rows == 0) {
            free(cust_row_of);
            return 0;
        }
        const int64_t size = o_rows + 1;
        int32_t* sum = (int32_t*)calloc((size_t)size, sizeof(int32_t));
        uint8_t* qualified = (uint8_t*)calloc((size_t)size, 1);
        int32_t* keys = (int32_t*)malloc(sizeof(int32_t) * (size_t)size);
        int64_t nq = 0;
        for (int64_t i = 0; i < li_rows; ++i) {
            const int32_t k = l_orderkey[i];
            if (k < 0 || k >= size) continue;
            const int32_t upd = (int32_t)((uint32_t)sum[k] + (uint32_t)(int32_t)l_qty[i]);
            sum[k] = upd;
            /* planted mistake 10: >= instead of > */
            const int over = oracle_mutation() == 10 ? upd >= threshold : upd > threshold;
            if (over && !qualified[k]) {
                qualified[k] = 1;
                keys[nq++] = k;
            }
        }
        for (int64_t j = 0; j < nq; ++j) {
            const int64_t orow = (int64_t)keys[j] - 1;
            if (orow < 0 || orow >= o_rows) continue;
            const int32_t ck = o_custkey[orow];
            int64_t crow;
            if (dense_cust) {
                crow = (int64_t)ck - 1;
                if (crow < 0 || crow >= c_rows) continue;
            } else {
                if (ck < 0 || ck >= cust_map || cust_row_of[ck] < 0) continue;
                crow = cust_row_of[ck];
            }
            if (n < max_out) {
                oracle_q18_row* r = &out[n];
                r->custkey = ck;
                r->orderkey = keys[j];
                r->orderdate = o_orderdate[orow];
                r->totalprice = o_totalprice[orow];
                r->sum_qty = sum[keys[j]];
                r->cust_row = (int32_t)crow;
                r->order_row = (int32_t)orow;
            }
            ++n;
        }
        free(sum);
        free(qualified);
        free(keys);
    } else {
        int32_t mk = max_i32(o_orderkey, o_rows);
        const int32_t ml = max_i32(l_orderkey, li_rows);
        if (ml > mk) mk = ml;
        if (mk < 0) {
            free(cust_row_of);
            return 0;
        }
        const int64_t size = (int64_t)mk + 1;
        int32_t* sum = (int32_t*)calloc((size_t)size, sizeof(int32_t));
        for (int64_t i = 0; i < li_rows; ++i) {
            const int32_t k = l_orderkey[i];
            if (k < 0) continue;
            sum[k] = (int32_t)((uint32_t)sum[k] + (uint32_t)(int32_t)l_qty[i]);
        }
        for (int64_t o = 0; o < o_rows; ++o) {
            const int32_t k = o_orderkey[o];
            if (k < 0) continue;
            const int32_t total = sum[k];
            if (oracle_mutation() == 10 ? total < threshold : total <= threshold) continue;
            const int32_t ck = o_custkey[o];
            if (ck < 0 || ck >= cust_map || cust_row_of[ck] < 0) continue;
            if (n < max_out) {
                oracle_q18_row* r = &out[n];
                r->custkey = ck;
                r->orderkey = k;
                r->orderdate = o_orderdate[o];
                r->totalprice = o_totalprice[o];
                r->sum_qty = total;
                r->cust_row = cust_row_of[ck];
                r->order_row = (int32_t)o;
            }
            ++n;
        }
        free(sum);
    }
    free(cust_row_of);
    if (n <= max_out) qsort(out, (size_t)n, sizeof(oracle_q18_row), cmp_q18);
    if (oracle_mutation() == 11 && n > 1 && n <= max_out) {  /* planted: first/last swapped */
        oracle_q18_row t = out[0];
        out[0] = out[n - 1];
        out[n - 1] = t;
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* Build-time derivations of the lineitem/orders layout.                      */
/* ------------------------------------------------------------------------ */
/* discounted_price = price * (100 - disc) / 100 truncated toward zero, and   */
/* the combined returnflag/linestatus code rf * n_ls + ls                     */
/* (builder_impl.cpp:523-532, 552-567).  Returns 1 when the order keys are    */
/* non-decreasing (builder_impl.cpp:569-577).                                 */
int32_t oracle_derive_lineitem(int64_t rows, const int32_t* price, const uint8_t* disc,
                               const uint8_t* rf, const uint8_t* ls, int32_t n_ls,
                               const int32_t* orderkey, int32_t* disc_price, uint8_t* code) {
    int32_t sorted = 1;
    int32_t prev = INT32_MIN;
    for (int64_t i = 0; i < rows; ++i) {
        disc_price[i] = (int32_t)((int64_t)price[i] * (100 - (int32_t)disc[i]) / 100);
        code[i] = (uint8_t)((uint32_t)rf[i] * (uint32_t)n_ls + (uint32_t)ls[i]);
        if (orderkey[i] < prev) sorted = 0;
        prev = orderkey[i];
    }
    return sorted;
}

/* Stable counting-sort order of the order dates (builder_impl.cpp:374-405):
   order[k] = row placed k-th. */
void oracle_date_order(int64_t rows, const int16_t* date, uint32_t* order) {
    if (rows <= 0) return;
    int32_t lo = date[0], hi = date[0];
    for (int64_t i = 0; i < rows; ++i) {
        if (date[i] < lo) lo = date[i];
        if (date[i] > hi) hi = date[i];
    }
    const int64_t range = (int64_t)hi - lo + 1;
    uint32_t* off = (uint32_t*)calloc((size_t)range + 1, sizeof(uint32_t));
    for (int64_t i = 0; i < rows; ++i) off[date[i] - lo + 1] += 1;
    for (int64_t b = 1; b <= range; ++b) off[b] += off[b - 1];
    for (int64_t i = 0; i < rows; ++i) order[off[date[i] - lo]++] = (uint32_t)i;
    free(off);
}

/* orderkey -> order row over 0..max key (-1 when absent; the last row with a
   key wins) and each order's [start, end) range of lineitem rows
   (builder_impl.cpp:610-630, 929-991).  `map` holds max key + 1 entries. */
void oracle_order_ranges(int64_t o_rows, const int32_t* o_orderkey, int64_t map_size,
                         int32_t* map, int64_t li_rows, const int32_t* l_orderkey,
                         int32_t li_sorted, uint32_t* start, uint32_t* end) {
    for (int64_t k = 0; k < map_size; ++k) map[k] = -1;
    for (int64_t r = 0; r < o_rows; ++r) {
        const int32_t k = o_orderkey[r];
        if (k >= 0 && k < map_size) map[k] = (int32_t)r;
    }
    for (int64_t r = 0; r < o_rows; ++r) start[r] = end[r] = 0;
    if (map_size == 0 || li_rows == 0) return;
    if (li_sorted) {
        uint32_t s = 0;
        int32_t cur = l_orderkey[0];
        for (int64_t i = 1; i <= li_rows; ++i) {
            if (i < li_rows && l_orderkey[i] == cur) continue;
            if (cur >= 0 && cur < map_size) {
                const int32_t row = map[cur];
                if (row >= 0 && row < o_rows) {
                    start[row] = s;
                    end[row] = (uint32_t)i;
                }
            }
            if (i < li_rows) {
                s = (uint32_t)i;
                cur = l_orderkey[i];
            }
        }
    } else {
        for (int64_t i = 0; i < li_rows; ++i) {
            const int32_t k = l_orderkey[i];
            if (k < 0 || k >= map_size) continue;
            const int32_t row = map[k];
            if (row < 0 || row >= o_rows) continue;
            if (end[row] == 0) {
                start[row] = (uint32_t)i;
                end[row] = (uint32_t)i + 1;
            } else {
                if ((uint32_t)i < start[row]) start[row] = (uint32_t)i;
                if ((uint32_t)i + 1 > end[row]) end[row] = (uint32_t)i + 1;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Q14: bespoke_tpch/query_q14.cpp:176-365.  partkey -> promo flag over      */
/* 0..max part key (-1 absent; the last part row with a key wins), then every */
/* lineitem row with ship_lo <= ship < ship_hi whose part is known adds its   */
/* discounted_price to the total and, for promo parts, to the promo sum.      */
/* promo_revenue = 100 * promo / total (0 when total <= 0).                   */
/* ------------------------------------------------------------------------ */
void oracle_q14(int64_t li_rows, const int16_t* ship, const int32_t* partkey,
                const int32_t* disc_price, int32_t ship_lo, int32_t ship_hi, int64_t p_rows,
                const int32_t* p_partkey, const uint8_t* p_promo, oracle_q14_out* out) {
    memset(out, 0, sizeof(*out));
    const int32_t mp = max_i32(p_partkey, p_rows);
    if (mp < 0) return;
    const int64_t size = (int64_t)mp + 1;
    int8_t* flag = (int8_t*)malloc((size_t)size);
    memset(flag, -1, (size_t)size);
    for (int64_t r = 0; r < p_rows; ++r)
        if (p_partkey[r] >= 0) flag[p_partkey[r]] = (int8_t)(p_promo[r] ? 1 : 0);
    int64_t promo = 0, total = 0, emitted = 0;
    for (int64_t i = 0; i < li_rows; ++i) {
        /* planted mistake 12: an inclusive end date */
        if (ship[i] < ship_lo || (oracle_mutation() == 12 ? ship[i] > ship_hi : ship[i] >= ship_hi))
            continue;
        const int32_t k = partkey[i];
        if (k < 0 || k >= size) continue;
        const int8_t f = flag[k];
        if (f < 0) continue;
        total += disc_price[i];
        if (f) promo += disc_price[i];
        emitted += 1;
    }
    free(flag);
    out->promo_sum = promo;
    out->total_sum = total;
    out->rows_emitted = emitted;
    out->promo_revenue = total > 0 ? 100.0 * (double)promo / (double)total : 0.0;
}

/* ------------------------------------------------------------------------ */
/* Q17: bespoke_tpch/query_q17.cpp:119-309.  Parts passing the predicates     */
/* set bits over 0..max part key; every lineitem row of such a part adds its  */
/* quantity to sum_qty[part] and 1 to cnt[part] and is kept; a kept row adds  */
/* its extendedprice when quantity * 5 * cnt < sum_qty (0.2 * average,       */
/* :288-291).                                                                 */
/* ------------------------------------------------------------------------ */
void oracle_q17(int64_t p_rows, const int32_t* p_partkey, int32_t n_preds,
                const oracle_pred* preds, int64_t li_rows, const int32_t* l_partkey,
                const int16_t* l_qty, const int32_t* l_price, oracle_q17_out* out) {
    memset(out, 0, sizeof(*out));
    const int32_t mp = max_i32(p_partkey, p_rows);
    if (mp < 0) return;
    const int64_t size = (int64_t)mp + 1;
    uint8_t* match = (uint8_t*)calloc((size_t)size, 1);
    int64_t parts = 0;
    for (int64_t r = 0; r < p_rows; ++r) {
        const int32_t k = p_partkey[r];
        if (k < 0 || k > mp) continue;
        int ok = 1;
        for (int32_t j = 0; j < n_preds; ++j) ok = ok && oracle_pred_row(&preds[j], r);
        if (ok) {
            match[k] = 1;
            parts += 1;
        }
    }
    out->parts = parts;
    if (parts == 0) {
        free(match);
        return;
    }
    int64_t* sum = (int64_t*)calloc((size_t)size, sizeof(int64_t));
    int32_t* cnt = (int32_t*)calloc((size_t)size, sizeof(int32_t));
    int64_t* kept = (int64_t*)malloc(sizeof(int64_t) * (size_t)(li_rows > 0 ? li_rows : 1));
    int64_t nk = 0;
    for (int64_t i = 0; i < li_rows; ++i) {
        const uint32_t k = (uint32_t)l_partkey[i];
        if (k >= (uint64_t)size || !match[k]) continue;
        sum[k] += l_qty[i];
        cnt[k] += 1;
        kept[nk++] = i;
    }
    int64_t ext = 0, emitted = 0;
    for (int64_t j = 0; j < nk; ++j) {
        const int64_t i = kept[j];
        const int32_t k = l_partkey[i];
        const int64_t lhs = (int64_t)l_qty[i] * 5 * cnt[k];
        /* planted mistake 13: quantity <= 0.2 * average */
        if (oracle_mutation() == 13 ? lhs > sum[k] : lhs >= sum[k]) continue;
        ext += l_price[i];
        emitted += 1;
    }
    free(match);
    free(sum);
    free(cnt);
    free(kept);
    out->matches = nk;
    out->rows_emitted = emitted;
    out->sum_extendedprice = ext;
}

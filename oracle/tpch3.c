/*
 * tpch3.c — CPU restatement of the reference's supplier-side TPC-H pipelines
 * Q2, Q11, Q15 and Q20 over the reference Database layout (see oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's cpu_baseline leg
 * as the checker / timed CPU baseline.  The product library never links or
 * calls anything here.  Reference paths are relative to the
 * DataManagementLab/BespokeOLAP_Artifacts root.
 *
 * Layout assumed (bespoke_tpch/builder_impl.hpp:120-140, builder_impl.cpp:
 * 718-778): partsupp stably sorted by partkey; supplier rows in table order;
 * supplycost and acctbal in cents; partsupp.supplycost_availqty the int32
 * product of the two columns.  String work (type suffixes, name prefixes,
 * names in the output) is the caller's: it arrives here as per-row or
 * per-dictionary-entry flags, as the reference decides it once per string.
 *
 * Where the reference indexes an array with a key it never checked (Q11's
 * supp_in_nation, Q15's revenue_by_suppkey, Q20's colour bitset), a key
 * outside the array is treated as absent; the reference's behaviour there is
 * undefined.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Q2 (bespoke_tpch/query_q2.cpp:146-405): a part row qualifies when its     */
/* size equals `size` and type_ok[type code] (type ends with TYPE, decided   */
/* per dictionary entry at :219-231).  Its key is taken as row + 1 (:295).   */
/* Its partsupp rows are the run of that key in the sorted table (the        */
/* stride shortcut at :288-312 and the lower_bound at :313-324 find the same */
/* run).  Rows whose supplier row suppkey - 1 is outside the table or not in */
/* the region (s_ok) are skipped; the others keep the minimum supplycost and */
/* every row at that minimum, in row order (:331-357).  Each kept row is     */
/* emitted as (part row, supplier row) in part-row order.  Returns the count  */
/* (rows beyond max_out are counted, not written).                           */
/* ------------------------------------------------------------------------ */
static int64_t lower_bound_i32(const int32_t* v, int64_t n, int32_t x) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (v[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

int64_t oracle_q2(int64_t p_rows, const int32_t* p_size, const uint16_t* p_type,
                  const uint8_t* type_ok, int32_t size, int64_t ps_rows, const int32_t* ps_partkey,
                  const int32_t* ps_suppkey, const int32_t* ps_supplycost, int64_t s_rows,
                  const uint8_t* s_ok, int32_t* out_prow, int32_t* out_srow, int64_t max_out) {
    int64_t n = 0;
    for (int64_t prow = 0; prow < p_rows; ++prow) {
        if (p_size[prow] != size || !type_ok[p_type[prow]]) continue;
        const int32_t key = (int32_t)(prow + 1);
        const int64_t s = lower_bound_i32(ps_partkey, ps_rows, key);
        int64_t e = s;
        while (e < ps_rows && ps_partkey[e] == key) ++e;
        int32_t min_cost = 0x7fffffff;
        for (int64_t r = s; r < e; ++r) {
            const uint32_t srow = (uint32_t)(ps_suppkey[r] - 1);
            if (srow >= (uint64_t)s_rows || (!s_ok[srow] && oracle_mutation() != 27)) continue;
            if (ps_supplycost[r] < min_cost) min_cost = ps_supplycost[r];
        }
        if (min_cost == 0x7fffffff) continue;
        for (int64_t r = s; r < e; ++r) {
            const uint32_t srow = (uint32_t)(ps_suppkey[r] - 1);
            if (srow >= (uint64_t)s_rows || (!s_ok[srow] && oracle_mutation() != 27)) continue;
            if (ps_supplycost[r] != min_cost) continue;
            if (n < max_out) {
                out_prow[n] = (int32_t)prow;
                out_srow[n] = (int32_t)srow;
            }
            ++n;
        }
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* Q11 (bespoke_tpch/query_q11.cpp:119-217): suppliers of nation `nation`    */
/* mark their key (keys 0..max supplier key).  Every run of equal partkeys   */
/* in partsupp row order sums supplycost_availqty over its rows whose        */
/* supplier is marked; a run with a non-zero sum is a group (:163-186).       */
/* total = the sum over every marked row; a group is kept when               */
/* (double)sum > (double)total * fraction (:194-203).  The kept groups are   */
/* sorted by value descending, then partkey (:207-214).  Returns the count   */
/* of kept groups; *total_out = total; *groups_out = groups formed.          */
/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t partkey;
    int64_t value;
} q11_group;

static int cmp_q11(const void* pa, const void* pb) {
    const q11_group* a = (const q11_group*)pa;
    const q11_group* b = (const q11_group*)pb;
    if (a->value != b->value) return a->value > b->value ? -1 : 1;
    return (a->partkey > b->partkey) - (a->partkey < b->partkey);
}

int64_t oracle_q11(int64_t s_rows, const int32_t* s_suppkey, const int32_t* s_nationkey,
                   int32_t nation, int64_t ps_rows, const int32_t* ps_partkey,
                   const int32_t* ps_suppkey, const int32_t* ps_value, double fraction,
                   int32_t* out_partkey, int64_t* out_value, int64_t max_out, int64_t* total_out,
                   int64_t* groups_out) {
    int32_t smax = 0;
    for (int64_t i = 0; i < s_rows; ++i)
        if (i == 0 || s_suppkey[i] > smax) smax = s_suppkey[i];
    const int64_t size = s_rows > 0 ? (int64_t)smax + 1 : 1;
    uint8_t* mark = (uint8_t*)calloc((size_t)(size > 0 ? size : 1), 1);
    for (int64_t i = 0; i < s_rows; ++i)
        if (s_nationkey[i] == nation && s_suppkey[i] >= 0 && s_suppkey[i] < size)
            mark[s_suppkey[i]] = 1;
    q11_group* g = (q11_group*)malloc(sizeof(q11_group) * (size_t)(ps_rows > 0 ? ps_rows : 1));
    int64_t ng = 0, total = 0;
    for (int64_t i = 0; i < ps_rows;) {
        const int32_t key = ps_partkey[i];
        int64_t sum = 0;
        for (; i < ps_rows && ps_partkey[i] == key; ++i) {
            const int32_t sk = ps_suppkey[i];
            if (sk < 0 || sk >= size || !mark[sk]) continue;
            sum += ps_value[i];
            total += ps_value[i];
        }
        if (sum != 0 || oracle_mutation() == 28) g[ng++] = (q11_group){key, sum};
    }
    const double threshold = (double)total * fraction;
    int64_t nk = 0;
    for (int64_t j = 0; j < ng; ++j)
        if ((double)g[j].value > threshold) g[nk++] = g[j];
    qsort(g, (size_t)nk, sizeof(q11_group), cmp_q11);
    for (int64_t j = 0; j < nk && j < max_out; ++j) {
        out_partkey[j] = g[j].partkey;
        out_value[j] = g[j].value;
    }
    *total_out = total;
    *groups_out = ng;
    free(mark);
    free(g);
    return nk;
}

/* ------------------------------------------------------------------------ */
/* Q15 (bespoke_tpch/query_q15.cpp:215-534, the loop without shards): every  */
/* lineitem with (uint32)(ship - date_lo) < (uint32)(date_hi - date_lo)      */
/* adds discounted_price to revenue[suppkey] (suppkeys 0..max supplier key,  */
/* at least one slot).  No row in the window, or no positive revenue: no     */
/* result.  Otherwise every suppkey (ascending) whose revenue equals the     */
/* largest positive revenue and that has a supplier row (the last row of the  */
/* key, :478-490) is a result (suppkey, row, revenue).  Returns the count;   */
/* *emitted_out = lineitems in the window.                                   */
/* ------------------------------------------------------------------------ */
int64_t oracle_q15(int64_t li_rows, const int16_t* l_shipdate, const int32_t* l_suppkey,
                   const int32_t* l_disc_price, int32_t date_lo, int32_t date_hi,
                   int64_t s_rows, const int32_t* s_suppkey, int32_t* out_suppkey,
                   int32_t* out_row, int64_t* out_revenue, int64_t max_out,
                   int64_t* emitted_out) {
    int32_t smax = -1;
    for (int64_t i = 0; i < s_rows; ++i)
        if (s_suppkey[i] > smax) smax = s_suppkey[i];
    const int64_t size = smax + 1 > 1 ? (int64_t)smax + 1 : 1;
    int64_t* rev = (int64_t*)calloc((size_t)size, sizeof(int64_t));
    const uint32_t span = (uint32_t)(date_hi - date_lo)
                          + (oracle_mutation() == 29 ? 1u : 0u);
    int64_t emitted = 0;
    for (int64_t i = 0; i < li_rows; ++i) {
        if ((uint32_t)((int32_t)l_shipdate[i] - date_lo) >= span) continue;
        ++emitted;
        const uint32_t sk = (uint32_t)l_suppkey[i];
        if (sk < (uint64_t)size) rev[sk] += l_disc_price[i];
    }
    *emitted_out = emitted;
    int64_t n = 0;
    if (emitted > 0) {
        int64_t mx = 0;
        for (int64_t k = 0; k < size; ++k)
            if (rev[k] > mx) mx = rev[k];
        if (mx > 0) {
            int32_t* row = (int32_t*)malloc(sizeof(int32_t) * (size_t)size);
            for (int64_t k = 0; k < size; ++k) row[k] = -1;
            for (int64_t i = 0; i < s_rows; ++i)
                if (s_suppkey[i] >= 0 && s_suppkey[i] < size) row[s_suppkey[i]] = (int32_t)i;
            for (int64_t k = 0; k < size; ++k) {
                if (rev[k] != mx || row[k] < 0) continue;
                if (n < max_out) {
                    out_suppkey[n] = (int32_t)k;
                    out_row[n] = row[k];
                    out_revenue[n] = rev[k];
                }
                ++n;
            }
            free(row);
        }
    }
    free(rev);
    return n;
}

/* ------------------------------------------------------------------------ */
/* Q20 (bespoke_tpch/query_q20.cpp:331-620, the loop without shards): part   */
/* rows with p_ok_row (name starts with COLOR) and key 0..max part key mark   */
/* their key.  Lineitems with (uint32)(ship - date_lo) < (uint32)(date_hi -   */
/* date_lo), supp_nationkey == nation and a marked partkey sum quantity per   */
/* (partkey, suppkey).  A partsupp row with a marked partkey, a supplier of   */
/* `nation` (supp_nation_by_key, keys below its size) and a sum for its pair  */
/* qualifies its suppkey (<= max_suppkey) when availqty * 2 * 100 > sum.      */
/* qualified[] (max_suppkey + 1 entries) is written; returns how many keys    */
/* qualified.  *pairs_out = (partkey, suppkey) pairs summed.                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint64_t key;
    int64_t qty;
} q20_item;

static int cmp_q20(const void* pa, const void* pb) {
    const uint64_t a = ((const q20_item*)pa)->key, b = ((const q20_item*)pb)->key;
    return (a > b) - (a < b);
}

int64_t oracle_q20(int64_t p_rows, const int32_t* p_partkey, const uint8_t* p_ok_row,
                   int64_t li_rows, const int16_t* l_shipdate, const int32_t* l_partkey,
                   const int32_t* l_suppkey, const int16_t* l_supp_nation,
                   const int16_t* l_quantity, int32_t date_lo, int32_t date_hi, int32_t nation,
                   int64_t ps_rows, const int32_t* ps_partkey, const int32_t* ps_suppkey,
                   const int32_t* ps_availqty, const int16_t* supp_nation_by_key,
                   int64_t n_supp_map, int32_t max_suppkey, uint8_t* qualified,
                   int64_t* pairs_out) {
    int32_t pmax = -1;
    for (int64_t i = 0; i < p_rows; ++i)
        if (p_partkey[i] > pmax) pmax = p_partkey[i];
    *pairs_out = 0;
    if (max_suppkey >= 0) memset(qualified, 0, (size_t)max_suppkey + 1);
    if (pmax < 0) return 0;
    uint8_t* color = (uint8_t*)calloc((size_t)pmax + 1, 1);
    for (int64_t i = 0; i < p_rows; ++i)
        if (p_partkey[i] >= 0 && p_ok_row[i]) color[p_partkey[i]] = 1;
    const uint32_t span = (uint32_t)(date_hi - date_lo);
    const int64_t factor = oracle_mutation() == 30 ? 1 : 2;
    q20_item* it = (q20_item*)malloc(sizeof(q20_item) * (size_t)(li_rows > 0 ? li_rows : 1));
    int64_t n = 0;
    for (int64_t i = 0; i < li_rows; ++i) {
        if ((uint32_t)((int32_t)l_shipdate[i] - date_lo) >= span) continue;
        if (l_supp_nation[i] != (int16_t)nation) continue;
        const uint32_t pk = (uint32_t)l_partkey[i];
        if (pk > (uint32_t)pmax || !color[pk]) continue;
        it[n].key = ((uint64_t)pk << 32) | (uint32_t)l_suppkey[i];
        it[n].qty = l_quantity[i];
        ++n;
    }
    qsort(it, (size_t)n, sizeof(q20_item), cmp_q20);
    int64_t m = 0;  /* merge equal keys */
    for (int64_t i = 0; i < n; ++i) {
        if (m > 0 && it[m - 1].key == it[i].key) it[m - 1].qty += it[i].qty;
        else it[m++] = it[i];
    }
    *pairs_out = m;
    int64_t count = 0;
    for (int64_t j = 0; j < ps_rows; ++j) {
        const uint32_t pk = (uint32_t)ps_partkey[j];
        if (pk > (uint32_t)pmax || !color[pk]) continue;
        const uint32_t sk = (uint32_t)ps_suppkey[j];
        if (sk >= (uint64_t)n_supp_map || supp_nation_by_key[sk] != (int16_t)nation) continue;
        const uint64_t key = ((uint64_t)pk << 32) | sk;
        int64_t lo = 0, hi = m;
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo) / 2;
            if (it[mid].key < key) lo = mid + 1;
            else hi = mid;
        }
        if (lo == m || it[lo].key != key) continue;
        if ((int64_t)ps_availqty[j] * factor * 100 > it[lo].qty && max_suppkey >= 0 &&
            sk <= (uint32_t)max_suppkey && !qualified[sk]) {
            qualified[sk] = 1;
            ++count;
        }
    }
    free(color);
    free(it);
    return count;
}

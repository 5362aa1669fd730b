/*
 * tpch2.c — CPU restatement of the reference's Q4, Q5, Q10, Q12 and Q19
 * pipelines over the reference Database layout (see oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's cpu_baseline leg
 * as the checker / timed CPU baseline.  The product library never links or
 * calls anything here.  Reference paths are relative to the
 * DataManagementLab/BespokeOLAP_Artifacts root; every function follows the
 * reference loop it cites, including its orderkey_sorted / unsorted split.
 *
 * Layout assumed (bespoke_tpch/builder_impl.hpp:39-84): orders sorted by
 * orderdate (builder_impl.cpp:597-608); lineitem_ranges[o] = [start, end) of
 * order row o's lineitem rows, {0, 0} for none (:929-991); with unsorted
 * lineitem the ranges are min/max spans and the queries check each row's
 * orderkey.  Dictionary codes are passed as uint16 like DictionaryColumn.
 */
#include "oracle.h"

#include <stdlib.h>

#include <string.h>

/* lower_bound over a non-decreasing int16 column (std::lower_bound). */
static int64_t lower_bound_i16(const int16_t* v, int64_t n, int32_t x) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if ((int32_t)v[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

/* ------------------------------------------------------------------------ */
/* Q4 (bespoke_tpch/query_q4.cpp:140-268): orders with orderdate in          */
/* [date_lo, date_hi) (lower_bound on the sorted dates) that have a lineitem  */
/* with commitdate < receiptdate, compared as the packed uint16 halves of     */
/* commit_receipt (q4_is_late, :70-72), counted per orderpriority code.       */
/* ------------------------------------------------------------------------ */
void oracle_q4(int64_t o_rows, const int32_t* o_orderkey, const int16_t* o_orderdate,
               const uint16_t* o_priority, int32_t n_priority, const uint32_t* r_start,
               const uint32_t* r_end, const int32_t* l_orderkey, int32_t l_sorted,
               const uint32_t* l_commit_receipt, int32_t date_lo, int32_t date_hi,
               int64_t* counts) {
    for (int32_t c = 0; c < n_priority; ++c) counts[c] = 0;
    const int64_t s = lower_bound_i16(o_orderdate, o_rows, date_lo);
    const int64_t e = lower_bound_i16(o_orderdate, o_rows, date_hi);
    for (int64_t o = s; o < e; ++o) {
        if (r_end[o] == 0) continue;
        int late = 0;
        for (uint32_t r = r_start[o]; r < r_end[o]; ++r) {
            if (!l_sorted && l_orderkey[r] != o_orderkey[o]) continue;
            const uint32_t p = l_commit_receipt[r];
            const int is_late = oracle_mutation() == 15 ? (uint16_t)p <= (uint16_t)(p >> 16)
                                                        : (uint16_t)p < (uint16_t)(p >> 16);
            if (is_late) {
                late = 1;
                break;
            }
        }
        if (late && o_priority[o] < n_priority) counts[o_priority[o]] += 1;
    }
}

/* ------------------------------------------------------------------------ */
/* Q5 (bespoke_tpch/query_q5.cpp:190-440): orders with orderdate in           */
/* [date_lo, date_hi) whose customer nation (orders.cust_nationkey,           */
/* builder_impl.cpp:913-928) is in the region; per such order the sum of      */
/* discounted_price over its lineitems whose supplier nation                   */
/* (lineitem.supp_nationkey, :838-853) equals the customer's, added to that   */
/* nation.  revenue[n] for n < n_nation (max nation key + 1); a customer      */
/* nation outside [0, n_nation) counts as not in the region.                  */
/* ------------------------------------------------------------------------ */
void oracle_q5(int64_t o_rows, const int32_t* o_orderkey, const int16_t* o_orderdate,
               const int16_t* o_cust_nation, const uint32_t* r_start, const uint32_t* r_end,
               const int32_t* l_orderkey, int32_t l_sorted, const int16_t* l_supp_nation,
               const int32_t* l_disc_price, const uint8_t* nation_in_region, int32_t n_nation,
               int32_t date_lo, int32_t date_hi, int64_t* revenue) {
    for (int32_t n = 0; n < n_nation; ++n) revenue[n] = 0;
    const int64_t s = lower_bound_i16(o_orderdate, o_rows, date_lo);
    const int64_t e = lower_bound_i16(o_orderdate, o_rows, date_hi);
    for (int64_t o = s; o < e; ++o) {
        const int32_t cn = o_cust_nation[o];
        if (cn < 0 || cn >= n_nation || !nation_in_region[cn]) continue;
        if (r_end[o] == 0) continue;
        int64_t local = 0;
        for (uint32_t r = r_start[o]; r < r_end[o]; ++r) {
            if (!l_sorted && l_orderkey[r] != o_orderkey[o]) continue;
            const int same = oracle_mutation() == 16 ? 1 : l_supp_nation[r] == cn;
            if (!same) continue;
            local += l_disc_price[r];
        }
        revenue[cn] += local;
    }
}

/* ------------------------------------------------------------------------ */
/* Q10 (bespoke_tpch/query_q10.cpp:375-560): orders with orderdate in         */
/* [date_lo, date_hi) and a customer key in [0, max customer key]; per order  */
/* the sum of discounted_price over its lineitems whose returnflag_linestatus */
/* code c has (u32)(c - code_lo) < code_span (returnflag 'R'); orders with a  */
/* nonzero sum are grouped by customer; a group is kept when its customer     */
/* has a row (custkey_to_row, last row of a key wins) whose nation key is in  */
/* [0, n_nation_names).  Rows ordered by the reference's LSD passes: revenue */
/* descending as an unsigned 64-bit value (~revenue ascending), ties by       */
/* customer key.  Returns the number of groups.                               */
/* ------------------------------------------------------------------------ */
static int cmp_q10(const void* pa, const void* pb) {
    const oracle_q10_row* a = (const oracle_q10_row*)pa;
    const oracle_q10_row* b = (const oracle_q10_row*)pb;
    const uint64_t ra = (uint64_t)a->revenue, rb = (uint64_t)b->revenue;
    if (ra != rb) return ra > rb ? -1 : 1;
    if (a->custkey != b->custkey) return a->custkey < b->custkey ? -1 : 1;
    return 0;
}

int64_t oracle_q10(int64_t o_rows, const int32_t* o_orderkey, const int16_t* o_orderdate,
                   const int32_t* o_custkey, const uint32_t* r_start, const uint32_t* r_end,
                   const int32_t* l_orderkey, int32_t l_sorted, const uint8_t* l_code,
                   int32_t code_lo, int32_t code_span, const int32_t* l_disc_price,
                   int32_t date_lo, int32_t date_hi, int64_t c_rows, const int32_t* c_custkey,
                   const int32_t* c_nationkey, int32_t n_nation_names, oracle_q10_row* out,
                   int64_t max_out) {
    if (code_lo < 0) return 0;  /* no 'R' in the returnflag dictionary */
    int32_t max_ck = 0;         /* q10::max_value: 0 for an empty table */
    for (int64_t i = 0; i < c_rows; ++i)
        if (i == 0 || c_custkey[i] > max_ck) max_ck = c_custkey[i];
    const int64_t size = (int64_t)max_ck + 1;
    int32_t* row_of = (int32_t*)malloc(sizeof(int32_t) * (size_t)(size > 0 ? size : 1));
    for (int64_t k = 0; k < size; ++k) row_of[k] = -1;
    for (int64_t i = 0; i < c_rows; ++i)
        if (c_custkey[i] >= 0) row_of[c_custkey[i]] = (int32_t)i;
    uint64_t* rev = (uint64_t*)calloc((size_t)(size > 0 ? size : 1), sizeof(uint64_t));
    uint8_t* has = (uint8_t*)calloc((size_t)(size > 0 ? size : 1), 1);
    const int64_t s = lower_bound_i16(o_orderdate, o_rows, date_lo);
    const int64_t e = lower_bound_i16(o_orderdate, o_rows, date_hi);
    for (int64_t o = s; o < e; ++o) {
        const int32_t ck = o_custkey[o];
        if (ck < 0 || ck >= size) continue;
        if (r_end[o] == 0) continue;
        int64_t order_rev = 0;
        for (uint32_t r = r_start[o]; r < r_end[o]; ++r) {
            if (!l_sorted && l_orderkey[r] != o_orderkey[o]) continue;
            const uint32_t off = (uint32_t)((int32_t)l_code[r] - code_lo);
            if (oracle_mutation() == 17 ? off > (uint32_t)code_span : off >= (uint32_t)code_span)
                continue;
            order_rev += l_disc_price[r];
        }
        if (order_rev != 0) {
            rev[ck] += (uint64_t)order_rev;
            has[ck] = 1;
        }
    }
    int64_t n = 0;
    for (int64_t k = 0; k < size; ++k) {
        if (!has[k]) continue;
        const int32_t row = row_of[k];
        if (row < 0) continue;
        const int32_t nk = c_nationkey[row];
        if (nk < 0 || nk >= n_nation_names) continue;
        if (n < max_out) {
            out[n].custkey = (int32_t)k;
            out[n].cust_row = row;
            out[n].revenue = (int64_t)rev[k];
        }
        ++n;
    }
    free(row_of);
    free(rev);
    free(has);
    const int64_t kept = n < max_out ? n : max_out;
    qsort(out, (size_t)kept, sizeof(oracle_q10_row), cmp_q10);
    return n;
}

/* ------------------------------------------------------------------------ */
/* Q12 (bespoke_tpch/query_q12.cpp:150-250, the loop without shards): a       */
/* lineitem counts when its shipmode code is mode1 or mode2, commitdate <     */
/* receiptdate and shipdate < commitdate (the packed halves as int16), and    */
/* (u32)(receiptdate - date_lo) < (u32)(date_hi - date_lo); its order row is  */
/* orderkey_to_row[orderkey] (keys outside the map or without a row are       */
/* skipped), and it adds to high[slot] when that order's priority code is     */
/* high1 or high2, else to low[slot] (slot 0 for mode1, 1 for mode2).         */
/* The reference's shard loop prunes shards whose shipdates all precede       */
/* date_lo - 30; for data with receiptdate <= shipdate + 30 (TPC-H) both      */
/* loops count the same rows.                                                  */
/* ------------------------------------------------------------------------ */
void oracle_q12(int64_t li_rows, const uint16_t* l_shipmode, const uint32_t* l_commit_receipt,
                const int16_t* l_shipdate, const int32_t* l_orderkey, int32_t mode1,
                int32_t mode2, int32_t date_lo, int32_t date_hi, const int32_t* order_map,
                int64_t map_size, const uint16_t* o_priority, int32_t high1, int32_t high2,
                int64_t* high, int64_t* low) {
    high[0] = high[1] = low[0] = low[1] = 0;
    const uint32_t span = (uint32_t)(date_hi - date_lo);
    for (int64_t i = 0; i < li_rows; ++i) {
        const int32_t m = l_shipmode[i];
        if (m != mode1 && m != mode2) continue;
        const uint32_t p = l_commit_receipt[i];
        const int32_t commit = (int16_t)(p & 0xffff);
        const int32_t receipt = (int16_t)(p >> 16);
        if (commit >= receipt) continue;
        if ((int32_t)l_shipdate[i] >= commit) continue;
        if ((uint32_t)(receipt - date_lo) >= span) continue;
        const int32_t k = l_orderkey[i];
        if ((uint32_t)k >= (uint64_t)map_size) continue;
        const int32_t row = order_map[k];
        if (row < 0) continue;
        const int32_t pr = o_priority[row];
        const int slot = m == mode1 ? 0 : 1;
        const int is_high = pr == high1 || pr == high2;
        if (oracle_mutation() == 18 ? !is_high : is_high) high[slot] += 1;
        else low[slot] += 1;
    }
}

/* ------------------------------------------------------------------------ */
/* Q19 (bespoke_tpch/query_q19.cpp:130-365).  Part scan: for every part row   */
/* in order, mask[partkey] = OR over the enabled groups g of (brand ==        */
/* brand[g] and container in containers[g] and 1 <= size <= size_hi[g]) << g  */
/* (a plain assignment, so the last row of a duplicated key wins, even with  */
/* mask 0).  Lineitem scan: shipmode in {mode1, mode2} (the codes >= 0),     */
/* shipinstruct == instruct, mask[partkey] != 0 and the quantity in          */
/* [qty_lo[g], qty_lo[g] + 1000] for one of the mask's groups; revenue sums  */
/* extendedprice * (100 - discount) exactly.  Returns the revenue.            */
/* ------------------------------------------------------------------------ */
int64_t oracle_q19(int64_t p_rows, const int32_t* p_partkey, const uint16_t* p_brand,
                   const uint16_t* p_container, const int32_t* p_size, int64_t li_rows,
                   const int32_t* l_partkey, const int16_t* l_quantity, const int32_t* l_price,
                   const uint8_t* l_disc, const uint16_t* l_shipmode,
                   const uint16_t* l_shipinstruct, const oracle_q19_group* groups, int32_t mode1,
                   int32_t mode2, int32_t instruct) {
    const int any = groups[0].enabled || groups[1].enabled || groups[2].enabled;
    if (!any) return 0;
    int32_t max_key = -1;
    for (int64_t i = 0; i < p_rows; ++i)
        if (p_partkey[i] > max_key) max_key = p_partkey[i];
    for (int64_t i = 0; i < li_rows; ++i)
        if (l_partkey[i] > max_key) max_key = l_partkey[i];
    if (max_key < 0) return 0;
    uint8_t* mask = (uint8_t*)calloc((size_t)max_key + 1, 1);
    for (int64_t i = 0; i < p_rows; ++i) {
        uint8_t m = 0;
        for (int g = 0; g < 3; ++g) {
            const oracle_q19_group* G = &groups[g];
            if (!G->enabled || (int32_t)p_brand[i] != G->brand) continue;
            int in = 0;
            for (int c = 0; c < 4; ++c) in |= G->containers[c] >= 0 &&
                                             (int32_t)p_container[i] == G->containers[c];
            if (!in) continue;
            if (p_size[i] < 1 || p_size[i] > G->size_hi) continue;
            m |= (uint8_t)(1u << g);
        }
        if (p_partkey[i] >= 0) mask[p_partkey[i]] = m;
    }
    int64_t revenue = 0;
    if (instruct >= 0 && (mode1 >= 0 || mode2 >= 0)) {
        for (int64_t i = 0; i < li_rows; ++i) {
            const int32_t sm = l_shipmode[i];
            if (sm != mode1 && sm != mode2) continue;
            if ((int32_t)l_shipinstruct[i] != instruct) continue;
            const int32_t pk = l_partkey[i];
            if (pk < 0) continue;
            const uint8_t m = mask[pk];
            if (!m) continue;
            const int32_t q = l_quantity[i];
            int ok = 0;
            for (int g = 0; g < 3; ++g) {
                const int32_t hi = groups[g].qty_lo + (oracle_mutation() == 19 ? 999 : 1000);
                if ((m >> g) & 1) ok |= q >= groups[g].qty_lo && q <= hi;
            }
            if (!ok) continue;
            revenue += (int64_t)l_price[i] * (int64_t)(100 - (int32_t)l_disc[i]);
        }
    }
    free(mask);
    return revenue;
}

/* ------------------------------------------------------------------------ */
/* Q21 (bespoke_tpch/query_q21.cpp:140-380): every order with status code    */
/* final_code and a range of more than one row; over its lineitems (with    */
/* the key check when unsorted) whose suppkey is in [0, n_suppkeys): the    */
/* order has several suppliers when a suppkey differs from the first, and   */
/* a row is late when receiptdate > commitdate (packed halves, unsigned);   */
/* when late rows come from two suppliers the order is dropped (the walk    */
/* stops there).  An order with several suppliers and exactly one late      */
/* supplier adds its late-row count to that supplier when the supplier is   */
/* in the target nation (is_target[suppkey]).  wait[n_suppkeys] out.        */
/* ------------------------------------------------------------------------ */
void oracle_q21(int64_t o_rows, const int32_t* o_orderkey, const uint16_t* o_status,
                int32_t final_code, const uint32_t* r_start, const uint32_t* r_end,
                const int32_t* l_orderkey, int32_t l_sorted, const int32_t* l_suppkey,
                const uint32_t* l_commit_receipt, const uint8_t* is_target, int64_t n_suppkeys,
                int64_t* wait) {
    for (int64_t k = 0; k < n_suppkeys; ++k) wait[k] = 0;
    for (int64_t o = 0; o < o_rows; ++o) {
        if ((int32_t)o_status[o] != final_code) continue;
        if (r_end[o] == 0) continue;
        if (r_end[o] - r_start[o] <= 1) continue;
        int32_t first = -1, late_supp = -1, late_n = 0;
        int multi = 0, multi_late = 0;
        for (uint32_t r = r_start[o]; r < r_end[o]; ++r) {
            if (!l_sorted && l_orderkey[r] != o_orderkey[o]) continue;
            const int32_t s = l_suppkey[r];
            if ((uint32_t)s >= (uint64_t)n_suppkeys) continue;
            if (first < 0) first = s;
            else if (s != first) multi = 1;
            const uint32_t p = l_commit_receipt[r];
            const int late = oracle_mutation() == 20 ? (p >> 16) >= (p & 0xffffu)
                                                     : (p >> 16) > (p & 0xffffu);
            if (late) {
                if (late_supp == -1 || late_supp == s) {
                    late_supp = s;
                    late_n += 1;
                } else {
                    multi_late = 1;
                    break;
                }
            }
        }
        if (multi_late || !multi || late_supp < 0) continue;
        if (!is_target[late_supp]) continue;
        wait[late_supp] += late_n;
    }
}

/* ------------------------------------------------------------------------ */
/* Q22 (bespoke_tpch/query_q22.cpp:190-520): a customer is a candidate when */
/* its phone prefix is in the filter — a two-digit prefix code (phone_code, */
/* 255 when the phone does not start with two digits, builder_impl.cpp:     */
/* 152-178) with numeric_allowed[code], or else its first two characters    */
/* (phone_pair = c0 << 8 | c1, -1 for phones shorter than two) equal one of */
/* the extra prefixes — and acctbal > 0.  avg = floor(sum / count) of the   */
/* candidates' acctbal; candidates with acctbal > avg and no orders count   */
/* per prefix: numeric[2 * code] (customers), numeric[2 * code + 1]          */
/* (acctbal sum), extra[2 * i], extra[2 * i + 1].  A customer has orders    */
/* when an order carries its key (the reference's orders_by_customer         */
/* offsets, builder_impl.cpp:625-668; a key no order carries has none).      */
/* ------------------------------------------------------------------------ */
void oracle_q22(int64_t c_rows, const int32_t* c_custkey, const uint8_t* c_phone_code,
                const int32_t* c_phone_pair, const int32_t* c_acctbal,
                const uint8_t* numeric_allowed, int32_t n_extra, const int32_t* extra_pairs,
                int64_t o_rows, const int32_t* o_custkey, int64_t* numeric, int64_t* extra) {
    for (int i = 0; i < 200; ++i) numeric[i] = 0;
    for (int i = 0; i < 2 * n_extra; ++i) extra[i] = 0;
    int32_t omax = -1;
    for (int64_t i = 0; i < o_rows; ++i)
        if (o_custkey[i] > omax) omax = o_custkey[i];
    uint8_t* has = (uint8_t*)calloc((size_t)(omax + 2), 1);
    for (int64_t i = 0; i < o_rows; ++i)
        if (o_custkey[i] >= 0) has[o_custkey[i]] = 1;
    int64_t sum = 0, cnt = 0;
    int32_t* slot = (int32_t*)malloc(sizeof(int32_t) * (size_t)(c_rows > 0 ? c_rows : 1));
    for (int64_t i = 0; i < c_rows; ++i) {
        slot[i] = -1;
        const uint8_t code = c_phone_code[i];
        if (code != 255) {
            if (code < 100 && numeric_allowed[code]) slot[i] = code;
        } else if (c_phone_pair[i] >= 0) {
            for (int32_t e = 0; e < n_extra; ++e)
                if (extra_pairs[e] == c_phone_pair[i]) {
                    slot[i] = 100 + e;
                    break;
                }
        }
        if (slot[i] < 0) continue;
        if (c_acctbal[i] <= 0) {
            slot[i] = -1;
            continue;
        }
        sum += c_acctbal[i];
        cnt += 1;
    }
    if (cnt > 0) {
        const int64_t avg = sum / cnt;
        for (int64_t i = 0; i < c_rows; ++i) {
            if (slot[i] < 0) continue;
            if (c_acctbal[i] <= avg) continue;
            const int32_t k = c_custkey[i];
            if (oracle_mutation() != 21 && k >= 0 && k <= omax && has[k]) continue;
            int64_t* agg = slot[i] < 100 ? numeric + 2 * slot[i] : extra + 2 * (slot[i] - 100);
            agg[0] += 1;
            agg[1] += c_acctbal[i];
        }
    }
    free(has);
    free(slot);
}

/* ------------------------------------------------------------------------ */
/* Q7 (bespoke_tpch/query_q7.cpp:150-330).  Supplier tags: a supplier row   */
/* with nation n1 tags its key 0, else (n1 != n2) with nation n2 tags it 1; */
/* other rows leave the tag as it is (-1 at first).  Over every order whose */
/* customer nation is n1 (index 0) or, when n1 != n2, n2 (index 1) and whose */
/* range is not empty: lineitems with shipdate in [ship_lo, ship_hi]        */
/* (inclusive) and a tagged supplier of the other index (both 0 when        */
/* n1 == n2) add discounted_price to rev[supp * 4 + cust * 2 + year], year  */
/* = shipdate >= split.  Lineitem suppkeys outside the tag array count as   */
/* untagged.                                                                 */
/* ------------------------------------------------------------------------ */
void oracle_q7(int64_t o_rows, const int32_t* o_orderkey, const int16_t* o_cust_nation,
               const uint32_t* r_start, const uint32_t* r_end, const int32_t* l_orderkey,
               int32_t l_sorted, const int16_t* l_shipdate, const int32_t* l_suppkey,
               const int32_t* l_disc_price, int64_t s_rows, const int32_t* s_suppkey,
               const int32_t* s_nationkey, int32_t n1, int32_t n2, int32_t ship_lo,
               int32_t ship_hi, int32_t split, int64_t* rev) {
    for (int i = 0; i < 8; ++i) rev[i] = 0;
    const int same = n1 == n2;
    int32_t smax = -1;
    for (int64_t i = 0; i < s_rows; ++i)
        if (s_suppkey[i] > smax) smax = s_suppkey[i];
    int8_t* tag = (int8_t*)malloc((size_t)(smax + 1 > 0 ? smax + 1 : 1));
    for (int32_t k = 0; k <= smax; ++k) tag[k] = -1;
    for (int64_t i = 0; i < s_rows; ++i) {
        if (s_suppkey[i] < 0) continue;
        if (s_nationkey[i] == n1) tag[s_suppkey[i]] = 0;
        else if (!same && s_nationkey[i] == n2) tag[s_suppkey[i]] = 1;
    }
    for (int64_t o = 0; o < o_rows; ++o) {
        const int32_t cn = o_cust_nation[o];
        int ci;
        if (cn == n1) ci = 0;
        else if (!same && cn == n2) ci = 1;
        else continue;
        if (r_end[o] <= r_start[o]) continue;
        for (uint32_t r = r_start[o]; r < r_end[o]; ++r) {
            if (!l_sorted && l_orderkey[r] != o_orderkey[o]) continue;
            const int32_t ship = l_shipdate[r];
            if ((uint32_t)(ship - ship_lo) > (uint32_t)(ship_hi - ship_lo)) continue;
            const int32_t sk = l_suppkey[r];
            const int si = (sk >= 0 && sk <= smax) ? tag[sk] : -1;
            if (si < 0) continue;
            if (oracle_mutation() != 22 && (same ? (si != 0 || ci != 0) : si == ci)) continue;
            const int y = ship >= split;
            rev[si * 4 + ci * 2 + y] += l_disc_price[r];
        }
    }
    free(tag);
}

/* ------------------------------------------------------------------------ */
/* Q8 (bespoke_tpch/query_q8.cpp:150-330): customers whose nation is in the */
/* region mark their key, suppliers of the target nation mark theirs, parts */
/* of type code type_code mark theirs (marks are never cleared).  Orders of */
/* [date_lo, date_hi] (inclusive; lower_bound / upper_bound on the sorted   */
/* dates) with a marked customer: every lineitem of a marked part adds its  */
/* volume price * (100 - discount) to total[year] and, for a marked         */
/* supplier, to nation[year]; year = orderdate >= split.                     */
/* ------------------------------------------------------------------------ */
static uint8_t* key_marks(int64_t rows, const int32_t* keys, int32_t* size_out) {
    int32_t mx = -1;
    for (int64_t i = 0; i < rows; ++i)
        if (keys[i] > mx) mx = keys[i];
    *size_out = mx + 1;
    return (uint8_t*)calloc((size_t)(mx + 1 > 0 ? mx + 1 : 1), 1);
}

void oracle_q8(int64_t o_rows, const int32_t* o_orderkey, const int16_t* o_orderdate,
               const int32_t* o_custkey, const uint32_t* r_start, const uint32_t* r_end,
               const int32_t* l_orderkey, int32_t l_sorted, const int32_t* l_partkey,
               const int32_t* l_suppkey, const int32_t* l_price, const uint8_t* l_disc,
               int64_t c_rows, const int32_t* c_custkey, const int32_t* c_nationkey,
               const uint8_t* nation_in_region, int32_t n_nation, int64_t s_rows,
               const int32_t* s_suppkey, const int32_t* s_nationkey, int32_t target,
               int64_t p_rows, const int32_t* p_partkey, const uint16_t* p_type,
               int32_t type_code, int32_t date_lo, int32_t date_hi, int32_t split,
               int64_t* total, int64_t* nation) {
    total[0] = total[1] = nation[0] = nation[1] = 0;
    int32_t cn, sn, pn;
    uint8_t* cmark = key_marks(c_rows, c_custkey, &cn);
    uint8_t* smark = key_marks(s_rows, s_suppkey, &sn);
    uint8_t* pmark = key_marks(p_rows, p_partkey, &pn);
    for (int64_t i = 0; i < c_rows; ++i) {
        const int32_t nk = c_nationkey[i];
        if (c_custkey[i] >= 0 && nk >= 0 && nk < n_nation && nation_in_region[nk])
            cmark[c_custkey[i]] = 1;
    }
    for (int64_t i = 0; i < s_rows; ++i)
        if (s_suppkey[i] >= 0 && s_nationkey[i] == target) smark[s_suppkey[i]] = 1;
    for (int64_t i = 0; i < p_rows; ++i)
        if (p_partkey[i] >= 0 && type_code >= 0 && (int32_t)p_type[i] == type_code)
            pmark[p_partkey[i]] = 1;
    const int64_t s = lower_bound_i16(o_orderdate, o_rows, date_lo);
    const int64_t e = lower_bound_i16(o_orderdate, o_rows, date_hi + 1);
    for (int64_t o = s; o < e; ++o) {
        const int32_t ck = o_custkey[o];
        if (ck < 0 || ck >= cn || !cmark[ck]) continue;
        if (r_end[o] <= r_start[o]) continue;
        const int y = o_orderdate[o] >= split;
        for (uint32_t r = r_start[o]; r < r_end[o]; ++r) {
            if (!l_sorted && l_orderkey[r] != o_orderkey[o]) continue;
            const int32_t pk = l_partkey[r];
            if (pk < 0 || pk >= pn || !pmark[pk]) continue;
            const int64_t vol = (int64_t)l_price[r] * (100 - (int32_t)l_disc[r]);
            total[y] += vol;
            const int32_t sk = l_suppkey[r];
            const int hit = sk >= 0 && sk < sn && smark[sk];
            if (oracle_mutation() == 23 ? !hit : hit) nation[y] += vol;
        }
    }
    free(cmark);
    free(smark);
    free(pmark);
}

/* ------------------------------------------------------------------------ */
/* Q9 (bespoke_tpch/query_q9.cpp:150-400): parts whose name contains the    */
/* colour (p_ok_row, decided on the host as the reference does per name)    */
/* mark their key.  Every lineitem of a marked part finds its supply cost   */
/* in the first partsupp row (row order) with its partkey and suppkey, its  */
/* supplier nation (lineitem.supp_nationkey, >= 0) and its order's year     */
/* (the last order row of its key; rows without an order are skipped), and  */
/* adds discounted_price - supplycost * quantity / 100 (C division) to      */
/* totals[nation * year_span + year - min_year], marking used[] too.         */
/* min_year / year_span come from the earliest and latest orderdate.        */
/* ------------------------------------------------------------------------ */
static int32_t base_days(void) { return (int32_t)oracle_days_from_civil(1992, 1, 1); }

static int32_t year_of(int32_t offset) {
    int32_t y, m, d;
    oracle_civil_from_days(base_days() + offset, &y, &m, &d);
    return y;
}

int32_t oracle_q9(int64_t p_rows, const int32_t* p_partkey, const uint8_t* p_ok_row,
                  int64_t ps_rows, const int32_t* ps_partkey, const int32_t* ps_suppkey,
                  const int32_t* ps_supplycost, int64_t o_rows, const int32_t* o_orderkey,
                  const int16_t* o_orderdate, int64_t li_rows, const int32_t* l_partkey,
                  const int32_t* l_suppkey, const int32_t* l_orderkey, const int16_t* l_quantity,
                  const int32_t* l_disc_price, const int16_t* l_supp_nation, int32_t n_nation,
                  int64_t* totals, uint8_t* used, int32_t* min_year_out) {
    if (o_rows == 0) return 0;
    int32_t dmin = o_orderdate[0], dmax = o_orderdate[0];
    for (int64_t i = 0; i < o_rows; ++i) {
        if (o_orderdate[i] < dmin) dmin = o_orderdate[i];
        if (o_orderdate[i] > dmax) dmax = o_orderdate[i];
    }
    const int32_t y0 = year_of(dmin), span = year_of(dmax) - y0 + 1;
    *min_year_out = y0;
    for (int64_t i = 0; i < (int64_t)n_nation * span; ++i) {
        totals[i] = 0;
        used[i] = 0;
    }
    int32_t pn = 0;
    uint8_t* pok = key_marks(p_rows, p_partkey, &pn);
    for (int64_t i = 0; i < p_rows; ++i)
        if (p_partkey[i] >= 0 && p_ok_row[i]) pok[p_partkey[i]] = 1;
    /* partsupp rows grouped by partkey, in row order within a key */
    int32_t psn = 0;
    free(key_marks(ps_rows, ps_partkey, &psn));
    int64_t* ps_off = (int64_t*)calloc((size_t)psn + 2, sizeof(int64_t));
    int64_t* ps_idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ps_rows > 0 ? ps_rows : 1));
    for (int64_t j = 0; j < ps_rows; ++j)
        if (ps_partkey[j] >= 0) ps_off[ps_partkey[j] + 1] += 1;
    for (int32_t k = 0; k < psn; ++k) ps_off[k + 1] += ps_off[k];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * ((size_t)psn + 1));
    for (int32_t k = 0; k <= psn; ++k) fill[k] = ps_off[k];
    for (int64_t j = 0; j < ps_rows; ++j)
        if (ps_partkey[j] >= 0) ps_idx[fill[ps_partkey[j]]++] = j;
    free(fill);
    int32_t on = 0;
    uint8_t* dummy = key_marks(o_rows, o_orderkey, &on);
    free(dummy);
    int32_t* orow = (int32_t*)malloc(sizeof(int32_t) * (size_t)(on > 0 ? on : 1));
    for (int32_t k = 0; k < on; ++k) orow[k] = -1;
    for (int64_t i = 0; i < o_rows; ++i)
        if (o_orderkey[i] >= 0) orow[o_orderkey[i]] = (int32_t)i;
    for (int64_t i = 0; i < li_rows; ++i) {
        const int32_t pk = l_partkey[i];
        if (pk < 0 || pk >= pn || !pok[pk]) continue;
        const int32_t sk = l_suppkey[i];
        int64_t cost = 0;
        int found = 0;
        if (pk < psn)  /* first partsupp row of (pk, sk), in row order */
            for (int64_t j = ps_off[pk]; j < ps_off[pk + 1]; ++j)
                if (ps_suppkey[ps_idx[j]] == sk) {
                    cost = ps_supplycost[ps_idx[j]];
                    found = 1;
                    break;
                }
        if (!found) continue;
        const int32_t nk = l_supp_nation[i];
        if (nk < 0 || nk >= n_nation) continue;
        const int32_t ok = l_orderkey[i];
        if (ok < 0 || ok >= on || orow[ok] < 0) continue;
        const int32_t yi = year_of(o_orderdate[orow[ok]]) - y0;
        const int64_t div = oracle_mutation() == 24 ? 10 : 100;
        const int64_t profit = (int64_t)l_disc_price[i] - cost * (int64_t)l_quantity[i] / div;
        totals[(int64_t)nk * span + yi] += profit;
        used[(int64_t)nk * span + yi] = 1;
    }
    free(pok);
    free(orow);
    free(ps_off);
    free(ps_idx);
    return span;
}

/* ------------------------------------------------------------------------ */
/* Q13 (bespoke_tpch/query_q13.cpp:150-250): an order matches when WORD1    */
/* occurs in its comment and WORD2 occurs after that occurrence's end       */
/* (matches_pattern, :50-82: an empty word is found at the search start,   */
/* which must not be the comment's end).  Every order that does not match  */
/* adds one to its customer's count (keys in [0, max customer key]); the    */
/* result is dist[c] = customer rows with count c, c <= max count.  Returns */
/* the largest count.                                                        */
/* ------------------------------------------------------------------------ */
static const char* find_sub(const char* s, const char* e, const char* w, size_t n) {
    if (n == 0) return s;
    for (const char* p = s; p + n <= e; ++p)
        if (memcmp(p, w, n) == 0) return p;
    return e;
}

int32_t oracle_q13(int64_t o_rows, const int32_t* o_custkey, const uint8_t* chars,
                   const uint32_t* offsets, const char* w1, const char* w2, int64_t c_rows,
                   const int32_t* c_custkey, int64_t* dist, int64_t max_dist) {
    int32_t cmax = 0;
    for (int64_t i = 0; i < c_rows; ++i)
        if (i == 0 || c_custkey[i] > cmax) cmax = c_custkey[i];
    const int64_t size = (int64_t)cmax + 1;
    int32_t* cnt = (int32_t*)calloc((size_t)(size > 0 ? size : 1), sizeof(int32_t));
    const size_t n1 = strlen(w1), n2 = strlen(w2);
    for (int64_t o = 0; o < o_rows; ++o) {
        const char* s = (const char*)chars + offsets[o];
        const char* e = (const char*)chars + offsets[o + 1];
        const char* p1 = find_sub(s, e, w1, n1);
        int match = 0;
        if (p1 != e) {
            const char* p2 = find_sub(p1 + (oracle_mutation() == 25 ? 0 : n1), e, w2, n2);
            match = p2 != e;
        }
        if (match) continue;
        const int32_t k = o_custkey[o];
        if (k >= 0 && k < size) cnt[k] += 1;
    }
    int32_t mx = 0;
    for (int64_t k = 0; k < size; ++k)
        if (cnt[k] > mx) mx = cnt[k];
    for (int64_t c = 0; c < max_dist; ++c) dist[c] = 0;
    for (int64_t i = 0; i < c_rows; ++i) {
        const int32_t k = c_custkey[i];
        const int32_t c = (k >= 0 && k < size) ? cnt[k] : 0;
        if (c < max_dist) dist[c] += 1;
    }
    free(cnt);
    return mx;
}

/* ------------------------------------------------------------------------ */
/* Q16 (bespoke_tpch/query_q16.cpp:180-480): a part row is eligible when    */
/* its brand code is not banned_brand, its size is one of the sizes         */
/* (size_mask, sizes 0..50) and its type code is not banned (type_banned[], */
/* types starting with TYPE); its key (in [0, n_part_keys)) then belongs to */
/* group (brand, type, size) - the last eligible row of a key decides.      */
/* Every partsupp row of a grouped key whose supplier (keys in             */
/* [0, n_supp_keys) can be banned, supp_banned[]) is not banned adds its   */
/* supplier to the group's distinct set.  counts[(brand * n_types + type)   */
/* * 51 + size] = distinct suppliers; returns the number of groups with    */
/* suppliers.                                                                */
/* ------------------------------------------------------------------------ */
static int cmp_u64(const void* a, const void* b) {
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : x > y;
}

int64_t oracle_q16(int64_t p_rows, const int32_t* p_partkey, const uint16_t* p_brand,
                   const uint16_t* p_type, const int32_t* p_size, int32_t n_types,
                   int32_t banned_brand, const uint8_t* type_banned, uint64_t size_mask,
                   int64_t n_part_keys, int64_t ps_rows, const int32_t* ps_partkey,
                   const int32_t* ps_suppkey, const uint8_t* supp_banned, int64_t n_supp_keys,
                   int64_t* counts, int64_t n_counts) {
    int64_t* grp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n_part_keys > 0 ? n_part_keys : 1));
    for (int64_t k = 0; k < n_part_keys; ++k) grp[k] = -1;
    for (int64_t i = 0; i < p_rows; ++i) {
        const int32_t b = p_brand[i];
        if (b == banned_brand) continue;
        const int32_t sz = p_size[i];
        if (sz < 0 || sz > 50 || !((size_mask >> sz) & 1)) continue;
        if (type_banned[p_type[i]]) continue;
        const int32_t k = p_partkey[i];
        if (k < 0 || k >= n_part_keys) continue;
        grp[k] = ((int64_t)b * n_types + p_type[i]) * 51 + sz;
    }
    uint64_t* pairs = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(ps_rows > 0 ? ps_rows : 1));
    int64_t n = 0;
    for (int64_t j = 0; j < ps_rows; ++j) {
        const int32_t k = ps_partkey[j];
        if (k < 0 || k >= n_part_keys || grp[k] < 0) continue;
        const int32_t s = ps_suppkey[j];
        if (s >= 0 && s < n_supp_keys && supp_banned[s] && oracle_mutation() != 26) continue;
        pairs[n++] = ((uint64_t)grp[k] << 32) | (uint32_t)s;
    }
    qsort(pairs, (size_t)n, sizeof(uint64_t), cmp_u64);
    for (int64_t g = 0; g < n_counts; ++g) counts[g] = 0;
    int64_t groups = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (i > 0 && pairs[i] == pairs[i - 1]) continue;
        const int64_t g = (int64_t)(pairs[i] >> 32);
        if (g < n_counts) {
            if (counts[g] == 0) ++groups;
            counts[g] += 1;
        }
    }
    free(grp);
    free(pairs);
    return groups;
}

/*
 * oracle.h — CPU restatement of the reference pipelines.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker
 * (or the timed CPU baseline).  The product path (libbolap) never links,
 * loads or calls it.
 *
 * Every function is single-threaded plain C over host arrays and cites the
 * reference file:line it restates (paths relative to the reference root).
 * Integer accumulation uses unsigned 64/128-bit arithmetic so overflow
 * wraps exactly like the two's-complement sums of the GPU path.
 */
#ifndef BOLAP_ORACLE_H
#define BOLAP_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* --- seeded generators (CPU mirror of the device generators) ----------- */
uint64_t og_mix64(uint64_t z);
uint64_t og_hash(uint64_t seed, uint32_t stream, uint64_t row);
int64_t og_uniform(uint64_t h, int64_t lo, int64_t hi);

void og_lineitem(uint64_t seed, int64_t rows, int64_t row0, int32_t* ship, int64_t* qty,
                 int64_t* price, int64_t* disc, int64_t* tax, uint8_t* flag_a,
                 uint8_t* flag_b);
void og_dim(uint64_t seed, int64_t rows, int32_t* key, uint8_t* cat, uint8_t* cont);
void og_fact(uint64_t seed, int64_t rows, int64_t row0, int64_t dim_rows, int32_t* fk,
             int64_t* price, int64_t* disc);
void og_customers(uint64_t seed, int64_t rows, int32_t* key, uint8_t* segment);
int64_t og_orders(uint64_t seed, int64_t rows, int64_t cust_rows, int32_t* key,
                  int32_t* day, int32_t* cust, int64_t* fact_offsets);
void og_order_facts(uint64_t seed, int64_t order_rows, const int64_t* fact_offsets,
                    int32_t* fk_order, int64_t* price, int64_t* disc);

/* --- Q1: filter-scan + low-cardinality group-by ------------------------ */
typedef struct {
    int64_t sum_qty, sum_base_price, sum_disc_price;
    uint64_t sum_charge_lo;
    int64_t sum_charge_hi;
    int64_t sum_disc, count, reserved;
} oracle_q1_group;

void oracle_q1_wide(int64_t rows, const int32_t* ship, int64_t ship_lo, int64_t ship_hi,
                    const int64_t* qty, const int64_t* price, const int64_t* disc,
                    const int64_t* tax, const uint8_t* flag_a, const uint8_t* flag_b,
                    int32_t n_a, int32_t n_b, oracle_q1_group* out);
void oracle_q1_ref(int64_t rows, const int16_t* ship, int64_t ship_lo, int64_t ship_hi,
                   const int16_t* qty, const int32_t* price, const int32_t* disc_price,
                   const uint8_t* disc, const uint8_t* tax, const uint8_t* code,
                   int32_t n_groups, oracle_q1_group* out);

/* --- predicates ---------------------------------------------------------- */
typedef struct {
    const void* col;
    int32_t dtype;   /* 0 u8, 1 i16, 2 i32, 3 i64 */
    int32_t op;      /* 0 range, 1 in-set */
    int64_t lo, hi;
    uint64_t inline_set[4];
    const uint32_t* set;
    int64_t set_bits;
} oracle_pred;

int oracle_pred_row(const oracle_pred* p, int64_t row);

/* --- join + scalar sum (Q19 shape) -------------------------------------- */
typedef struct {
    int64_t revenue, build_rows, matches;
} oracle_join_sum_out;

void oracle_join_sum(int64_t dim_rows, const int32_t* dim_key, int32_t n_preds,
                     const oracle_pred* preds, int64_t fact_rows, const int32_t* fact_fk,
                     const void* fact_price, int32_t price_dtype, const void* fact_disc,
                     int32_t disc_dtype, oracle_join_sum_out* out);

/* --- chain join + group-by (Q3 shape) ------------------------------------ */
typedef struct {
    int32_t key;
    int32_t day;
    int64_t revenue;
} oracle_group_row;

typedef struct {
    int64_t cust_pass, order_pass, fact_matches, groups;
} oracle_chain_stats;

int64_t oracle_chain_groups(int64_t cust_rows, const int32_t* cust_key,
                            const uint8_t* cust_segment, int32_t segment, int64_t order_rows,
                            const int32_t* order_key, const int32_t* order_day,
                            const int32_t* order_cust, int32_t day_lt, int64_t fact_rows,
                            const int32_t* fact_order, const void* fact_price,
                            int32_t price_dtype, const void* fact_disc, int32_t disc_dtype,
                            const int32_t* fact_disc_price, const int16_t* fact_ship, int32_t ship_gt, oracle_group_row* out,
                            int64_t max_out, oracle_chain_stats* stats);
void oracle_sort_groups(oracle_group_row* rows, int64_t n);
int64_t oracle_q3_ranges(int64_t cust_rows, const int32_t* cust_key, const uint8_t* cust_segment,
                         int32_t segment, int64_t order_rows, const int32_t* order_key,
                         const int32_t* order_day, const int32_t* order_cust,
                         const uint32_t* range_start, const uint32_t* range_end, int32_t day_lt,
                         const int64_t* fact_price, const int64_t* fact_disc,
                         oracle_group_row* out, int64_t max_out);

/* --- radix sort / top-k --------------------------------------------------- */
void oracle_radix_sort_u64(uint64_t* keys, uint32_t* vals, int64_t n, int32_t key_bits);
void oracle_topk_u64(const uint64_t* keys, const uint32_t* vals, int64_t n, int32_t k,
                     uint64_t* keys_out, uint32_t* vals_out);

/* --- star join + group-by and estimator statistics (star.c) ------------- */
void og_uniform_i32(uint64_t seed, uint32_t stream, int64_t rows, int32_t lo, int32_t hi,
                    int32_t* out);
void og_uniform_u8(uint64_t seed, uint32_t stream, int64_t rows, int32_t lo, int32_t hi,
                   uint8_t* out);
void og_zipf_i32(uint64_t seed, uint32_t stream, int64_t rows, const int32_t* quantiles,
                 int64_t n_keys, uint64_t mul, int32_t* out);
void og_price_disc(uint64_t seed, uint32_t stream, int64_t rows, int64_t* price,
                   int64_t* disc);
void oracle_hll_i32(const int32_t* col, int64_t n, uint8_t* regs);

typedef struct {
    int64_t rows;
    const int32_t* key;
    int32_t key_min, key_max;
    int32_t n_preds;
    const oracle_pred* preds;
    const int32_t* fact_fk;
} oracle_star_dim;

typedef struct {
    int64_t dim_pass[3];
    int64_t matches, groups, revenue;
} oracle_star_stats;

int64_t oracle_star_groupby(int32_t n_dims, const oracle_star_dim* dims, int64_t fact_rows,
                            const int32_t* group_key, int32_t group_min, int32_t group_max,
                            const void* price, int32_t price_dtype, const void* disc,
                            int32_t disc_dtype, int32_t* keys_out, int64_t* rev_out,
                            int64_t max_out, oracle_star_stats* stats);
void oracle_star_sample_counts(int32_t n_dims, const oracle_star_dim* dims, int64_t fact_rows,
                               const int32_t* group_key, uint64_t seed, int64_t* counts);

/* --- CEB star queries (ceb.c) --------------------------------------------- */
typedef struct {
    int64_t title_rows;
    const int32_t *t_year, *t_kind, *t_code;
    int64_t ci_rows;
    const int32_t *ci_movie, *ci_role;
    int64_t mi_rows;
    const int32_t *mi_movie, *mi_type, *mi_info;
    int64_t mk_rows;
    const int32_t *mk_movie, *mk_keyword;
} oracle_ceb_db;

typedef struct {
    int32_t tables, year_lo, year_hi;
    uint32_t kind_mask, role_mask;
    const uint8_t* title_mask; int64_t title_n;  /* masks over codes / ids, or null */
    const uint8_t* info_mask; int64_t info_n;
    const uint8_t* type_mask; int64_t type_n;
    const uint8_t* kw_mask; int64_t kw_n;
} oracle_ceb_query;

/* out: count_lo, count_hi, min_year, title_pass, joined_movies */
void oracle_ceb_run(const oracle_ceb_db* db, const oracle_ceb_query* q, uint64_t* out);

/* Grouped MIN/MAX over join rows expanded through per-item ranges
   (bespoke_ceb/query_q10a.cpp:838-1107).  out: 5 int32 per row (group, key,
   has, min, max) in (group, key) order.  Returns the row count, -1 when a
   group, item, range or key is out of bounds, -2 when max_rows is short. */
int64_t oracle_group_minmax(int64_t pairs, const int32_t* pair_group, const int32_t* pair_item,
                            int64_t items, const int32_t* item_offset,
                            const int32_t* item_count, const uint8_t* item_has,
                            const int32_t* item_value, int64_t values,
                            const int32_t* range_value, int32_t groups, int32_t keys,
                            int64_t dense_max, int32_t* out, int64_t max_rows);

/* Row order of a table sorted ascending by (primary, secondary) (secondary
   may be null), ties in input order: sort_table_by_i32_column(s),
   bespoke_ceb/builder_impl.cpp:265-340 (std::sort leaves ties open). */
void oracle_sort_order_i32(const int32_t* primary, const int32_t* secondary, int64_t rows,
                           uint32_t* order);

/* --- TPC-H Q6 / Q18 and build-time derivations (tpch.c) ----------------- */
int64_t oracle_days_from_civil(int32_t y, int32_t m, int32_t d);
void oracle_civil_from_days(int32_t z, int32_t* y, int32_t* m, int32_t* d);

typedef struct {
    int64_t revenue, rows_emitted, rows_scanned;
} oracle_q6_out;

void oracle_q6_flat(int64_t rows, const int16_t* ship, const uint8_t* disc, const int16_t* qty,
                    const int32_t* price, int32_t ship_lo, int32_t ship_hi, int32_t disc_lo,
                    int32_t disc_hi, int32_t qty_lt, oracle_q6_out* out);
void oracle_q6_sharded(int64_t rows, const int16_t* ship, const uint8_t* disc,
                       const int16_t* qty, const int32_t* price, int32_t ship_lo,
                       int32_t ship_hi, int32_t disc_lo, int32_t disc_hi, int32_t qty_lt,
                       int32_t month_lo, int32_t month_hi, oracle_q6_out* out);

typedef struct {
    int32_t custkey, orderkey, orderdate, totalprice;
    int64_t sum_qty;
    int32_t cust_row, order_row;
} oracle_q18_row;

int64_t oracle_q18(int64_t li_rows, const int32_t* l_orderkey, const int16_t* l_qty,
                   int64_t o_rows, const int32_t* o_orderkey, const int32_t* o_custkey,
                   const int16_t* o_orderdate, const int32_t* o_totalprice, int64_t c_rows,
                   const int32_t* c_custkey, int32_t threshold, oracle_q18_row* out,
                   int64_t max_out, int32_t* dense_path);

int32_t oracle_derive_lineitem(int64_t rows, const int32_t* price, const uint8_t* disc,
                               const uint8_t* rf, const uint8_t* ls, int32_t n_ls,
                               const int32_t* orderkey, int32_t* disc_price, uint8_t* code);
void oracle_date_order(int64_t rows, const int16_t* date, uint32_t* order);
void oracle_order_ranges(int64_t o_rows, const int32_t* o_orderkey, int64_t map_size,
                         int32_t* map, int64_t li_rows, const int32_t* l_orderkey,
                         int32_t li_sorted, uint32_t* start, uint32_t* end);

typedef struct {
    int64_t promo_sum, total_sum, rows_emitted;
    double promo_revenue;
} oracle_q14_out;

void oracle_q14(int64_t li_rows, const int16_t* ship, const int32_t* partkey,
                const int32_t* disc_price, int32_t ship_lo, int32_t ship_hi, int64_t p_rows,
                const int32_t* p_partkey, const uint8_t* p_promo, oracle_q14_out* out);

typedef struct {
    int64_t sum_extendedprice, parts, matches, rows_emitted;
} oracle_q17_out;

void oracle_q17(int64_t p_rows, const int32_t* p_partkey, int32_t n_preds,
                const oracle_pred* preds, int64_t li_rows, const int32_t* l_partkey,
                const int16_t* l_qty, const int32_t* l_price, oracle_q17_out* out);

/* --- planted mistakes, for tests that show the checks have teeth -------- */
void oracle_set_mutation(int32_t m);
int32_t oracle_mutation(void);

#ifdef __cplusplus
}
#endif
/* tpch2.c: Q4, Q5, Q10, Q12, Q19 over the reference Database layout. */
typedef struct {
    int32_t custkey;
    int32_t cust_row;
    int64_t revenue;
} oracle_q10_row;

typedef struct {
    int32_t enabled;        /* QUANTITYg and BRANDg both given */
    int32_t brand;          /* brand code, -1 when not in the dictionary */
    int32_t containers[4];  /* container codes, -1 when not in the dictionary */
    int32_t size_hi;        /* 5, 10, 15 */
    int32_t qty_lo;         /* llround(QUANTITYg * 100) */
} oracle_q19_group;

void oracle_q4(int64_t o_rows, const int32_t* o_orderkey, const int16_t* o_orderdate,
               const uint16_t* o_priority, int32_t n_priority, const uint32_t* r_start,
               const uint32_t* r_end, const int32_t* l_orderkey, int32_t l_sorted,
               const uint32_t* l_commit_receipt, int32_t date_lo, int32_t date_hi,
               int64_t* counts);
void oracle_q5(int64_t o_rows, const int32_t* o_orderkey, const int16_t* o_orderdate,
               const int16_t* o_cust_nation, const uint32_t* r_start, const uint32_t* r_end,
               const int32_t* l_orderkey, int32_t l_sorted, const int16_t* l_supp_nation,
               const int32_t* l_disc_price, const uint8_t* nation_in_region, int32_t n_nation,
               int32_t date_lo, int32_t date_hi, int64_t* revenue);
int64_t oracle_q10(int64_t o_rows, const int32_t* o_orderkey, const int16_t* o_orderdate,
                   const int32_t* o_custkey, const uint32_t* r_start, const uint32_t* r_end,
                   const int32_t* l_orderkey, int32_t l_sorted, const uint8_t* l_code,
                   int32_t code_lo, int32_t code_span, const int32_t* l_disc_price,
                   int32_t date_lo, int32_t date_hi, int64_t c_rows, const int32_t* c_custkey,
                   const int32_t* c_nationkey, int32_t n_nation_names, oracle_q10_row* out,
                   int64_t max_out);
void oracle_q12(int64_t li_rows, const uint16_t* l_shipmode, const uint32_t* l_commit_receipt,
                const int16_t* l_shipdate, const int32_t* l_orderkey, int32_t mode1,
                int32_t mode2, int32_t date_lo, int32_t date_hi, const int32_t* order_map,
                int64_t map_size, const uint16_t* o_priority, int32_t high1, int32_t high2,
                int64_t* high, int64_t* low);
int64_t oracle_q19(int64_t p_rows, const int32_t* p_partkey, const uint16_t* p_brand,
                   const uint16_t* p_container, const int32_t* p_size, int64_t li_rows,
                   const int32_t* l_partkey, const int16_t* l_quantity, const int32_t* l_price,
                   const uint8_t* l_disc, const uint16_t* l_shipmode,
                   const uint16_t* l_shipinstruct, const oracle_q19_group* groups, int32_t mode1,
                   int32_t mode2, int32_t instruct);

void oracle_q21(int64_t o_rows, const int32_t* o_orderkey, const uint16_t* o_status,
                int32_t final_code, const uint32_t* r_start, const uint32_t* r_end,
                const int32_t* l_orderkey, int32_t l_sorted, const int32_t* l_suppkey,
                const uint32_t* l_commit_receipt, const uint8_t* is_target, int64_t n_suppkeys,
                int64_t* wait);
void oracle_q22(int64_t c_rows, const int32_t* c_custkey, const uint8_t* c_phone_code,
                const int32_t* c_phone_pair, const int32_t* c_acctbal,
                const uint8_t* numeric_allowed, int32_t n_extra, const int32_t* extra_pairs,
                int64_t o_rows, const int32_t* o_custkey, int64_t* numeric, int64_t* extra);

void oracle_q7(int64_t o_rows, const int32_t* o_orderkey, const int16_t* o_cust_nation,
               const uint32_t* r_start, const uint32_t* r_end, const int32_t* l_orderkey,
               int32_t l_sorted, const int16_t* l_shipdate, const int32_t* l_suppkey,
               const int32_t* l_disc_price, int64_t s_rows, const int32_t* s_suppkey,
               const int32_t* s_nationkey, int32_t n1, int32_t n2, int32_t ship_lo,
               int32_t ship_hi, int32_t split, int64_t* rev);
void oracle_q8(int64_t o_rows, const int32_t* o_orderkey, const int16_t* o_orderdate,
               const int32_t* o_custkey, const uint32_t* r_start, const uint32_t* r_end,
               const int32_t* l_orderkey, int32_t l_sorted, const int32_t* l_partkey,
               const int32_t* l_suppkey, const int32_t* l_price, const uint8_t* l_disc,
               int64_t c_rows, const int32_t* c_custkey, const int32_t* c_nationkey,
               const uint8_t* nation_in_region, int32_t n_nation, int64_t s_rows,
               const int32_t* s_suppkey, const int32_t* s_nationkey, int32_t target,
               int64_t p_rows, const int32_t* p_partkey, const uint16_t* p_type,
               int32_t type_code, int32_t date_lo, int32_t date_hi, int32_t split,
               int64_t* total, int64_t* nation);
int32_t oracle_q9(int64_t p_rows, const int32_t* p_partkey, const uint8_t* p_ok_row,
                  int64_t ps_rows, const int32_t* ps_partkey, const int32_t* ps_suppkey,
                  const int32_t* ps_supplycost, int64_t o_rows, const int32_t* o_orderkey,
                  const int16_t* o_orderdate, int64_t li_rows, const int32_t* l_partkey,
                  const int32_t* l_suppkey, const int32_t* l_orderkey, const int16_t* l_quantity,
                  const int32_t* l_disc_price, const int16_t* l_supp_nation, int32_t n_nation,
                  int64_t* totals, uint8_t* used, int32_t* min_year_out);

int32_t oracle_q13(int64_t o_rows, const int32_t* o_custkey, const uint8_t* chars,
                   const uint32_t* offsets, const char* w1, const char* w2, int64_t c_rows,
                   const int32_t* c_custkey, int64_t* dist, int64_t max_dist);
int64_t oracle_q16(int64_t p_rows, const int32_t* p_partkey, const uint16_t* p_brand,
                   const uint16_t* p_type, const int32_t* p_size, int32_t n_types,
                   int32_t banned_brand, const uint8_t* type_banned, uint64_t size_mask,
                   int64_t n_part_keys, int64_t ps_rows, const int32_t* ps_partkey,
                   const int32_t* ps_suppkey, const uint8_t* supp_banned, int64_t n_supp_keys,
                   int64_t* counts, int64_t n_counts);

/* tpch3.c: Q2, Q11, Q15, Q20 over the reference Database layout. */
int64_t oracle_q2(int64_t p_rows, const int32_t* p_size, const uint16_t* p_type,
                  const uint8_t* type_ok, int32_t size, int64_t ps_rows, const int32_t* ps_partkey,
                  const int32_t* ps_suppkey, const int32_t* ps_supplycost, int64_t s_rows,
                  const uint8_t* s_ok, int32_t* out_prow, int32_t* out_srow, int64_t max_out);
int64_t oracle_q11(int64_t s_rows, const int32_t* s_suppkey, const int32_t* s_nationkey,
                   int32_t nation, int64_t ps_rows, const int32_t* ps_partkey,
                   const int32_t* ps_suppkey, const int32_t* ps_value, double fraction,
                   int32_t* out_partkey, int64_t* out_value, int64_t max_out, int64_t* total_out,
                   int64_t* groups_out);
int64_t oracle_q15(int64_t li_rows, const int16_t* l_shipdate, const int32_t* l_suppkey,
                   const int32_t* l_disc_price, int32_t date_lo, int32_t date_hi,
                   int64_t s_rows, const int32_t* s_suppkey, int32_t* out_suppkey,
                   int32_t* out_row, int64_t* out_revenue, int64_t max_out,
                   int64_t* emitted_out);
int64_t oracle_q20(int64_t p_rows, const int32_t* p_partkey, const uint8_t* p_ok_row,
                   int64_t li_rows, const int16_t* l_shipdate, const int32_t* l_partkey,
                   const int32_t* l_suppkey, const int16_t* l_supp_nation,
                   const int16_t* l_quantity, int32_t date_lo, int32_t date_hi, int32_t nation,
                   int64_t ps_rows, const int32_t* ps_partkey, const int32_t* ps_suppkey,
                   const int32_t* ps_availqty, const int16_t* supp_nation_by_key,
                   int64_t n_supp_map, int32_t max_suppkey, uint8_t* qualified,
                   int64_t* pairs_out);

#endif

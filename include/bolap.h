/*
 * bolap.h — C ABI of the MI355X (gfx950) execution layer for the bespoke
 * per-query OLAP pipelines of DataManagementLab/BespokeOLAP_Artifacts.
 *
 * The reference runs each query as a single-threaded C++ loop nest behind
 *   std::vector<QNResultRow> run_qN(const Database&, const QNArgs&)
 * (e.g. bespoke_tpch/query_q1.hpp:22, query_q3.hpp:17, query_q19.hpp:14).
 * This header is the seam a re-implemented run_qN body calls instead of
 * its scalar loops: plain pointers and sizes, no C++ or torch types.
 *
 * Conventions
 *  - Every column pointer is a DEVICE pointer (HBM-resident column, as the
 *    device mirror of the reference Database holds it).  Columns read with
 *    128-bit loads must be 16-byte aligned; unaligned columns take a
 *    slower scalar path, results are identical.
 *  - Functions return 0 (BOLAP_OK) or a bolap_status; bolap_last_error()
 *    gives a thread-local message.  The C++ seam rethrows it as
 *    std::runtime_error, which is how the reference reports errors.
 *  - Capacity overflow of an estimate-sized hash table is NOT an error: the
 *    pipeline re-runs with exact sizing and reports it in `fallbacks`.
 *  - All work is issued on the context's stream.  Functions that return
 *    host results synchronise that stream before returning.
 */
#ifndef BOLAP_H
#define BOLAP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BOLAP_ABI_VERSION 1

typedef enum {
    BOLAP_OK = 0,
    BOLAP_ERR_ARG = 1,          /* bad argument (null pointer, size, dtype) */
    BOLAP_ERR_HIP = 2,          /* HIP runtime error */
    BOLAP_ERR_UNSUPPORTED = 3,  /* valid request this build does not cover */
    BOLAP_ERR_CAPACITY = 4,     /* a fixed maximum was exceeded */
    BOLAP_ERR_NO_DEVICE = 5,
    BOLAP_ERR_INTERNAL = 6      /* a self-check failed (e.g. a tuned variant disagreed) */
} bolap_status;

typedef enum {
    BOLAP_U8 = 0,
    BOLAP_I16 = 1,
    BOLAP_I32 = 2,
    BOLAP_I64 = 3
} bolap_dtype;

typedef enum {
    BOLAP_PLACE_AUTO = 0,
    BOLAP_PLACE_LDS = 1,   /* hash table staged per workgroup in LDS */
    BOLAP_PLACE_HBM = 2    /* open-addressing table in HBM, probed through L2 */
} bolap_placement;

/* ------------------------------------------------------------------------ */
/* Context: one per (host thread, device, stream).                            */
/* ------------------------------------------------------------------------ */
typedef struct bolap_ctx bolap_ctx;

typedef struct {
    int32_t device;
    int32_t compute_units;
    int32_t lds_bytes_per_cu;
    int32_t wavefront;
    int64_t hbm_bytes;
    char arch[32];
} bolap_device_info;

/* `stream` is a hipStream_t (NULL = the null stream). */
int bolap_ctx_create(int device, void* stream, bolap_ctx** out);
int bolap_ctx_destroy(bolap_ctx* ctx);
int bolap_ctx_sync(bolap_ctx* ctx);
int bolap_ctx_device_info(bolap_ctx* ctx, bolap_device_info* out);
const char* bolap_last_error(void);
int bolap_abi_version(void);

/* Stage tracing: the counterpart of the reference's -DTRACE output
   (bespoke_tpch/trace_utils.hpp:18-53, "PROFILE name ns" / "COUNT name n").
   When enabled, every pipeline stage records a pair of HIP events on the
   context stream and every pipeline records its row counts.  Reading the
   trace synchronises the stream, returns the entries in issue order and
   clears them. */
typedef struct {
    char name[40];
    double ms;        /* GPU time between the stage's events; -1 for counts */
    int64_t count;    /* row count for COUNT entries; 0 for timings */
} bolap_trace_entry;

int bolap_ctx_set_trace(bolap_ctx* ctx, int enabled);

/* Tuning knobs (kernel variant, tile size, workgroups per CU) read by the
   pipelines at launch; 0 or unset = the library's default choice.  The
   autotuner writes them; no reference counterpart (the reference hard-codes
   its constants, e.g. the 16-bit radix of bespoke_tpch/query_q10.cpp:139). */
int bolap_tune_set(bolap_ctx* ctx, const char* knob, int64_t value);
int bolap_tune_get(bolap_ctx* ctx, const char* knob, int64_t* value);
/* The knob table persisted as text, one "name value" line per knob ('#'
   starts a comment); load adds to (and overrides) the context's knobs and
   leaves them unchanged when a line is malformed. */
int bolap_tune_save(bolap_ctx* ctx, const char* path);
int bolap_tune_load(bolap_ctx* ctx, const char* path);
int bolap_ctx_read_trace(bolap_ctx* ctx, bolap_trace_entry* out, int32_t max_entries,
                         int32_t* n_out);

/* ------------------------------------------------------------------------ */
/* Predicates (dimension-side filters).                                       */
/* A row passes when every predicate holds (conjunction).                     */
/* ------------------------------------------------------------------------ */
typedef enum {
    BOLAP_OP_RANGE = 0,   /* lo <= v <= hi  (equality: lo == hi)            */
    BOLAP_OP_IN_SET = 1   /* bit v of the set is 1; v outside the set fails */
} bolap_op;

#define BOLAP_MAX_PREDS 4

typedef struct {
    const void* col;          /* device column */
    int32_t dtype;            /* bolap_dtype */
    int32_t op;               /* bolap_op */
    int64_t lo, hi;           /* BOLAP_OP_RANGE bounds, inclusive */
    uint64_t inline_set[4];   /* IN_SET over values 0..255 when `set` is NULL */
    const uint32_t* set;      /* device bitmap for IN_SET over 0..set_bits-1 */
    int64_t set_bits;
} bolap_pred;

/* ------------------------------------------------------------------------ */
/* Q1 shape: fused filter-scan + low-cardinality group-by aggregate.          */
/* Replaces the scan loop of run_q1 (bespoke_tpch/query_q1.cpp:104-228) and  */
/* its Q1Agg accumulators (query_q1.cpp:27-34).                               */
/* ------------------------------------------------------------------------ */
typedef enum {
    /* declared-width layout: ship i32, qty/price/disc/tax i64, flag_a and
       flag_b u8 dictionary codes; disc_price = price * (100 - disc) exact. */
    BOLAP_Q1_LAYOUT_WIDE = 0,
    /* reference Database layout (bespoke_tpch/builder_impl.hpp:39-66):
       shipdate i16, quantity i16 (x100), extendedprice i32,
       discounted_price i32 (truncated at build, builder_impl.cpp:523-532),
       discount u8, tax u8, returnflag_linestatus u8 group code.          */
    BOLAP_Q1_LAYOUT_REF = 1
} bolap_q1_layout;

typedef struct {
    int32_t layout;           /* bolap_q1_layout */
    int64_t rows;
    const void* ship;         /* filter column */
    int64_t ship_lo, ship_hi; /* row passes when ship_lo <= ship <= ship_hi */
    const void* qty;
    const void* price;
    const void* disc_price;   /* REF layout only */
    const void* disc;
    const void* tax;
    const uint8_t* flag_a;    /* WIDE: flag_a code; REF: combined group code */
    const uint8_t* flag_b;    /* WIDE only */
    int32_t n_a, n_b;         /* dictionary sizes; group = a * n_b + b   */
} bolap_q1_args;

/* Exact integer accumulators of one group (Q1Agg, query_q1.cpp:27-34).
   sum_charge is a two's-complement 128-bit integer (lo, hi). */
typedef struct {
    int64_t sum_qty;
    int64_t sum_base_price;
    int64_t sum_disc_price;
    uint64_t sum_charge_lo;
    int64_t sum_charge_hi;
    int64_t sum_disc;
    int64_t count;
    int64_t reserved;
} bolap_q1_group;

#define BOLAP_Q1_MAX_GROUPS 16

/* Host result: `out` holds n_a*n_b groups, indexed a*n_b+b. */
int bolap_q1_scan_agg(bolap_ctx* ctx, const bolap_q1_args* args, bolap_q1_group* out);
/* Stream-ordered variant: `out_dev` is a device array of n_a*n_b groups.
   No host synchronisation. */
int bolap_q1_scan_agg_async(bolap_ctx* ctx, const bolap_q1_args* args,
                            bolap_q1_group* out_dev);

/* Q1 result rows as run_q1 finalises them (query_q1.cpp:234-275): sums
   divided by their fixed-point scale, averages over the count, empty groups
   dropped, rows ordered by (returnflag, linestatus) strings.  groups holds
   n_rf * n_ls accumulators indexed rf * n_ls + ls; rf_dict / ls_dict are the
   dictionary strings by code.  Scales: REF layout quantity x100, prices and
   discounted_price in cents, charge in cents x hundredths; WIDE layout
   quantity in units, disc_price in cents x hundredths, charge x another 100.
   rows_out holds n_rf * n_ls rows.  Host only. */
typedef struct {
    char returnflag[8];
    char linestatus[8];
    double sum_qty, sum_base_price, sum_disc_price, sum_charge;
    double avg_qty, avg_price, avg_disc;
    int64_t count_order;
} bolap_q1_row;

int bolap_q1_finalize(const bolap_q1_group* groups, int32_t layout, int32_t n_rf,
                      int32_t n_ls, const char* const* rf_dict, const char* const* ls_dict,
                      bolap_q1_row* rows_out, int32_t* n_out);
/* write_q1_csv (query_q1.cpp:283-298): fixed notation, 2 decimals, avg_disc 4. */
int bolap_q1_write_csv(const char* path, const bolap_q1_row* rows, int32_t n);

/* ------------------------------------------------------------------------ */
/* Q19/Q17/Q14 shape: filtered dimension -> hash build -> fact probe ->       */
/* scalar SUM(price * (100 - disc)).                                          */
/* Replaces run_q19's part scan + lineitem probe                              */
/* (bespoke_tpch/query_q19.cpp:210-345) with a hash join on the key.          */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t dim_rows;
    const int32_t* dim_key;          /* unique keys */
    int32_t n_dim_preds;
    bolap_pred dim_preds[BOLAP_MAX_PREDS];
    int64_t fact_rows;
    const int32_t* fact_fk;
    const void* fact_price;          /* i64 (WIDE) or i32 (reference extendedprice) */
    int32_t price_dtype;
    const void* fact_disc;           /* i64 (WIDE) or u8 (reference discount) */
    int32_t disc_dtype;
    /* Sizing.  capacity_hint == 0: exact path (count pass, host reads the
       count, then build).  > 0: expected build rows from an estimator; the
       table is sized from it without a count pass and an overflow flag
       triggers the exact path. */
    int64_t capacity_hint;
    int32_t placement;               /* bolap_placement */
} bolap_join_sum_args;

typedef struct {
    int64_t revenue;        /* SUM(price * (100 - disc)), exact int64 */
    int64_t build_rows;     /* dimension rows that passed the filter */
    int64_t matches;        /* fact rows whose key found a build row */
    int64_t capacity;       /* hash slots of the table that produced revenue */
    int32_t placement;      /* BOLAP_PLACE_LDS or BOLAP_PLACE_HBM */
    int32_t fallbacks;      /* 1 when an estimate-sized table overflowed */
} bolap_join_sum_result;

int bolap_join_sum(bolap_ctx* ctx, const bolap_join_sum_args* args,
                   bolap_join_sum_result* out);

/* Autotuner of the join: times bolap_join_sum on these arguments under each
   probe loop (knob join.probe_variant 1..5), LDS filter bits per build key
   (join.lds_bloom_bits 6..8) and probe workgroups per CU (join.blocks_per_cu
   1..2), median of reps calls each, and keeps the fastest in the context's
   knobs; a candidate whose result differs from the first call's is an error
   (BOLAP_ERR_INTERNAL).  No reference counterpart (its probe loop and unroll
   are fixed, bespoke_tpch/query_q19.cpp:286-345). */
int bolap_join_autotune(bolap_ctx* ctx, const bolap_join_sum_args* args, int32_t reps,
                        double* best_ms);

/* Allocation-free variant of bolap_join_sum.  No counterpart in the
   reference: it is for callers that cannot allocate device or pinned memory
   while they run (code under HIP graph capture, tight serving loops,
   harnesses that forbid hipMalloc).  It runs the same kernels with the same
   exact, device-side sizing (capacity_hint is ignored) and returns the same
   result, but calls no hipMalloc, hipFree, hipHostMalloc or allocating
   library routine: its key list, hash table, filter, partials, counters and
   result row live in a fixed 64 MiB device array of the library (one copy
   per GPU, held until the process ends; its address is looked up by
   bolap_ctx_create).  A dimension whose worst-case build needs more than
   that (about 4 M rows) is rejected with BOLAP_ERR_CAPACITY.  The array
   keeps the last call's contents and every call on a GPU shares it: calls
   that overlap on one GPU (from several contexts or streams) race.  The
   result is copied back with a stream-ordered copy and a synchronisation
   of the context's stream. */
int bolap_join_sum_noalloc(bolap_ctx* ctx, const bolap_join_sum_args* args,
                           bolap_join_sum_result* out);

/* Q17: parts passing the predicates (brand / container codes) -> per-part
   SUM(quantity) and COUNT over their lineitem rows -> SUM(extendedprice) of
   the rows with quantity * 5 * count < SUM(quantity), i.e. below 0.2 times
   the part's average.  Replaces run_q17 (bespoke_tpch/query_q17.cpp:
   119-309); parts with negative keys are skipped as there (:168-171).  At
   most BOLAP_MAX_PREDS - 1 predicates. */
typedef struct {
    int64_t part_rows;
    const int32_t* p_partkey;
    int32_t n_part_preds;
    bolap_pred part_preds[BOLAP_MAX_PREDS];
    int64_t li_rows;
    const int32_t* l_partkey;
    const int16_t* l_quantity;       /* x 100 */
    const int32_t* l_price;          /* extendedprice, cents */
} bolap_q17_args;

typedef struct {
    int64_t sum_extendedprice;  /* exact; avg_yearly = this / 7 / 100 */
    int64_t parts;              /* part rows that passed */
    int64_t matches;            /* lineitem rows of passing parts */
    int64_t rows_emitted;       /* rows below the threshold */
} bolap_q17_result;

int bolap_q17(bolap_ctx* ctx, const bolap_q17_args* args, bolap_q17_result* out);

/* ------------------------------------------------------------------------ */
/* Q3 shape: customer filter -> orders filter + join -> fact join ->          */
/* GROUP BY order key SUM(price * (100 - disc)) -> ORDER BY revenue DESC,     */
/* day ASC, key ASC.  Replaces run_q3 (bespoke_tpch/query_q3.cpp:153-365);    */
/* the key tiebreak refines the reference's unspecified std::sort tie order  */
/* (query_q3.cpp:348-356).                                                    */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t cust_rows;
    const int32_t* cust_key;
    const uint8_t* cust_segment;
    int32_t segment;                 /* customer passes when segment == this */
    int64_t order_rows;
    const int32_t* order_key;        /* unique, >= 0 */
    const int32_t* order_day;
    const int32_t* order_cust;
    int32_t day_lt;                  /* order passes when day < day_lt */
    int64_t fact_rows;
    const int32_t* fact_order;
    const void* fact_price;          /* i64 or i32 */
    int32_t price_dtype;
    const void* fact_disc;           /* i64 or u8; ignored when disc_price set */
    int32_t disc_dtype;
    const int32_t* fact_disc_price;  /* optional: reference discounted_price
                                        (builder_impl.cpp:523-532); when set,
                                        revenue sums it instead */
    int64_t capacity_hint;           /* 0 = exact sizing; >0 = estimated groups */
    /* Optional, together: order row i's fact rows are exactly
       [order_fact_start[i], order_fact_end[i]) (the reference's
       lineitem_ranges, builder_impl.cpp:929-991, as bolap_order_ranges makes
       them; query_q3.cpp:219-260 reads the fact side through them).  The
       fact side is then read by ranges: fact_order and capacity_hint are
       unused and no group table is built. */
    const uint32_t* order_fact_start;
    const uint32_t* order_fact_end;
    /* Optional key ranges from the tables' statistics (a built database keeps
       per-column i32_min / i32_max, bespoke_ceb/builder_impl.hpp:100-101):
       with key_ranges_known = 1 the call skips its own min/max pass and the
       host round trip after it.  A customer or order key outside its range
       is an error (BOLAP_ERR_ARG), never a wrong result. */
    int32_t key_ranges_known;
    int32_t cust_key_min, cust_key_max;
    int32_t order_key_min, order_key_max;
    /* Optional fact-side filter, run_q3's l_shipdate > DATE (query_q3.cpp:
       236-241): with fact_ship set (int16 day offsets), only fact rows with
       fact_ship > ship_gt are summed and counted, and an order forms a group
       when one of its rows passes. */
    const int16_t* fact_ship;
    int32_t ship_gt;
    /* With order ranges only.  1: the fact rows are sorted by order key (the
       reference's lineitem.orderkey_sorted, which bolap_derive_lineitem
       reports), so every row of an order's range belongs to it.  0: the
       ranges may be min/max spans over unsorted rows (builder_impl.cpp:
       963-988), and rows whose fact_order differs from the order's key are
       skipped, as run_q3's unsorted branch does (query_q3.cpp:262-281);
       fact_order is then required. */
    int32_t fact_order_sorted;
} bolap_chain_args;

typedef struct {
    int32_t key;
    int32_t day;
    int64_t revenue;
} bolap_group_row;

typedef struct {
    int64_t cust_pass;
    int64_t order_pass;      /* orders that passed day + customer filter */
    int64_t fact_matches;    /* fact rows joined to a passing order */
    int64_t groups;          /* passing orders with >= 1 fact row */
    int64_t capacity;        /* group hash slots */
    int32_t fallbacks;
    int32_t sort_passes;     /* radix passes run by the ordering step */
} bolap_chain_stats;

/* Top-k rows in order; *n_out = min(k, groups). */
int bolap_chain_topk(bolap_ctx* ctx, const bolap_chain_args* args, int32_t k,
                     bolap_group_row* out, int32_t* n_out, bolap_chain_stats* stats);
/* All groups in order (the reference returns every group, query_q3.cpp:348).
   `out` must hold `max_rows`; *n_out = groups (error when > max_rows). */
int bolap_chain_groups_sorted(bolap_ctx* ctx, const bolap_chain_args* args,
                              bolap_group_row* out, int64_t max_rows, int64_t* n_out,
                              bolap_chain_stats* stats);

/* ------------------------------------------------------------------------ */
/* Star join + hash group-by: a fact table semi-joined with 1..3 filtered      */
/* dimensions, grouped on one fact column, SUM(price * (100 - disc)) per      */
/* group.  Shape of the CEB star queries (bespoke_ceb/query_q1a.cpp:241-621:  */
/* dimension masks, then a fact scan) and of the estimate-sized group maps    */
/* (bespoke_tpch/query_q20.cpp:132-223, 405-417; bespoke_ceb/query_q11b.cpp:  */
/* 273-362).  Sizing: with every hint 0 the pipeline counts first (each       */
/* dimension filter, then the join rows) and sizes from the counts; with      */
/* hints it skips the counting passes.  A hint that proves too small (a       */
/* table over half full, or an insert past the probe bound) reruns the whole  */
/* query with exact sizing and reports fallbacks = 1; results never change.   */
/* ------------------------------------------------------------------------ */
#define BOLAP_STAR_MAX_DIMS 3
#define BOLAP_STAR_LDS_KEYS 1024   /* largest key set staged in LDS per dimension */

typedef struct {
    int64_t rows;
    const int32_t* key;          /* unique keys within [key_min, key_max] */
    int32_t key_min, key_max;    /* key range from the table's statistics */
    int32_t n_preds;
    bolap_pred preds[BOLAP_MAX_PREDS];
    const int32_t* fact_fk;      /* fact column joined to key */
    int64_t pass_hint;           /* estimated passing rows; 0 = count exactly */
} bolap_star_dim;

typedef struct {
    int32_t n_dims;              /* 1..BOLAP_STAR_MAX_DIMS */
    bolap_star_dim dims[BOLAP_STAR_MAX_DIMS];
    int64_t fact_rows;
    const int32_t* group_key;    /* fact column grouped on */
    int32_t group_min, group_max;/* its range from the table's statistics */
    const void* fact_price;      /* BOLAP_I64 or BOLAP_I32 */
    int32_t price_dtype;
    const void* fact_disc;       /* BOLAP_I64 or BOLAP_U8 */
    int32_t disc_dtype;
    int64_t group_hint;          /* estimated group count; 0 = count join rows first */
} bolap_star_args;

typedef struct {
    int64_t dim_pass[BOLAP_STAR_MAX_DIMS];
    int32_t dim_placement[BOLAP_STAR_MAX_DIMS];  /* LDS (hash set) or HBM (bitmap) */
    int32_t group_placement;     /* LDS (per-workgroup pre-aggregation) or HBM */
    int64_t matches;             /* join rows */
    int64_t groups;
    int64_t group_capacity;
    int64_t revenue;             /* SUM over every group */
    int32_t fallbacks;
    int32_t count_passes;        /* counting passes that ran (0 when fully hinted) */
} bolap_star_stats;

/* Groups in ascending key order into keys_out / rev_out (device arrays of
   max_rows; both may be null to get the statistics only). */
int bolap_star_groupby(bolap_ctx* ctx, const bolap_star_args* args, int32_t* keys_out,
                       int64_t* rev_out, int64_t max_rows, int64_t* n_out,
                       bolap_star_stats* stats);

/* ------------------------------------------------------------------------ */
/* CEB-style star queries over movie ids (batched parameterised workload).   */
/* Restates the factorised COUNT of bespoke_ceb/query_q1a.cpp:69-152          */
/* (dimension masks), :241-596 (per-movie count arrays) and :602-621          */
/* (COUNT(*) = sum over qualifying movies of the product of the per-table     */
/* counts), with MIN(production_year) as in query_q10a.cpp:334 ff.  Fact      */
/* tables must be sorted by movie_id (the reference builder's order,          */
/* bespoke_ceb/builder_impl.cpp:432-452); title ids are 1..title_rows.        */
/* String predicates arrive as bitmaps over dictionary codes, computed on     */
/* the host (bolap_ilike_mask).                                               */
/* ------------------------------------------------------------------------ */
#define BOLAP_CEB_CI 1             /* cast_info takes part */
#define BOLAP_CEB_MI 2             /* movie_info (and info_type) */
#define BOLAP_CEB_MK 4             /* movie_keyword (and keyword) */
#define BOLAP_CEB_COUNT 0
#define BOLAP_CEB_MIN_YEAR 1

typedef struct {
    int64_t title_rows;
    const int32_t* t_year;       /* title.production_year */
    const int32_t* t_kind;       /* title.kind_id */
    const int32_t* t_code;       /* title.title dictionary code */
    int64_t ci_rows;
    const int32_t* ci_movie;     /* cast_info.movie_id (sorted) */
    const int32_t* ci_role;      /* cast_info.role_id */
    int64_t mi_rows;
    const int32_t* mi_movie;     /* movie_info.movie_id (sorted) */
    const int32_t* mi_type;      /* movie_info.info_type_id */
    const int32_t* mi_info;      /* movie_info.info dictionary code */
    int64_t mk_rows;
    const int32_t* mk_movie;     /* movie_keyword.movie_id (sorted) */
    const int32_t* mk_keyword;   /* movie_keyword.keyword_id */
} bolap_ceb_db;

typedef struct {
    int32_t tables;              /* BOLAP_CEB_CI | BOLAP_CEB_MI | BOLAP_CEB_MK, at least one */
    int32_t agg;                 /* BOLAP_CEB_COUNT or BOLAP_CEB_MIN_YEAR */
    int32_t year_lo, year_hi;    /* title.production_year in [lo, hi] */
    uint32_t kind_mask;          /* bit k: title.kind_id == k passes (k < 32) */
    uint32_t role_mask;          /* bit r: cast_info.role_id == r passes (r < 32) */
    const uint32_t* title_bits;  /* device bitmap over title codes, or null (no predicate) */
    int64_t title_bits_n;
    const uint32_t* info_bits;   /* device bitmap over movie_info.info codes, or null */
    int64_t info_bits_n;
    const uint32_t* type_bits;   /* device bitmap over info_type ids, or null */
    int64_t type_bits_n;
    const uint32_t* kw_bits;     /* device bitmap over keyword ids, or null */
    int64_t kw_bits_n;
} bolap_ceb_query;

typedef struct {
    uint64_t count_lo;           /* COUNT(*) as an unsigned 128-bit integer */
    uint64_t count_hi;
    int64_t min_year;            /* MIN(production_year); INT64_MAX when no row joins */
    int64_t title_pass;          /* movies passing the title predicates */
    int64_t joined_movies;       /* of those, movies with a match in every table */
} bolap_ceb_result;

/* One query, executed alone (the plain path). */
int bolap_ceb_run(bolap_ctx* ctx, const bolap_ceb_db* db, const bolap_ceb_query* q,
                  bolap_ceb_result* out);
/* A batch of queries through persistent stage kernels that share each table
   scan among the queries reading it, replayed from a captured hipGraph.
   out receives n results. */
int bolap_ceb_run_batch(bolap_ctx* ctx, const bolap_ceb_db* db, const bolap_ceb_query* qs,
                        int32_t n, bolap_ceb_result* out);
/* Times the batch under each candidate stage configuration and keeps the
   fastest in the context's knobs (ceb.blocks_per_cu, ceb.qchunk); best_ms
   (optional) receives its batch time. */
int bolap_ceb_autotune(bolap_ctx* ctx, const bolap_ceb_db* db, const bolap_ceb_query* qs,
                       int32_t n, double* best_ms);

/* ------------------------------------------------------------------------ */
/* CEB factorised COUNT over the reference's nullable columns.  Replaces the */
/* scans of run_q1a (bespoke_ceb/query_q1a.cpp:241-621), run_q2a            */
/* (query_q2a.cpp:240-503), run_q2b, run_q2c and run_q8a: a title row passes */
/* when its id is non-null and in [0, movie_cap), its production year is    */
/* non-null and in [year_lo, year_hi] (when t_year is set) and every title   */
/* predicate holds; a term row passes when its movie passed the title scan,  */
/* its movie column is non-null and every predicate holds, and adds `weight` */
/* to its movie's u32 counter.  count = sum over passing title rows (in row  */
/* order, duplicates included, as the reference loops over its title list)   */
/* of the product of that movie's counters, as wrapping 64-bit integers.     */
/* A predicate holds when the column value is non-null (is_null NULL: the    */
/* column has no nulls) and 0 <= v < mask_n with mask[v] != 0.  All columns  */
/* and masks are device arrays; columns 4-byte aligned.                      */
/* ------------------------------------------------------------------------ */
#define BOLAP_CEB_MAX_TERM_PREDS 4
#define BOLAP_CEB_MAX_TITLE_PREDS 2
#define BOLAP_CEB_MAX_TERMS 6

typedef struct {
    const int32_t* col;
    const uint8_t* is_null;
    const uint8_t* mask;
    int64_t mask_n;
} bolap_ceb_pred;

typedef struct {
    int64_t rows;
    const int32_t* movie;
    const uint8_t* movie_null;
    int32_t n_preds;
    uint32_t weight;            /* added per passing row (1 for COUNT(*)) */
    bolap_ceb_pred preds[BOLAP_CEB_MAX_TERM_PREDS];
} bolap_ceb_term;

typedef struct {
    int64_t title_rows;
    const int32_t* t_id;
    const uint8_t* t_id_null;
    const int32_t* t_year;      /* NULL: no year window */
    const uint8_t* t_year_null;
    int32_t year_lo, year_hi;
    int32_t n_title_preds;
    int32_t movie_cap;          /* above every title id (the id column's max + 1) */
    bolap_ceb_pred title_preds[BOLAP_CEB_MAX_TITLE_PREDS];
    int32_t n_terms;
    int32_t reserved;
    bolap_ceb_term terms[BOLAP_CEB_MAX_TERMS];
} bolap_ceb_count_args;

typedef struct {
    int64_t count;
    int64_t title_pass;         /* title rows that passed */
    int64_t joined;             /* of those, rows whose every counter is nonzero */
    int64_t term_rows[BOLAP_CEB_MAX_TERMS];   /* rows that passed, per term */
} bolap_ceb_count_result;

int bolap_ceb_count(bolap_ctx* ctx, const bolap_ceb_count_args* args,
                    bolap_ceb_count_result* out);

/* Host-side ILIKE over a string dictionary: mask[i] = 1 when string i
   (chars[offsets[i] .. offsets[i+1])) matches pattern, '%' matching any
   run of characters and '_' any one character, ASCII case-insensitive
   (bespoke_ceb/query_q10a.cpp:43-209). */
int bolap_ilike_mask(const char* chars, const int64_t* offsets, int64_t n,
                     const char* pattern, uint8_t* mask);

/* ------------------------------------------------------------------------ */
/* Grouped MIN/MAX of a nullable per-item value over join rows that expand  */
/* through per-item value ranges (CSR).  Restates the result aggregation of  */
/* bespoke_ceb/query_q10a.cpp:838-1107: each (person, movie) join row whose  */
/* person has a name group (pair_group >= 0; -1 drops the row, :1078-1080)    */
/* expands into its movie's info values (item_offset / item_count into        */
/* range_value, :1081-1090); every (group, info) pair is a group holding      */
/* {has, min, max} of the movie years, created as {0, 0, 0} by a movie        */
/* without a year (:880-900).  Placement as the reference: a dense matrix     */
/* when groups x keys <= dense_max_slots (0: 5,000,000, :843-856), else (the  */
/* reference's unordered_map) the expanded rows sorted by (group, key) and    */
/* reduced run by run.  Rows come out in (group, key) order (the reference's  */
/* order is first-seen or hash order, :1094-1120).  All pointers are device   */
/* pointers; a group, item, range or key outside its bounds is an error.     */
/* ------------------------------------------------------------------------ */
#define BOLAP_MINMAX_DENSE 0
#define BOLAP_MINMAX_HASH 1   /* beyond the dense limit (sort + run reduce) */

typedef struct {
    int64_t pairs;               /* join rows */
    const int32_t* pair_group;   /* group of the row's person, [0, groups) or -1 */
    const int32_t* pair_item;    /* the row's movie: index into the item arrays */
    int64_t items;               /* length of the item arrays */
    const int32_t* item_offset;  /* first range value of the item */
    const int32_t* item_count;   /* number of range values of the item */
    const uint8_t* item_has;     /* 1 when the item's value is present (null: all present) */
    const int32_t* item_value;   /* the aggregated value (production year) */
    int64_t values;              /* length of range_value */
    const int32_t* range_value;  /* second group key (dense info id), [0, keys) */
    int32_t groups;              /* first key domain (name groups) */
    int32_t keys;                /* second key domain (info ids) */
    int64_t dense_max_slots;     /* 0: 5,000,000 */
} bolap_minmax_args;

typedef struct {
    int32_t group, key;          /* the group */
    int32_t has;                 /* 1 when some row of the group had a value */
    int32_t min, max;            /* 0, 0 when has == 0 */
} bolap_minmax_row;

typedef struct {
    int64_t expanded_rows;       /* join rows x range lengths (rows aggregated) */
    int64_t groups;
    int32_t placement;           /* BOLAP_MINMAX_DENSE / BOLAP_MINMAX_HASH */
    int32_t pad;
    int64_t slots;               /* dense matrix slots (0 when hashed) */
} bolap_minmax_stats;

/* Rows into rows_out (device array of max_rows; null for the statistics
   only).  Synchronous: rows_out is complete when the call returns. */
int bolap_group_minmax(bolap_ctx* ctx, const bolap_minmax_args* args,
                       bolap_minmax_row* rows_out, int64_t max_rows, int64_t* n_out,
                       bolap_minmax_stats* stats);

/* ------------------------------------------------------------------------ */
/* CEB query 10a (bespoke_ceb/query_q10a.cpp:334-1120) over device columns: */
/* the title, movie_info, cast_info and name passes of run_q10a and its      */
/* MIN/MAX(production_year) per (name, info) through the aggregation above.  */
/* The caller resolves the IN lists as the reference does (kind / role /     */
/* info strings through the string pool, :211-265, and the ID list, :211-   */
/* 225) into masks indexed like the reference's arrays (kind_type, role_type */
/* and info_type rows + 1 entries, ids outside them fail), a dense info     */
/* index per string id, and the name ILIKE as a byte per string id.  Rows   */
/* carry string ids: group = the name's, key = the info's; in the order the  */
/* reference emits them when the (name, info) matrix is dense (<=           */
/* dense_max_slots, 0: 5,000,000), else in (name rank, dense info) order.    */
/* All pointers are device pointers except rows_out (host, max_rows rows).  */
/* ------------------------------------------------------------------------ */
typedef struct {
    const int32_t* v;            /* values (string ids for string columns) */
    const uint8_t* null;         /* 1 = null; NULL: the column has no nulls */
} bolap_ceb_col;

typedef struct {
    int64_t title_rows;
    bolap_ceb_col t_id, t_kind_id, t_year;
    int64_t mi_rows;
    bolap_ceb_col mi_movie, mi_type, mi_info;
    int64_t ci_rows;
    bolap_ceb_col ci_movie, ci_person, ci_role;
    int64_t name_rows;
    bolap_ceb_col n_id, n_name;
    const uint8_t* kind_ok;  int64_t kind_n;   /* kind_type rows + 1 entries */
    const uint8_t* role_ok;  int64_t role_n;   /* role_type rows + 1 entries */
    const uint8_t* type_ok;  int64_t type_n;   /* info_type rows + 1 entries */
    int64_t pool_n;                            /* strings in the pool */
    const int32_t* info_dense;   /* pool_n entries: dense index of an INFO value, else -1 */
    const int32_t* info_ids;     /* n_infos entries: string id of each dense index */
    const uint8_t* name_ok;      /* pool_n entries: 1 when the name matches; NULL: all */
    int32_t n_infos;
    int32_t reserved;
    int64_t dense_max_slots;
} bolap_ceb_q10a_args;

int bolap_ceb_q10a(bolap_ctx* ctx, const bolap_ceb_q10a_args* args, bolap_minmax_row* rows_out,
                   int64_t max_rows, int64_t* n_out, bolap_minmax_stats* stats);

/* bolap_ilike_mask over a dictionary in device memory (chars, offsets and
   mask are device pointers; the pattern is a host string), same answers. */
int bolap_ilike_mask_device(bolap_ctx* ctx, const char* chars, const int64_t* offsets, int64_t n,
                            const char* pattern, uint8_t* mask);

/* ------------------------------------------------------------------------ */
/* CEB build-time data (bespoke_ceb/builder_impl.hpp:14-89, builder_impl.cpp  */
/* :364-466): one string pool mapping every string column to int32 ids, and  */
/* per-table sort orders.                                                    */
/* ------------------------------------------------------------------------ */
typedef struct bolap_string_pool bolap_string_pool;

/* Replaces StringPool (builder_impl.hpp:14-89).  Host memory only. */
int bolap_string_pool_create(bolap_string_pool** out);
void bolap_string_pool_destroy(bolap_string_pool* pool);
/* ids_out[i] = id of string i (chars[offsets[i] .. offsets[i+1])), a new id
   (the pool's size) for a string not seen before; -1 where valid[i] == 0
   (valid may be null).  StringPool::intern, builder_impl.hpp:26-39, with ids
   in first-appearance order. */
int bolap_string_pool_intern(bolap_string_pool* pool, const char* chars, const int64_t* offsets,
                             int64_t n, const uint8_t* valid, int32_t* ids_out);
int bolap_string_pool_size(const bolap_string_pool* pool, int64_t* n_strings, int64_t* n_chars);
/* The dictionary in id order (chars of n_chars bytes, offsets of n_strings +
   1), ASCII lower-cased when lower != 0 (StringPool::finalize and
   get_lower, builder_impl.hpp:41-82). */
int bolap_string_pool_export(const bolap_string_pool* pool, int32_t lower, char* chars,
                             int64_t* offsets);

/* order_out[i] = the input row at position i of the table sorted ascending by
   (primary, secondary) as signed int32 (secondary may be null); ties keep
   input order.  Replaces sort_table_by_i32_column(s), builder_impl.cpp
   :265-340 (std::sort, ties unspecified).  Device pointers; synchronous. */
int bolap_sort_order_i32(bolap_ctx* ctx, const int32_t* primary, const int32_t* secondary,
                         int64_t rows, uint32_t* order_out);
/* dst[i] = src[order[i]] for elements of elem_bytes (1, 2, 4 or 8): applies a
   sort order to one column (reorder_vector, builder_impl.cpp:255-263).  An
   order entry >= src_rows is an error.  Device pointers; synchronous. */
int bolap_gather_rows(bolap_ctx* ctx, const void* src, int64_t src_rows, int32_t elem_bytes,
                      const uint32_t* order, int64_t rows, void* dst);

/* ------------------------------------------------------------------------ */
/* Table statistics for the estimator (computed once per column, at load).   */
/* ------------------------------------------------------------------------ */
#define BOLAP_HLL_REGS 4096        /* HyperLogLog, p = 12 */
#define BOLAP_SAMPLE_ROWS 4096
#define BOLAP_SAMPLE_COUNTS 16

/* HyperLogLog registers of an int32 column (host array of 4096 bytes):
   h = mix64(value), register h >> 52, rank = leading zeros of (h << 12) + 1. */
int bolap_hll_i32(bolap_ctx* ctx, const int32_t* col, int64_t n, uint8_t* regs_host);

/* Sample counts of n_plans star plans on 4096 sampled rows (row i of a table
   of n rows is ((mix-hash(seed, stream, i) >> 32) * n) >> 32).  Per plan,
   BOLAP_SAMPLE_COUNTS int64 (host): [0..2] sampled dimension rows passing
   each dimension filter; [3..5] sampled fact rows whose key passes each
   dimension; [6] sampled fact rows passing every dimension; [7] distinct
   group keys among those; [8] of them seen once; [9] seen twice.  Needs
   dimension keys equal to key_min + row. */
int bolap_star_sample_counts(bolap_ctx* ctx, const bolap_star_args* plans, int32_t n_plans,
                             uint64_t seed, int64_t* counts_host);

/* ------------------------------------------------------------------------ */
/* LSD radix sort of 64-bit keys with 32-bit payloads (ascending, stable).    */
/* Replaces the 16-bit-digit counting sorts of bespoke_tpch/query_q10.cpp:    */
/* 134-166 and 318-373.  keys/vals are device arrays, sorted in place.        */
/* key_bits: only the low key_bits bits take part (passes = ceil(bits/8)).    */
/* ------------------------------------------------------------------------ */
int bolap_radix_sort_u64(bolap_ctx* ctx, uint64_t* keys, uint32_t* vals, int64_t n,
                         int32_t key_bits);
/* Autotuner of the radix digit: sorts copies of the pairs (the arrays are not
   modified) with 8-, 7-, 6- and 5-bit digits (knob radix.digit_bits; unset,
   the sort takes the fewest 8-bit passes and the narrowest digit that still
   fits them), median of reps sorts each, checks every width's output against
   the 8-bit one and keeps the fastest width.  No reference counterpart (the
   reference's digit is fixed at 16 bits, bespoke_tpch/query_q10.cpp:139). */
int bolap_sort_autotune(bolap_ctx* ctx, const uint64_t* keys, const uint32_t* vals, int64_t n,
                        int32_t key_bits, int32_t reps, double* best_ms);
/* k smallest keys (ascending) with their payloads, by LDS-histogram radix
   select; keys_out/vals_out are device arrays of k entries. */
int bolap_topk_u64(bolap_ctx* ctx, const uint64_t* keys, const uint32_t* vals, int64_t n,
                   int32_t k, uint64_t* keys_out, uint32_t* vals_out);

/* ------------------------------------------------------------------------ */
/* Learned cardinality estimator: MLP 256-128-64-3 (ReLU), BF16 MFMA GEMMs    */
/* with FP32 accumulation, trained with Adam on the device.  No reference     */
/* counterpart: the reference sizes tables from a hard-coded selectivity      */
/* (bespoke_tpch/query_q20.cpp:16, 405-417) and grows them on overflow        */
/* (:158-160).  Outputs: log selectivity, log join size, log group count.     */
/* Inputs/outputs are device fp32 row-major arrays (n x 256, n x 3).          */
/* ------------------------------------------------------------------------ */
#define BOLAP_MLP_IN 256
#define BOLAP_MLP_OUT 3
typedef struct bolap_mlp bolap_mlp;

int bolap_mlp_create(bolap_ctx* ctx, uint64_t seed, bolap_mlp** out);
int bolap_mlp_destroy(bolap_mlp* mlp);
/* Parameter count and host copies, ordered W1t[128][256], b1[128],
   W2t[64][128], b2[64], W3t[3][64], b3[3] (fp32 master weights). */
int bolap_mlp_num_params(void);
int bolap_mlp_get_params(bolap_mlp* mlp, float* host_params);
int bolap_mlp_set_params(bolap_mlp* mlp, const float* host_params);
int bolap_mlp_predict(bolap_mlp* mlp, const float* x_dev, int64_t n, float* y_dev);
/* One Adam step on the mean squared error of a batch; *loss_out (optional,
   host) receives the batch loss before the step. */
int bolap_mlp_train_step(bolap_mlp* mlp, const float* x_dev, const float* t_dev, int64_t n,
                         float lr, float* loss_out);
/* Gradients of the batch loss at the current weights (host, parameter order),
   no update. */
int bolap_mlp_gradients(bolap_mlp* mlp, const float* x_dev, const float* t_dev, int64_t n,
                        float* grads_host, float* loss_out);

/* ------------------------------------------------------------------------ */
/* Learned sizing as a library call (no reference counterpart; the reference  */
/* sizes tables from fixed selectivities, bespoke_tpch/query_q20.cpp:16,      */
/* 405-417).  An estimator is the MLP's parameters, the standardisation of    */
/* its three outputs (mu, sd) and the two sizing margins, saved to and loaded */
/* from a file.  bolap_estimator_hints fills pass_hint of every dimension and */
/* group_hint of each plan from the plan's sample counts, its predicate      */
/* constants and per-column statistics registered by device pointer: every   */
/* dimension's fact key (FK: HyperLogLog registers of bolap_hll_i32, span =  */
/* dimension rows), the group column (GROUP, span = groups; or one of the FK  */
/* columns), and every predicate column (VALUE for range-filtered values,     */
/* CATEGORY for categorical codes; span = the domain size).  Hints only size: */
/* bolap_star_groupby's results do not depend on them.                        */
/* ------------------------------------------------------------------------ */
#define BOLAP_EST_FK 0
#define BOLAP_EST_GROUP 1
#define BOLAP_EST_VALUE 2
#define BOLAP_EST_CATEGORY 3
#define BOLAP_EST_SAMPLE_SEED 0x5EEDull   /* the seed the sample counts use */

typedef struct {
    const void* col;             /* device column, as the plans point at it */
    int32_t kind;                /* BOLAP_EST_* */
    int32_t pad;
    int64_t span;                /* see above */
    uint8_t hll[BOLAP_HLL_REGS]; /* FK / GROUP: HyperLogLog registers */
} bolap_est_column;

typedef struct bolap_estimator bolap_estimator;

/* params: bolap_mlp_num_params() floats (bolap_mlp_get_params order);
   mu_sd: mu[3] then sd[3]; margins: dimension margin, group margin. */
int bolap_estimator_create(bolap_ctx* ctx, const float* params, const float* mu_sd,
                           const double* margins, bolap_estimator** out);
int bolap_estimator_destroy(bolap_estimator* est);
int bolap_estimator_save(const bolap_estimator* est, const char* path);
int bolap_estimator_load(bolap_ctx* ctx, const char* path, bolap_estimator** out);
/* plans are host structs (their columns on the device); hints written in place. */
int bolap_estimator_hints(bolap_estimator* est, const bolap_est_column* cols, int32_t n_cols,
                          bolap_star_args* plans, int32_t n_plans);

/* ------------------------------------------------------------------------ */
/* Q6 shape: three-predicate range scan + SUM(extendedprice * discount) over  */
/* the reference layout.  Replaces run_q6's scan (bespoke_tpch/query_q6.cpp:  */
/* 155-397): the flat loop (:197-223), or, with a zone map, the shard-pruned  */
/* loops (:224-379) over columns clustered by the shard key.                  */
/* ------------------------------------------------------------------------ */
#define BOLAP_ZONE_ROWS 2048

/* min/max of the Q6 predicate columns over BOLAP_ZONE_ROWS consecutive rows
   (the last zone may be shorter); 16 bytes. */
typedef struct {
    int16_t ship_min, ship_max;
    int16_t qty_min, qty_max;
    uint8_t disc_min, disc_max;
    uint8_t reserved[6];
} bolap_zone;

typedef struct {
    int64_t rows;
    const int16_t* ship;       /* shipdate, days since 1992-01-01 */
    const uint8_t* disc;       /* discount in hundredths (8-byte aligned) */
    const int16_t* qty;        /* quantity x 100 */
    const int32_t* price;      /* extendedprice in cents */
    int32_t ship_lo, ship_hi;  /* a row passes when ship_lo <= ship < ship_hi, */
    int32_t disc_lo, disc_hi;  /* disc_lo <= disc <= disc_hi                    */
    int32_t qty_lt;            /* and qty < qty_lt                              */
    const bolap_zone* zones;   /* null: flat scan; else ceil(rows / ZONE_ROWS)
                                  zones of these columns (bolap_q6_zones) */
} bolap_q6_args;

typedef struct {
    int64_t revenue;       /* SUM(extendedprice * discount), exact int64 */
    int64_t rows_emitted;  /* rows that passed */
    int64_t rows_scanned;  /* rows of the blocks that were read */
} bolap_q6_result;

/* Q6 arguments from the query's strings as run_q6 derives them
   (query_q6.cpp:171-190, 92-95): [DATE, DATE + 1 year) as day offsets with
   the day clamped to the month; DISCOUNT +- 0.01 in hundredths; QUANTITY
   x 100; month_lo/hi = the shard month-index window (year * 12 + month - 1).
   Host only. */
typedef struct {
    int32_t ship_lo, ship_hi, disc_lo, disc_hi, qty_lt, month_lo, month_hi;
} bolap_q6_params;

int bolap_q6_parse(const char* date, const char* discount, const char* quantity,
                   bolap_q6_params* out);
int bolap_q6(bolap_ctx* ctx, const bolap_q6_args* args, bolap_q6_result* out);
/* Zone map of (possibly clustered) columns: zones_out holds
   ceil(rows / BOLAP_ZONE_ROWS) entries. */
int bolap_q6_zones(bolap_ctx* ctx, int64_t rows, const int16_t* ship, const uint8_t* disc,
                   const int16_t* qty, bolap_zone* zones_out);
/* Copies of the Q6 columns ordered by the reference's shard key (ship month,
   discount, quantity / 100; builder_impl.cpp:855-911), stable within a key;
   perm_out (optional) receives the source row of each output row. */
int bolap_q6_cluster(bolap_ctx* ctx, int64_t rows, const int16_t* ship, const uint8_t* disc,
                     const int16_t* qty, const int32_t* price, int16_t* ship_out,
                     uint8_t* disc_out, int16_t* qty_out, int32_t* price_out,
                     uint32_t* perm_out);

/* ------------------------------------------------------------------------ */
/* Q18 shape: SUM(quantity) per order key into a dense array, orders above a  */
/* threshold, joined to orders and customers, ordered by totalprice desc,     */
/* orderdate asc, orderkey asc.  Replaces run_q18 (bespoke_tpch/query_q18.cpp:*/
/* 183-455): the dense-key path when orders/customers carry keys 1..N (first  */
/* and last key checked, :194-199), else the max-key-sized path.  An order    */
/* qualifies when its total exceeds the threshold; the reference's dense path */
/* tests running sums, which is the same for non-negative quantities.         */
/* Remaining ties (duplicate order keys) are broken by order row.             */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t li_rows;
    const int32_t* l_orderkey;
    const int16_t* l_quantity;      /* x 100 */
    int32_t orderkey_sorted;        /* lineitem sorted by order key: run sums
                                       without atomics (checked on the device;
                                       out of order reruns with atomics) */
    int64_t order_rows;
    const int32_t* o_orderkey;
    const int32_t* o_custkey;
    const int16_t* o_orderdate;
    const int32_t* o_totalprice;    /* cents */
    int64_t cust_rows;
    const int32_t* c_custkey;
    int32_t quantity_threshold;     /* x 100 */
} bolap_q18_args;

typedef struct {
    int32_t custkey, orderkey, orderdate, totalprice;
    int64_t sum_qty;
    int32_t cust_row;   /* customer row (for c_name on the host) */
    int32_t order_row;
} bolap_q18_row;

typedef struct {
    int64_t rows;         /* result rows */
    int32_t dense_path;   /* 1: dense order keys */
    int32_t sorted_pass;  /* 1: run sums without atomics */
    int32_t fallbacks;    /* 1: sorted hint was wrong, rerun with atomics */
    int32_t reserved;
} bolap_q18_stats;

/* rows_out (host) holds max_rows rows; more result rows is an error
   (stats->rows says how many). */
int bolap_q18(bolap_ctx* ctx, const bolap_q18_args* args, bolap_q18_row* rows_out,
              int64_t max_rows, bolap_q18_stats* stats);

/* write_q3_csv (query_q3.cpp:367-380): rows as bolap_chain_topk /
   bolap_chain_groups_sorted return them (revenue in cents of the summed
   discounted_price, day = order date offset from 1992-01-01), 2 decimals,
   ISO dates; shippriority[i] per row (null: 0, the TPC-H value).  Host only. */
int bolap_q3_write_csv(const char* path, const bolap_group_row* rows, int64_t n,
                       const int32_t* shippriority);
/* write_q6_csv (query_q6.cpp:399-408) and write_q18_csv (query_q18.cpp:465-484:
   c_name quoted with \" escapes, ISO dates, 2 decimals); c_name[cust_row]
   gives each row's customer name.  Host only. */
int bolap_q6_write_csv(const char* path, const bolap_q6_result* result);
int bolap_q18_write_csv(const char* path, const bolap_q18_row* rows, int64_t n,
                        const char* const* c_name, int64_t n_names);

/* ------------------------------------------------------------------------ */
/* Q14 shape: lineitem rows shipped in a one-month window joined to part    */
/* through a dense partkey -> promo-flag array, SUM(discounted_price) in     */
/* total and for promo parts.  Replaces run_q14 (bespoke_tpch/query_q14.cpp: */
/* 176-365); the flag array is built per call over 0..max part key with the  */
/* last part row of a key winning, as there.                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t li_rows;
    const int16_t* l_ship;
    const int32_t* l_partkey;
    const int32_t* l_disc_price;   /* truncated discounted_price (cents) */
    int32_t ship_lo, ship_hi;      /* ship_lo <= ship < ship_hi */
    const bolap_zone* zones;       /* optional zone map of these rows (bolap_q6_zones) */
    int64_t part_rows;
    const int32_t* p_partkey;
    const uint8_t* p_promo;        /* 1 when the part's type starts with "PROMO" */
} bolap_q14_args;

typedef struct {
    int64_t promo_sum, total_sum, rows_emitted;
    double promo_revenue;          /* 100 * promo / total, 0 when total <= 0 */
} bolap_q14_result;

/* [DATE, DATE + 1 month) as day offsets, the day clamped to the month
   (query_q14.cpp:46-74).  Host only. */
int bolap_q14_parse(const char* date, int32_t* ship_lo, int32_t* ship_hi);
int bolap_q14(bolap_ctx* ctx, const bolap_q14_args* args, bolap_q14_result* out);
/* write_q14_csv (query_q14.cpp:367-375), write_q17_csv (query_q17.cpp:311-322).
   Host only. */
int bolap_q14_write_csv(const char* path, const bolap_q14_result* result);
int bolap_q17_write_csv(const char* path, const bolap_q17_result* result);

/* ------------------------------------------------------------------------ */
/* Order-range queries Q4, Q5, Q10 over the reference layout: orders sorted  */
/* by orderdate (builder_impl.cpp:597-608) and each order's lineitem rows as */
/* lineitem_ranges [start, end), {0, 0} for none (:929-991, as               */
/* bolap_order_ranges makes them).  A date window [date_lo, date_hi) is found */
/* on the device with a lower_bound over o_orderdate, as the reference's     */
/* std::lower_bound does.  With l_orderkey_sorted = 0 the ranges may be      */
/* min/max spans and each lineitem's key is checked against its order's, as  */
/* the queries' unsorted branches do.                                         */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t order_rows;
    const int16_t* o_orderdate;      /* non-decreasing day offsets */
    const int32_t* o_orderkey;       /* needed when l_orderkey_sorted == 0 */
    const uint32_t* range_start;
    const uint32_t* range_end;
    const int32_t* l_orderkey;       /* needed when l_orderkey_sorted == 0 */
    int32_t l_orderkey_sorted;       /* lineitem.orderkey_sorted */
} bolap_order_side;

/* Q4 (bespoke_tpch/query_q4.cpp:140-268): orders of [date_lo, date_hi) with
   a lineitem whose commitdate < receiptdate (the packed commit_receipt
   halves compared as uint16, :70-72), counted per priority code;
   counts[n_priority] (host).  Codes >= n_priority are not counted. */
typedef struct {
    bolap_order_side orders;
    const uint8_t* o_orderpriority;  /* codes */
    int32_t n_priority;              /* <= 256 */
    const uint32_t* l_commit_receipt; /* receipt << 16 | commit (builder_impl.cpp:536-544) */
    int32_t date_lo, date_hi;
} bolap_q4_args;
int bolap_q4(bolap_ctx* ctx, const bolap_q4_args* args, int64_t* counts);

/* Q5 (bespoke_tpch/query_q5.cpp:190-440): orders of [date_lo, date_hi) whose
   customer nation (orders.cust_nationkey, builder_impl.cpp:913-928) is
   marked in nation_in_region (a HOST array of n_nation bytes, n_nation =
   max nation key + 1 <= 256); per order the sum of discounted_price over its
   lineitems whose supplier nation (lineitem.supp_nationkey, :838-853) is the
   customer's.  revenue[n_nation] (host), exact int64. */
typedef struct {
    bolap_order_side orders;
    const int16_t* o_cust_nation;
    const int16_t* l_supp_nation;
    const int32_t* l_disc_price;
    const uint8_t* nation_in_region;
    int32_t n_nation;
    int32_t date_lo, date_hi;
} bolap_q5_args;
int bolap_q5(bolap_ctx* ctx, const bolap_q5_args* args, int64_t* revenue);

/* Q10 (bespoke_tpch/query_q10.cpp:375-560): orders of [date_lo, date_hi)
   with a customer key in [0, max c_custkey]; per order the sum of
   discounted_price over lineitems whose returnflag_linestatus code c has
   (u32)(c - code_lo) < code_span (returnflag 'R': code_lo = R * n_linestatus,
   code_span = n_linestatus; code_lo < 0 gives no rows); orders with a
   nonzero sum grouped by customer; a group is kept when its key has a
   customer row (the last row of a key) whose c_nationkey is in
   [0, n_nation_names).  Rows in the reference's order: revenue descending as
   an unsigned 64-bit value, ties by customer key ascending (its stable LSD
   passes, :318-373).  *n_rows = number of groups; BOLAP_ERR_CAPACITY when
   it exceeds max_rows. */
typedef struct {
    bolap_order_side orders;
    const int32_t* o_custkey;
    const uint8_t* l_code;           /* returnflag_linestatus */
    int32_t code_lo, code_span;
    const int32_t* l_disc_price;
    int32_t date_lo, date_hi;
    int64_t cust_rows;
    const int32_t* c_custkey;
    const int32_t* c_nationkey;
    int32_t n_nation_names;          /* max nation key + 1 */
} bolap_q10_args;
typedef struct {
    int32_t custkey;
    int32_t cust_row;                /* customer row (name, address, ... on the host) */
    int64_t revenue;                 /* discounted_price units (cents) */
} bolap_q10_row;
int bolap_q10(bolap_ctx* ctx, const bolap_q10_args* args, bolap_q10_row* rows,
              int64_t max_rows, int64_t* n_rows);

/* Q12 (bespoke_tpch/query_q12.cpp:150-250): lineitems with shipmode code
   mode1 or mode2 (-1 = absent from the dictionary), commitdate <
   receiptdate and shipdate < commitdate (packed halves as int16), and
   (u32)(receiptdate - date_lo) < (u32)(date_hi - date_lo); their order row
   is order_map[orderkey] (map_size entries, -1 = none, as
   bolap_order_ranges makes it) and they count as high when its priority
   code is high1 or high2.  The reference's shard loop prunes shards whose
   shipdates all precede date_lo - 30, which counts the same rows when
   receiptdate <= shipdate + 30 (TPC-H).  Columns l_shipmode, l_commit_receipt
   and l_shipdate must be 4-, 16- and 8-byte aligned. */
typedef struct {
    int64_t li_rows;
    const uint8_t* l_shipmode;
    const uint32_t* l_commit_receipt;
    const int16_t* l_shipdate;
    const int32_t* l_orderkey;
    int32_t mode1, mode2;
    int32_t date_lo, date_hi;
    const int32_t* order_map;
    int64_t map_size;
    const uint8_t* o_orderpriority;
    int32_t high1, high2;            /* codes of 1-URGENT and 2-HIGH (-1 = absent) */
} bolap_q12_args;
typedef struct {
    int64_t high[2];                 /* [0]: mode1, [1]: mode2 */
    int64_t low[2];
} bolap_q12_result;
int bolap_q12(bolap_ctx* ctx, const bolap_q12_args* args, bolap_q12_result* out);

/* Q19 (bespoke_tpch/query_q19.cpp:130-365).  A part row sets its key's mask
   (the last row of a key wins, also with mask 0): bit g when group g is
   enabled, brand == groups[g].brand, container is one of its four codes and
   1 <= size <= size_hi.  A lineitem adds price * (100 - discount) when its
   shipmode is mode1 or mode2, its shipinstruct is `shipinstruct`, and its
   quantity lies in [qty_lo, qty_lo + 1000] for a group of its part's mask.
   Codes are -1 when absent from their dictionary.  Part rows < 2^28 - 1.
   l_shipmode, l_shipinstruct and l_partkey must be 4-, 4- and 16-byte
   aligned. */
typedef struct {
    int32_t enabled;                 /* QUANTITYg and BRANDg given (not <<NULL>>) */
    int32_t brand;
    int32_t containers[4];
    int32_t size_hi;                 /* 5, 10, 15 */
    int32_t qty_lo;                  /* llround(QUANTITYg * 100) */
} bolap_q19_group;
typedef struct {
    int64_t part_rows;
    const int32_t* p_partkey;
    const void* p_brand;             /* u8 or i16 codes */
    int32_t p_brand_dtype;
    const void* p_container;
    int32_t p_container_dtype;
    const int32_t* p_size;
    int64_t li_rows;
    const int32_t* l_partkey;
    const int16_t* l_quantity;       /* x100 */
    const int32_t* l_price;          /* extendedprice, cents */
    const uint8_t* l_disc;
    const uint8_t* l_shipmode;
    const uint8_t* l_shipinstruct;
    bolap_q19_group groups[3];
    int32_t mode1, mode2;            /* "AIR", "AIR REG" */
    int32_t shipinstruct;            /* "DELIVER IN PERSON" */
} bolap_q19_args;
int bolap_q19(bolap_ctx* ctx, const bolap_q19_args* args, int64_t* revenue);

/* Q21 (bespoke_tpch/query_q21.cpp:140-380): over every order whose status
   code is final_code ("F"; -1 = absent) and whose lineitem range has more
   than one row, the lineitems with suppkey in [0, n_suppkeys) (key-checked
   when unsorted): the order counts when its lines have several suppliers
   and exactly one supplier with late lines (receiptdate > commitdate, the
   packed halves unsigned; a second late supplier ends the walk); that
   supplier's late-line count is added when the last supplier row of its
   key has s_nationkey == nation.  wait[n_suppkeys] (host) per supplier key;
   n_suppkeys = max s_suppkey + 1. */
typedef struct {
    int64_t order_rows;
    const uint8_t* o_orderstatus;    /* codes */
    int32_t final_code;
    const int32_t* o_orderkey;       /* needed when l_orderkey_sorted == 0 */
    const uint32_t* range_start;
    const uint32_t* range_end;
    const int32_t* l_orderkey;       /* needed when l_orderkey_sorted == 0 */
    int32_t l_orderkey_sorted;
    const int32_t* l_suppkey;
    const uint32_t* l_commit_receipt;
    int64_t supp_rows;
    const int32_t* s_suppkey;
    const int32_t* s_nationkey;
    int32_t nation;
} bolap_q21_args;
int bolap_q21(bolap_ctx* ctx, const bolap_q21_args* args, int64_t* wait, int64_t n_suppkeys);

/* Q22 (bespoke_tpch/query_q22.cpp:190-520): a customer is a candidate when
   its phone prefix is in the filter — the two-digit prefix code c_phone_code
   (255 when the phone does not start with two digits, builder_impl.cpp:
   152-178) with numeric_allowed[code], or else its first two characters
   c_phone_pair (c0 << 8 | c1, -1 when shorter) equal to one of extra_pairs —
   and c_acctbal > 0.  Candidates above floor(sum / count) of the candidates'
   acctbal whose key no order carries count per prefix. */
typedef struct {
    int64_t cust_rows;
    const int32_t* c_custkey;
    const uint8_t* c_phone_code;
    const int32_t* c_phone_pair;
    const int32_t* c_acctbal;        /* cents */
    uint8_t numeric_allowed[100];
    int32_t n_extra;                 /* <= 7 */
    int32_t extra_pairs[7];
    int64_t order_rows;
    const int32_t* o_custkey;
} bolap_q22_args;
typedef struct {
    int64_t numcust[100], totacctbal[100];          /* by two-digit code */
    int64_t extra_numcust[7], extra_totacctbal[7];  /* by extra prefix */
} bolap_q22_result;
int bolap_q22(bolap_ctx* ctx, const bolap_q22_args* args, bolap_q22_result* out);

/* Q7 (bespoke_tpch/query_q7.cpp:150-330): suppliers tagged 0 by nation1 or
   (nation1 != nation2) 1 by nation2 — the last row of a key with one of the
   two nations decides.  Over every order whose customer nation
   (o_cust_nation) is nation1 (index 0) or nation2 (index 1), its lineitems
   with shipdate in [ship_lo, ship_hi] (inclusive) and a tagged supplier of
   the other index (both 0 when the nations are equal) add discounted_price
   to rev[supp * 4 + cust * 2 + (shipdate >= split)].  n_suppkeys = max
   s_suppkey + 1; lineitem suppkeys outside it are untagged.  rev[8] host. */
typedef struct {
    bolap_order_side orders;         /* o_orderdate unused */
    const int16_t* o_cust_nation;
    const int16_t* l_shipdate;
    const int32_t* l_suppkey;
    const int32_t* l_disc_price;
    int64_t supp_rows;
    const int32_t* s_suppkey;
    const int32_t* s_nationkey;
    int64_t n_suppkeys;
    int32_t nation1, nation2;
    int32_t ship_lo, ship_hi, split;
} bolap_q7_args;
int bolap_q7(bolap_ctx* ctx, const bolap_q7_args* args, int64_t* rev);

/* Q8 (bespoke_tpch/query_q8.cpp:150-330): customer keys marked by rows whose
   nation is marked in nation_in_region (a HOST array of n_nation bytes),
   supplier keys by rows of target_nation, part keys by rows whose type code
   is type_code (-1: none); marks are never cleared.  Orders of
   [date_lo, date_hi] (inclusive) with a marked customer: every lineitem of a
   marked part adds its volume price * (100 - discount) to total[year] and,
   with a marked supplier, to nation[year]; year = orderdate >= split.
   n_*_keys = max key + 1 of each table (keys outside are unmarked). */
typedef struct {
    bolap_order_side orders;
    const int32_t* o_custkey;
    const int32_t* l_partkey;
    const int32_t* l_suppkey;
    const int32_t* l_price;
    const uint8_t* l_disc;
    int64_t cust_rows;
    const int32_t* c_custkey;
    const int32_t* c_nationkey;
    int64_t n_cust_keys;
    const uint8_t* nation_in_region;
    int32_t n_nation;
    int64_t supp_rows;
    const int32_t* s_suppkey;
    const int32_t* s_nationkey;
    int64_t n_supp_keys;
    int32_t target_nation;
    int64_t part_rows;
    const int32_t* p_partkey;
    const uint8_t* p_type;           /* codes */
    int64_t n_part_keys;
    int32_t type_code;
    int32_t date_lo, date_hi, split;
} bolap_q8_args;
int bolap_q8(bolap_ctx* ctx, const bolap_q8_args* args, int64_t* total, int64_t* nation);

/* Q9 (bespoke_tpch/query_q9.cpp:150-400): part keys marked by the rows with
   p_ok (the name contains the colour, decided on the host per name as the
   reference does).  Every lineitem of a marked part takes the supply cost of
   the first partsupp row (in row order) with its partkey and suppkey, its
   supplier nation (l_supp_nation in [0, n_nation)) and its order's year
   (order_map[orderkey], map_size entries; rows without an order are
   skipped) and adds discounted_price - supplycost * quantity / 100 (C
   division) to totals[nation * year_span + year - min_year], setting used[]
   too.  min_year / year_span come from the first and last o_orderdate (the
   orders are sorted by date); n_nation * year_span <= 2048.  totals and used
   (host) hold n_nation * year_span entries. */
typedef struct {
    int64_t part_rows;
    const int32_t* p_partkey;
    const uint8_t* p_ok;
    int64_t n_part_keys;             /* max p_partkey + 1 */
    int64_t ps_rows;
    const int32_t* ps_partkey;
    const int32_t* ps_suppkey;
    const int32_t* ps_supplycost;    /* cents */
    int64_t order_rows;
    const int16_t* o_orderdate;
    const int32_t* order_map;
    int64_t map_size;
    int64_t li_rows;
    const int32_t* l_partkey;
    const int32_t* l_suppkey;
    const int32_t* l_orderkey;
    const int16_t* l_quantity;
    const int32_t* l_disc_price;
    const int16_t* l_supp_nation;
    int32_t n_nation;
} bolap_q9_args;
int bolap_q9(bolap_ctx* ctx, const bolap_q9_args* args, int64_t* totals, uint8_t* used,
             int32_t* min_year, int32_t* year_span);

/* The builder's per-string prefilter masks of a string column (chars +
   n + 1 uint32 offsets, builder_impl.cpp:124-150): alpha_out[i] has bit
   c - 'a' for every letter (either case), bigram_out[i] bit
   (c0 * 131 + c1) & 63 for every pair of adjacent bytes.  Stream-ordered. */
int bolap_string_masks(bolap_ctx* ctx, const uint8_t* chars, const uint32_t* offsets, int64_t n,
                       uint32_t* alpha_out, uint64_t* bigram_out);

/* Q13 (bespoke_tpch/query_q13.cpp:150-250): an order matches when word1
   occurs in its comment and word2 after that occurrence's end (an empty word
   is found at the search start, which must not be the comment's end;
   :27-82); alpha / bigram (optional, bolap_string_masks) skip comments that
   cannot match.  Every order that does not match adds one to its customer
   key's count (keys in [0, max c_custkey]); dist[c] = customer rows whose
   key has count c.  *max_count = the largest count; BOLAP_ERR_CAPACITY when
   max_dist <= it.  Words are host strings shorter than 64 bytes. */
typedef struct {
    int64_t order_rows;
    const int32_t* o_custkey;
    const uint8_t* chars;            /* o_comment bytes */
    const uint32_t* offsets;         /* order_rows + 1 */
    const uint32_t* alpha;           /* optional prefilter masks */
    const uint64_t* bigram;
    const char* word1;
    const char* word2;
    int64_t cust_rows;
    const int32_t* c_custkey;
} bolap_q13_args;
int bolap_q13(bolap_ctx* ctx, const bolap_q13_args* args, int64_t* dist, int32_t max_dist,
              int32_t* max_count);

/* Q16 (bespoke_tpch/query_q16.cpp:180-480): a part row is eligible when its
   brand code is not banned_brand, its size (0..50) is in size_mask and its
   type code is not marked in type_banned (a HOST array of n_types bytes:
   types starting with TYPE); its key (in [0, n_part_keys)) joins group
   (brand * n_types + type) * 51 + size — the last eligible row of a key
   decides.  Every partsupp row of a grouped key whose supplier is not
   banned (s_banned per supplier row: the comment has "Customer" followed by
   "Complaints", decided on the host; keys in [0, n_supp_keys)) adds its
   supplier to the group's distinct set.  counts (host) has
   n_brands * n_types * 51 entries: distinct suppliers per group. */
typedef struct {
    int64_t part_rows;
    const int32_t* p_partkey;
    const uint8_t* p_brand;          /* codes */
    const uint8_t* p_type;
    const int32_t* p_size;
    int32_t n_brands, n_types;
    int32_t banned_brand;            /* -1: none */
    const uint8_t* type_banned;
    uint64_t size_mask;
    int64_t n_part_keys;
    int64_t ps_rows;
    const int32_t* ps_partkey;
    const int32_t* ps_suppkey;
    int64_t supp_rows;
    const int32_t* s_suppkey;
    const uint8_t* s_banned;
    int64_t n_supp_keys;
} bolap_q16_args;
int bolap_q16(bolap_ctx* ctx, const bolap_q16_args* args, int64_t* counts, int64_t n_counts);

/* Q2 (bespoke_tpch/query_q2.cpp:146-405), its scans: a part row qualifies
   when p_size == size and type_ok[p_type] (host array of n_types flags: the
   type ends with TYPE, :219-231).  Its key is taken as row + 1 (:295); its
   partsupp run (partsupp sorted by partkey, builder_impl.cpp:769-777) comes
   from the row's stride block when that block is exactly the key's run, else
   from a lower_bound (:288-324).  Rows whose supplier row suppkey - 1 lies in
   the supplier table and whose nation (s_nationkey, keys below n_nation)
   has nation_ok set (host array: the nation is in REGION) keep the minimum
   supplycost; every such row at the minimum is a result (:331-392).
   pairs_out receives (part row, supplier row) int32 pairs in the reference's
   emission order (part rows ascending, then partsupp rows); *rows = their
   number, BOLAP_ERR_CAPACITY when it exceeds max_rows.  The caller sorts by
   (s_acctbal desc, n_name, s_name, p_partkey) with the strings (:408-423). */
typedef struct {
    int64_t part_rows;
    const int32_t* p_size;
    const uint8_t* p_type;         /* dictionary codes (at most 256 types) */
    const uint8_t* type_ok;        /* host, n_types entries */
    int32_t n_types;
    int32_t size;                  /* stoi(SIZE) */
    int64_t ps_rows;
    const int32_t* ps_partkey;     /* sorted */
    const int32_t* ps_suppkey;
    const int32_t* ps_supplycost;  /* cents */
    int64_t supp_rows;
    const int32_t* s_nationkey;
    const uint8_t* nation_ok;      /* host, n_nation entries, by nation key */
    int32_t n_nation;
} bolap_q2_args;
int bolap_q2(bolap_ctx* ctx, const bolap_q2_args* args, int32_t* pairs_out, int64_t max_rows,
             int64_t* rows);

/* Q11 (bespoke_tpch/query_q11.cpp:119-217): suppliers with s_nationkey ==
   nation mark their key (keys below n_supp_keys = max s_suppkey + 1, or 1
   without suppliers).  Each run of equal ps_partkey in row order sums
   ps_value (supplycost_availqty, builder_impl.cpp:760-766) over its rows of
   marked suppliers; a non-zero sum is a group.  *total = the sum over all
   marked rows; groups with (double)sum > (double)total * fraction are
   written, sorted by value descending then partkey (:194-214).  *rows =
   their number, BOLAP_ERR_CAPACITY beyond max_rows. */
typedef struct {
    int64_t supp_rows;
    const int32_t* s_suppkey;
    const int32_t* s_nationkey;
    int32_t nation;
    int64_t n_supp_keys;
    int64_t ps_rows;
    const int32_t* ps_partkey;     /* sorted */
    const int32_t* ps_suppkey;
    const int32_t* ps_value;
    double fraction;               /* stod(FRACTION) */
} bolap_q11_args;
int bolap_q11(bolap_ctx* ctx, const bolap_q11_args* args, int32_t* partkey_out,
              int64_t* value_out, int64_t max_rows, int64_t* rows, int64_t* total);

/* Q15 (bespoke_tpch/query_q15.cpp:215-534): lineitems with
   (uint32)(l_shipdate - date_lo) < (uint32)(date_hi - date_lo) add
   discounted_price to revenue[l_suppkey] (keys below n_supp_keys = max
   s_suppkey + 1, at least 1; other keys are skipped).  *emitted = lineitems
   in the window.  Without one, or without a positive revenue, there is no
   result; otherwise the keys whose revenue equals the largest positive one
   and that have a supplier row are written ascending with their last
   supplier row and the revenue (:461-525).  *rows = their number. */
typedef struct {
    int64_t li_rows;
    const int16_t* l_shipdate;     /* 16-byte aligned */
    const int32_t* l_suppkey;
    const int32_t* l_disc_price;
    int32_t date_lo, date_hi;      /* int16 day offsets, as the reference parses them */
    int64_t supp_rows;
    const int32_t* s_suppkey;
    int64_t n_supp_keys;
} bolap_q15_args;
int bolap_q15(bolap_ctx* ctx, const bolap_q15_args* args, int32_t* suppkey_out, int32_t* row_out,
              int64_t* revenue_out, int64_t max_rows, int64_t* rows, int64_t* emitted);

/* Q20 (bespoke_tpch/query_q20.cpp:331-620): part rows with p_ok_row (device
   flags: p_name starts with COLOR) and a key in [0, max_partkey] colour
   their key.  Lineitems in [date_lo, date_hi) (the unsigned window test)
   whose supp_nationkey == nation and whose partkey is coloured sum
   l_quantity per (partkey, suppkey) in an open-addressing table of a power
   of two >= 2 x capacity_hint slots (the reference's FlatHashMap reserve,
   :145-152, with its estimate rows x 1/7 x colour selectivity, :399-409).
   capacity_hint = 0 counts the rows first and sizes exactly; a table that
   fills beyond 0.7 is rerun exactly sized (stats.fallbacks).  Partsupp rows
   with a coloured key, a supplier of the nation (supp_nation_by_key, keys
   below n_supp_map) and a sum set qualified[suppkey] (host array of
   max_suppkey + 1 bytes) when availqty * 2 * 100 > sum (:548-580). */
typedef struct {
    int64_t part_rows;
    const int32_t* p_partkey;
    const uint8_t* p_ok_row;       /* device */
    int32_t max_partkey;           /* -1 without parts */
    int64_t li_rows;
    const int16_t* l_shipdate;     /* 16-byte aligned */
    const int32_t* l_partkey;
    const int32_t* l_suppkey;
    const int16_t* l_supp_nation;  /* 16-byte aligned */
    const int16_t* l_quantity;     /* x 100 */
    int32_t date_lo, date_hi, nation;
    int64_t ps_rows;
    const int32_t* ps_partkey;
    const int32_t* ps_suppkey;
    const int32_t* ps_availqty;
    const int16_t* supp_nation_by_key;
    int64_t n_supp_map;
    int32_t max_suppkey;           /* -1 without suppliers */
    int64_t capacity_hint;
} bolap_q20_args;
typedef struct {
    int64_t pairs;      /* (partkey, suppkey) groups */
    int64_t capacity;   /* table slots of the final run */
    int32_t fallbacks;  /* 1 when the estimate was short */
    int64_t qualified;  /* suppkeys set */
} bolap_q20_stats;
int bolap_q20(bolap_ctx* ctx, const bolap_q20_args* args, uint8_t* qualified,
              bolap_q20_stats* stats);

/* ------------------------------------------------------------------------ */
/* Build-time derivations of the reference Database layout on the device.    */
/* ------------------------------------------------------------------------ */
/* discounted_price = extendedprice * (100 - discount) / 100 truncated, and
   returnflag_linestatus = rf * n_ls + ls (builder_impl.cpp:523-532, 552-567);
   *orderkey_sorted = keys non-decreasing (:569-577; 1 when orderkey is null). */
int bolap_derive_lineitem(bolap_ctx* ctx, int64_t rows, const int32_t* price,
                          const uint8_t* disc, const uint8_t* rf, const uint8_t* ls,
                          int32_t n_ls, const int32_t* orderkey, int32_t* disc_price_out,
                          uint8_t* code_out, int32_t* orderkey_sorted);
/* order_out[k] = row placed k-th by a stable sort on the date
   (build_date_sorted_indices, builder_impl.cpp:374-405). */
int bolap_date_order(bolap_ctx* ctx, int64_t rows, const int16_t* date, uint32_t* order_out);
/* map_out[key] = order row (-1 absent, last row wins; map_size = max key + 1)
   and each order's lineitem rows [start, end), {0, 0} when it has none
   (builder_impl.cpp:610-630, 929-991). */
int bolap_order_ranges(bolap_ctx* ctx, int64_t o_rows, const int32_t* o_orderkey,
                       int64_t map_size, int32_t* map_out, int64_t li_rows,
                       const int32_t* l_orderkey, uint32_t* start_out, uint32_t* end_out);

/* ------------------------------------------------------------------------ */
/* Seeded synthetic tables (counter-based SplitMix64, order independent).     */
/* Every output is a caller-allocated device array.                           */
/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t* ship;
    int64_t* qty;
    int64_t* price;
    int64_t* disc;
    int64_t* tax;
    uint8_t* flag_a;
    uint8_t* flag_b;
} bolap_gen_lineitem_cols;

int bolap_gen_lineitem(bolap_ctx* ctx, uint64_t seed, int64_t rows,
                       const bolap_gen_lineitem_cols* cols);
int bolap_gen_dim(bolap_ctx* ctx, uint64_t seed, int64_t rows, int32_t* key,
                  uint8_t* cat, uint8_t* cont);
int bolap_gen_fact(bolap_ctx* ctx, uint64_t seed, int64_t rows, int64_t dim_rows,
                   int32_t* fk, int64_t* price, int64_t* disc);
int bolap_gen_customers(bolap_ctx* ctx, uint64_t seed, int64_t rows, int32_t* key,
                        uint8_t* segment);
/* Orders plus per-order fact row counts (1..7); fact_offsets gets the
   exclusive prefix sum (rows+1 entries, last = fact rows). */
int bolap_gen_orders(bolap_ctx* ctx, uint64_t seed, int64_t rows, int64_t cust_rows,
                     int32_t* key, int32_t* day, int32_t* cust, int64_t* fact_offsets);
int bolap_gen_order_facts(bolap_ctx* ctx, uint64_t seed, int64_t order_rows,
                          const int64_t* fact_offsets, int64_t fact_rows,
                          int32_t* fk_order, int64_t* price, int64_t* disc);

/* Generic columns for the star schema.  Stream ids select independent
   sequences.  zipf: ranks drawn through a 65,537-entry inverse-CDF table
   (quantiles[i] = rank at cumulative probability i / 65536, 1-based,
   non-decreasing; linear interpolation between entries), then scattered over
   1..n_keys by key = (rank - 1) * mul % n_keys + 1 (mul coprime to n_keys). */
int bolap_gen_iota_i32(bolap_ctx* ctx, int64_t rows, int32_t first, int32_t* out);
int bolap_gen_uniform_i32(bolap_ctx* ctx, uint64_t seed, uint32_t stream, int64_t rows,
                          int32_t lo, int32_t hi, int32_t* out);
int bolap_gen_uniform_u8(bolap_ctx* ctx, uint64_t seed, uint32_t stream, int64_t rows,
                         int32_t lo, int32_t hi, uint8_t* out);
int bolap_gen_zipf_i32(bolap_ctx* ctx, uint64_t seed, uint32_t stream, int64_t rows,
                       const int32_t* quantiles_dev, int64_t n_keys, uint64_t mul,
                       int32_t* out);
/* price = U{1..50} * U{90000..200000} / 100, disc = U{0..10} (streams s, s+1, s+2). */
int bolap_gen_price_disc(bolap_ctx* ctx, uint64_t seed, uint32_t stream, int64_t rows,
                         int64_t* price, int64_t* disc);

#ifdef __cplusplus
}
#endif
#endif /* BOLAP_H */

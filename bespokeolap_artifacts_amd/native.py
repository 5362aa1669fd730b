"""ctypes view of the libbolap C ABI (include/bolap.h).

This module only mirrors the header: structures, prototypes and a loader
that fails loudly when the HIP library is missing.  There is no CPU
fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes as C
import pathlib

_PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = _PKG / "libbolap.so"

BOLAP_OK = 0
BOLAP_ERR_ARG = 1
BOLAP_ERR_HIP = 2
BOLAP_ERR_UNSUPPORTED = 3
BOLAP_ERR_CAPACITY = 4
BOLAP_ERR_NO_DEVICE = 5
BOLAP_ERR_INTERNAL = 6

U8, I16, I32, I64 = 0, 1, 2, 3
PLACE_AUTO, PLACE_LDS, PLACE_HBM = 0, 1, 2
OP_RANGE, OP_IN_SET = 0, 1
Q1_LAYOUT_WIDE, Q1_LAYOUT_REF = 0, 1
MAX_PREDS = 4
Q1_MAX_GROUPS = 16

vp = C.c_void_p


class DeviceInfo(C.Structure):
    _fields_ = [("device", C.c_int32), ("compute_units", C.c_int32),
                ("lds_bytes_per_cu", C.c_int32), ("wavefront", C.c_int32),
                ("hbm_bytes", C.c_int64), ("arch", C.c_char * 32)]


class TraceEntry(C.Structure):
    _fields_ = [("name", C.c_char * 40), ("ms", C.c_double), ("count", C.c_int64)]


class Pred(C.Structure):
    _fields_ = [("col", vp), ("dtype", C.c_int32), ("op", C.c_int32),
                ("lo", C.c_int64), ("hi", C.c_int64),
                ("inline_set", C.c_uint64 * 4), ("set", vp), ("set_bits", C.c_int64)]



STAR_MAX_DIMS = 3
STAR_LDS_KEYS = 1024
HLL_REGS = 4096
SAMPLE_ROWS = 4096
SAMPLE_COUNTS = 16


class StarDim(C.Structure):
    _fields_ = [("rows", C.c_int64), ("key", vp), ("key_min", C.c_int32),
                ("key_max", C.c_int32), ("n_preds", C.c_int32), ("preds", Pred * MAX_PREDS),
                ("fact_fk", vp), ("pass_hint", C.c_int64)]


class StarArgs(C.Structure):
    _fields_ = [("n_dims", C.c_int32), ("dims", StarDim * STAR_MAX_DIMS),
                ("fact_rows", C.c_int64), ("group_key", vp), ("group_min", C.c_int32),
                ("group_max", C.c_int32), ("fact_price", vp), ("price_dtype", C.c_int32),
                ("fact_disc", vp), ("disc_dtype", C.c_int32), ("group_hint", C.c_int64)]


class StarStats(C.Structure):
    _fields_ = [("dim_pass", C.c_int64 * STAR_MAX_DIMS),
                ("dim_placement", C.c_int32 * STAR_MAX_DIMS),
                ("group_placement", C.c_int32), ("matches", C.c_int64), ("groups", C.c_int64),
                ("group_capacity", C.c_int64), ("revenue", C.c_int64),
                ("fallbacks", C.c_int32), ("count_passes", C.c_int32)]


CEB_CI, CEB_MI, CEB_MK = 1, 2, 4
CEB_COUNT, CEB_MIN_YEAR = 0, 1


class CebDb(C.Structure):
    _fields_ = [("title_rows", C.c_int64), ("t_year", vp), ("t_kind", vp), ("t_code", vp),
                ("ci_rows", C.c_int64), ("ci_movie", vp), ("ci_role", vp),
                ("mi_rows", C.c_int64), ("mi_movie", vp), ("mi_type", vp), ("mi_info", vp),
                ("mk_rows", C.c_int64), ("mk_movie", vp), ("mk_keyword", vp)]


class CebQuery(C.Structure):
    _fields_ = [("tables", C.c_int32), ("agg", C.c_int32), ("year_lo", C.c_int32),
                ("year_hi", C.c_int32), ("kind_mask", C.c_uint32), ("role_mask", C.c_uint32),
                ("title_bits", vp), ("title_bits_n", C.c_int64),
                ("info_bits", vp), ("info_bits_n", C.c_int64),
                ("type_bits", vp), ("type_bits_n", C.c_int64),
                ("kw_bits", vp), ("kw_bits_n", C.c_int64)]


class CebResult(C.Structure):
    _fields_ = [("count_lo", C.c_uint64), ("count_hi", C.c_uint64), ("min_year", C.c_int64),
                ("title_pass", C.c_int64), ("joined_movies", C.c_int64)]

CEB_MAX_TERM_PREDS, CEB_MAX_TITLE_PREDS, CEB_MAX_TERMS = 4, 2, 6


class CebPred(C.Structure):
    _fields_ = [("col", vp), ("is_null", vp), ("mask", vp), ("mask_n", C.c_int64)]


class CebTerm(C.Structure):
    _fields_ = [("rows", C.c_int64), ("movie", vp), ("movie_null", vp), ("n_preds", C.c_int32),
                ("weight", C.c_uint32), ("preds", CebPred * CEB_MAX_TERM_PREDS)]


class CebCountArgs(C.Structure):
    _fields_ = [("title_rows", C.c_int64), ("t_id", vp), ("t_id_null", vp), ("t_year", vp),
                ("t_year_null", vp), ("year_lo", C.c_int32), ("year_hi", C.c_int32),
                ("n_title_preds", C.c_int32), ("movie_cap", C.c_int32),
                ("title_preds", CebPred * CEB_MAX_TITLE_PREDS), ("n_terms", C.c_int32),
                ("reserved", C.c_int32), ("terms", CebTerm * CEB_MAX_TERMS)]


class CebCountResult(C.Structure):
    _fields_ = [("count", C.c_int64), ("title_pass", C.c_int64), ("joined", C.c_int64),
                ("term_rows", C.c_int64 * CEB_MAX_TERMS)]


MINMAX_DENSE, MINMAX_HASH = 0, 1


class MinmaxArgs(C.Structure):
    _fields_ = [("pairs", C.c_int64), ("pair_group", vp), ("pair_item", vp),
                ("items", C.c_int64), ("item_offset", vp), ("item_count", vp),
                ("item_has", vp), ("item_value", vp), ("values", C.c_int64),
                ("range_value", vp), ("groups", C.c_int32), ("keys", C.c_int32),
                ("dense_max_slots", C.c_int64)]


class MinmaxRow(C.Structure):
    _fields_ = [("group", C.c_int32), ("key", C.c_int32), ("has", C.c_int32),
                ("min", C.c_int32), ("max", C.c_int32)]


class MinmaxStats(C.Structure):
    _fields_ = [("expanded_rows", C.c_int64), ("groups", C.c_int64),
                ("placement", C.c_int32), ("pad", C.c_int32), ("slots", C.c_int64)]


EST_FK, EST_GROUP, EST_VALUE, EST_CATEGORY = 0, 1, 2, 3
EST_SAMPLE_SEED = 0x5EED


class EstColumn(C.Structure):
    _fields_ = [("col", vp), ("kind", C.c_int32), ("pad", C.c_int32), ("span", C.c_int64),
                ("hll", C.c_uint8 * 4096)]


class CebCol(C.Structure):
    _fields_ = [("v", vp), ("null", vp)]


class CebQ10aArgs(C.Structure):
    _fields_ = [("title_rows", C.c_int64), ("t_id", CebCol), ("t_kind_id", CebCol),
                ("t_year", CebCol), ("mi_rows", C.c_int64), ("mi_movie", CebCol),
                ("mi_type", CebCol), ("mi_info", CebCol), ("ci_rows", C.c_int64),
                ("ci_movie", CebCol), ("ci_person", CebCol), ("ci_role", CebCol),
                ("name_rows", C.c_int64), ("n_id", CebCol), ("n_name", CebCol),
                ("kind_ok", vp), ("kind_n", C.c_int64), ("role_ok", vp), ("role_n", C.c_int64),
                ("type_ok", vp), ("type_n", C.c_int64), ("pool_n", C.c_int64),
                ("info_dense", vp), ("info_ids", vp), ("name_ok", vp), ("n_infos", C.c_int32),
                ("reserved", C.c_int32), ("dense_max_slots", C.c_int64)]


class Q1Args(C.Structure):
    _fields_ = [("layout", C.c_int32), ("rows", C.c_int64), ("ship", vp),
                ("ship_lo", C.c_int64), ("ship_hi", C.c_int64),
                ("qty", vp), ("price", vp), ("disc_price", vp), ("disc", vp), ("tax", vp),
                ("flag_a", vp), ("flag_b", vp), ("n_a", C.c_int32), ("n_b", C.c_int32)]


class Q1Group(C.Structure):
    _fields_ = [("sum_qty", C.c_int64), ("sum_base_price", C.c_int64),
                ("sum_disc_price", C.c_int64), ("sum_charge_lo", C.c_uint64),
                ("sum_charge_hi", C.c_int64), ("sum_disc", C.c_int64),
                ("count", C.c_int64), ("reserved", C.c_int64)]


class Q1Row(C.Structure):
    _fields_ = [("returnflag", C.c_char * 8), ("linestatus", C.c_char * 8),
                ("sum_qty", C.c_double), ("sum_base_price", C.c_double),
                ("sum_disc_price", C.c_double), ("sum_charge", C.c_double),
                ("avg_qty", C.c_double), ("avg_price", C.c_double), ("avg_disc", C.c_double),
                ("count_order", C.c_int64)]


class JoinSumArgs(C.Structure):
    _fields_ = [("dim_rows", C.c_int64), ("dim_key", vp), ("n_dim_preds", C.c_int32),
                ("dim_preds", Pred * MAX_PREDS), ("fact_rows", C.c_int64),
                ("fact_fk", vp), ("fact_price", vp), ("price_dtype", C.c_int32),
                ("fact_disc", vp), ("disc_dtype", C.c_int32),
                ("capacity_hint", C.c_int64), ("placement", C.c_int32)]


class JoinSumResult(C.Structure):
    _fields_ = [("revenue", C.c_int64), ("build_rows", C.c_int64), ("matches", C.c_int64),
                ("capacity", C.c_int64), ("placement", C.c_int32), ("fallbacks", C.c_int32)]


class ChainArgs(C.Structure):
    _fields_ = [("cust_rows", C.c_int64), ("cust_key", vp), ("cust_segment", vp),
                ("segment", C.c_int32), ("order_rows", C.c_int64), ("order_key", vp),
                ("order_day", vp), ("order_cust", vp), ("day_lt", C.c_int32),
                ("fact_rows", C.c_int64), ("fact_order", vp), ("fact_price", vp),
                ("price_dtype", C.c_int32), ("fact_disc", vp), ("disc_dtype", C.c_int32),
                ("fact_disc_price", vp), ("capacity_hint", C.c_int64),
                ("order_fact_start", vp), ("order_fact_end", vp),
                ("key_ranges_known", C.c_int32), ("cust_key_min", C.c_int32),
                ("cust_key_max", C.c_int32), ("order_key_min", C.c_int32),
                ("order_key_max", C.c_int32), ("fact_ship", vp), ("ship_gt", C.c_int32),
                ("fact_order_sorted", C.c_int32)]


class GroupRow(C.Structure):
    _fields_ = [("key", C.c_int32), ("day", C.c_int32), ("revenue", C.c_int64)]


class ChainStats(C.Structure):
    _fields_ = [("cust_pass", C.c_int64), ("order_pass", C.c_int64),
                ("fact_matches", C.c_int64), ("groups", C.c_int64), ("capacity", C.c_int64),
                ("fallbacks", C.c_int32), ("sort_passes", C.c_int32)]


class GenLineitemCols(C.Structure):
    _fields_ = [("ship", vp), ("qty", vp), ("price", vp), ("disc", vp), ("tax", vp),
                ("flag_a", vp), ("flag_b", vp)]


ZONE_ROWS = 2048


class Zone(C.Structure):
    _fields_ = [("ship_min", C.c_int16), ("ship_max", C.c_int16), ("qty_min", C.c_int16),
                ("qty_max", C.c_int16), ("disc_min", C.c_uint8), ("disc_max", C.c_uint8),
                ("reserved", C.c_uint8 * 6)]


class Q6Args(C.Structure):
    _fields_ = [("rows", C.c_int64), ("ship", vp), ("disc", vp), ("qty", vp), ("price", vp),
                ("ship_lo", C.c_int32), ("ship_hi", C.c_int32), ("disc_lo", C.c_int32),
                ("disc_hi", C.c_int32), ("qty_lt", C.c_int32), ("zones", vp)]


class Q6Result(C.Structure):
    _fields_ = [("revenue", C.c_int64), ("rows_emitted", C.c_int64),
                ("rows_scanned", C.c_int64)]


class Q6Params(C.Structure):
    _fields_ = [("ship_lo", C.c_int32), ("ship_hi", C.c_int32), ("disc_lo", C.c_int32),
                ("disc_hi", C.c_int32), ("qty_lt", C.c_int32), ("month_lo", C.c_int32),
                ("month_hi", C.c_int32)]


class Q18Args(C.Structure):
    _fields_ = [("li_rows", C.c_int64), ("l_orderkey", vp), ("l_quantity", vp),
                ("orderkey_sorted", C.c_int32), ("order_rows", C.c_int64), ("o_orderkey", vp),
                ("o_custkey", vp), ("o_orderdate", vp), ("o_totalprice", vp),
                ("cust_rows", C.c_int64), ("c_custkey", vp),
                ("quantity_threshold", C.c_int32)]


class Q18Row(C.Structure):
    _fields_ = [("custkey", C.c_int32), ("orderkey", C.c_int32), ("orderdate", C.c_int32),
                ("totalprice", C.c_int32), ("sum_qty", C.c_int64), ("cust_row", C.c_int32),
                ("order_row", C.c_int32)]


class Q18Stats(C.Structure):
    _fields_ = [("rows", C.c_int64), ("dense_path", C.c_int32), ("sorted_pass", C.c_int32),
                ("fallbacks", C.c_int32), ("reserved", C.c_int32)]


class Q14Args(C.Structure):
    _fields_ = [("li_rows", C.c_int64), ("l_ship", vp), ("l_partkey", vp),
                ("l_disc_price", vp), ("ship_lo", C.c_int32), ("ship_hi", C.c_int32),
                ("zones", vp), ("part_rows", C.c_int64), ("p_partkey", vp), ("p_promo", vp)]


class Q14Result(C.Structure):
    _fields_ = [("promo_sum", C.c_int64), ("total_sum", C.c_int64),
                ("rows_emitted", C.c_int64), ("promo_revenue", C.c_double)]


class Q17Args(C.Structure):
    _fields_ = [("part_rows", C.c_int64), ("p_partkey", vp), ("n_part_preds", C.c_int32),
                ("part_preds", Pred * MAX_PREDS), ("li_rows", C.c_int64), ("l_partkey", vp),
                ("l_quantity", vp), ("l_price", vp)]


class Q17Result(C.Structure):
    _fields_ = [("sum_extendedprice", C.c_int64), ("parts", C.c_int64),
                ("matches", C.c_int64), ("rows_emitted", C.c_int64)]


class OrderSide(C.Structure):
    _fields_ = [("order_rows", C.c_int64), ("o_orderdate", vp), ("o_orderkey", vp),
                ("range_start", vp), ("range_end", vp), ("l_orderkey", vp),
                ("l_orderkey_sorted", C.c_int32)]


class Q4Args(C.Structure):
    _fields_ = [("orders", OrderSide), ("o_orderpriority", vp), ("n_priority", C.c_int32),
                ("l_commit_receipt", vp), ("date_lo", C.c_int32), ("date_hi", C.c_int32)]


class Q5Args(C.Structure):
    _fields_ = [("orders", OrderSide), ("o_cust_nation", vp), ("l_supp_nation", vp),
                ("l_disc_price", vp), ("nation_in_region", vp), ("n_nation", C.c_int32),
                ("date_lo", C.c_int32), ("date_hi", C.c_int32)]


class Q10Args(C.Structure):
    _fields_ = [("orders", OrderSide), ("o_custkey", vp), ("l_code", vp),
                ("code_lo", C.c_int32), ("code_span", C.c_int32), ("l_disc_price", vp),
                ("date_lo", C.c_int32), ("date_hi", C.c_int32), ("cust_rows", C.c_int64),
                ("c_custkey", vp), ("c_nationkey", vp), ("n_nation_names", C.c_int32)]


class Q10Row(C.Structure):
    _fields_ = [("custkey", C.c_int32), ("cust_row", C.c_int32), ("revenue", C.c_int64)]


class Q12Args(C.Structure):
    _fields_ = [("li_rows", C.c_int64), ("l_shipmode", vp), ("l_commit_receipt", vp),
                ("l_shipdate", vp), ("l_orderkey", vp), ("mode1", C.c_int32),
                ("mode2", C.c_int32), ("date_lo", C.c_int32), ("date_hi", C.c_int32),
                ("order_map", vp), ("map_size", C.c_int64), ("o_orderpriority", vp),
                ("high1", C.c_int32), ("high2", C.c_int32)]


class Q12Result(C.Structure):
    _fields_ = [("high", C.c_int64 * 2), ("low", C.c_int64 * 2)]


class Q19Group(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("brand", C.c_int32), ("containers", C.c_int32 * 4),
                ("size_hi", C.c_int32), ("qty_lo", C.c_int32)]


class Q19Args(C.Structure):
    _fields_ = [("part_rows", C.c_int64), ("p_partkey", vp), ("p_brand", vp),
                ("p_brand_dtype", C.c_int32), ("p_container", vp),
                ("p_container_dtype", C.c_int32), ("p_size", vp), ("li_rows", C.c_int64),
                ("l_partkey", vp), ("l_quantity", vp), ("l_price", vp), ("l_disc", vp),
                ("l_shipmode", vp), ("l_shipinstruct", vp), ("groups", Q19Group * 3),
                ("mode1", C.c_int32), ("mode2", C.c_int32), ("shipinstruct", C.c_int32)]


class Q21Args(C.Structure):
    _fields_ = [("order_rows", C.c_int64), ("o_orderstatus", vp), ("final_code", C.c_int32),
                ("o_orderkey", vp), ("range_start", vp), ("range_end", vp), ("l_orderkey", vp),
                ("l_orderkey_sorted", C.c_int32), ("l_suppkey", vp), ("l_commit_receipt", vp),
                ("supp_rows", C.c_int64), ("s_suppkey", vp), ("s_nationkey", vp),
                ("nation", C.c_int32)]


class Q22Args(C.Structure):
    _fields_ = [("cust_rows", C.c_int64), ("c_custkey", vp), ("c_phone_code", vp),
                ("c_phone_pair", vp), ("c_acctbal", vp), ("numeric_allowed", C.c_uint8 * 100),
                ("n_extra", C.c_int32), ("extra_pairs", C.c_int32 * 7), ("order_rows", C.c_int64),
                ("o_custkey", vp)]


class Q22Result(C.Structure):
    _fields_ = [("numcust", C.c_int64 * 100), ("totacctbal", C.c_int64 * 100),
                ("extra_numcust", C.c_int64 * 7), ("extra_totacctbal", C.c_int64 * 7)]


class Q7Args(C.Structure):
    _fields_ = [("orders", OrderSide), ("o_cust_nation", vp), ("l_shipdate", vp),
                ("l_suppkey", vp), ("l_disc_price", vp), ("supp_rows", C.c_int64),
                ("s_suppkey", vp), ("s_nationkey", vp), ("n_suppkeys", C.c_int64),
                ("nation1", C.c_int32), ("nation2", C.c_int32), ("ship_lo", C.c_int32),
                ("ship_hi", C.c_int32), ("split", C.c_int32)]


class Q8Args(C.Structure):
    _fields_ = [("orders", OrderSide), ("o_custkey", vp), ("l_partkey", vp), ("l_suppkey", vp),
                ("l_price", vp), ("l_disc", vp), ("cust_rows", C.c_int64), ("c_custkey", vp),
                ("c_nationkey", vp), ("n_cust_keys", C.c_int64), ("nation_in_region", vp),
                ("n_nation", C.c_int32), ("supp_rows", C.c_int64), ("s_suppkey", vp),
                ("s_nationkey", vp), ("n_supp_keys", C.c_int64), ("target_nation", C.c_int32),
                ("part_rows", C.c_int64), ("p_partkey", vp), ("p_type", vp),
                ("n_part_keys", C.c_int64), ("type_code", C.c_int32), ("date_lo", C.c_int32),
                ("date_hi", C.c_int32), ("split", C.c_int32)]


class Q9Args(C.Structure):
    _fields_ = [("part_rows", C.c_int64), ("p_partkey", vp), ("p_ok", vp),
                ("n_part_keys", C.c_int64), ("ps_rows", C.c_int64), ("ps_partkey", vp),
                ("ps_suppkey", vp), ("ps_supplycost", vp), ("order_rows", C.c_int64),
                ("o_orderdate", vp), ("order_map", vp), ("map_size", C.c_int64),
                ("li_rows", C.c_int64), ("l_partkey", vp), ("l_suppkey", vp), ("l_orderkey", vp),
                ("l_quantity", vp), ("l_disc_price", vp), ("l_supp_nation", vp),
                ("n_nation", C.c_int32)]


class Q13Args(C.Structure):
    _fields_ = [("order_rows", C.c_int64), ("o_custkey", vp), ("chars", vp), ("offsets", vp),
                ("alpha", vp), ("bigram", vp), ("word1", C.c_char_p), ("word2", C.c_char_p),
                ("cust_rows", C.c_int64), ("c_custkey", vp)]


class Q16Args(C.Structure):
    _fields_ = [("part_rows", C.c_int64), ("p_partkey", vp), ("p_brand", vp), ("p_type", vp),
                ("p_size", vp), ("n_brands", C.c_int32), ("n_types", C.c_int32),
                ("banned_brand", C.c_int32), ("type_banned", vp), ("size_mask", C.c_uint64),
                ("n_part_keys", C.c_int64), ("ps_rows", C.c_int64), ("ps_partkey", vp),
                ("ps_suppkey", vp), ("supp_rows", C.c_int64), ("s_suppkey", vp), ("s_banned", vp),
                ("n_supp_keys", C.c_int64)]


class Q2Args(C.Structure):
    _fields_ = [("part_rows", C.c_int64), ("p_size", vp), ("p_type", vp), ("type_ok", vp),
                ("n_types", C.c_int32), ("size", C.c_int32), ("ps_rows", C.c_int64),
                ("ps_partkey", vp), ("ps_suppkey", vp), ("ps_supplycost", vp),
                ("supp_rows", C.c_int64), ("s_nationkey", vp), ("nation_ok", vp),
                ("n_nation", C.c_int32)]


class Q11Args(C.Structure):
    _fields_ = [("supp_rows", C.c_int64), ("s_suppkey", vp), ("s_nationkey", vp),
                ("nation", C.c_int32), ("n_supp_keys", C.c_int64), ("ps_rows", C.c_int64),
                ("ps_partkey", vp), ("ps_suppkey", vp), ("ps_value", vp),
                ("fraction", C.c_double)]


class Q15Args(C.Structure):
    _fields_ = [("li_rows", C.c_int64), ("l_shipdate", vp), ("l_suppkey", vp),
                ("l_disc_price", vp), ("date_lo", C.c_int32), ("date_hi", C.c_int32),
                ("supp_rows", C.c_int64), ("s_suppkey", vp), ("n_supp_keys", C.c_int64)]


class Q20Args(C.Structure):
    _fields_ = [("part_rows", C.c_int64), ("p_partkey", vp), ("p_ok_row", vp),
                ("max_partkey", C.c_int32), ("li_rows", C.c_int64), ("l_shipdate", vp),
                ("l_partkey", vp), ("l_suppkey", vp), ("l_supp_nation", vp), ("l_quantity", vp),
                ("date_lo", C.c_int32), ("date_hi", C.c_int32), ("nation", C.c_int32),
                ("ps_rows", C.c_int64), ("ps_partkey", vp), ("ps_suppkey", vp),
                ("ps_availqty", vp), ("supp_nation_by_key", vp), ("n_supp_map", C.c_int64),
                ("max_suppkey", C.c_int32), ("capacity_hint", C.c_int64)]


class Q20Stats(C.Structure):
    _fields_ = [("pairs", C.c_int64), ("capacity", C.c_int64), ("fallbacks", C.c_int32),
                ("qualified", C.c_int64)]


P = C.POINTER
i32, i64, u64 = C.c_int32, C.c_int64, C.c_uint64

# name -> (restype, argtypes); every entry must be exported by libbolap.so.
PROTOTYPES = {
    "bolap_abi_version": (C.c_int, []),
    "bolap_last_error": (C.c_char_p, []),
    "bolap_ctx_create": (C.c_int, [C.c_int, vp, P(vp)]),
    "bolap_ctx_destroy": (C.c_int, [vp]),
    "bolap_ctx_sync": (C.c_int, [vp]),
    "bolap_ctx_device_info": (C.c_int, [vp, P(DeviceInfo)]),
    "bolap_ctx_set_trace": (C.c_int, [vp, C.c_int]),
    "bolap_ctx_read_trace": (C.c_int, [vp, P(TraceEntry), i32, P(i32)]),
    "bolap_tune_set": (C.c_int, [vp, C.c_char_p, i64]),
    "bolap_tune_get": (C.c_int, [vp, C.c_char_p, P(i64)]),
    "bolap_tune_save": (C.c_int, [vp, C.c_char_p]),
    "bolap_tune_load": (C.c_int, [vp, C.c_char_p]),
    "bolap_join_autotune": (C.c_int, [vp, P(JoinSumArgs), i32, P(C.c_double)]),
    "bolap_sort_autotune": (C.c_int, [vp, vp, vp, i64, i32, i32, P(C.c_double)]),
    "bolap_q1_scan_agg": (C.c_int, [vp, P(Q1Args), P(Q1Group)]),
    "bolap_q1_scan_agg_async": (C.c_int, [vp, P(Q1Args), vp]),
    "bolap_join_sum": (C.c_int, [vp, P(JoinSumArgs), P(JoinSumResult)]),
    "bolap_join_sum_noalloc": (C.c_int, [vp, P(JoinSumArgs), P(JoinSumResult)]),
    "bolap_chain_topk": (C.c_int, [vp, P(ChainArgs), i32, P(GroupRow), P(i32),
                                   P(ChainStats)]),
    "bolap_chain_groups_sorted": (C.c_int, [vp, P(ChainArgs), P(GroupRow), i64, P(i64),
                                            P(ChainStats)]),
    "bolap_radix_sort_u64": (C.c_int, [vp, vp, vp, i64, i32]),
    "bolap_topk_u64": (C.c_int, [vp, vp, vp, i64, i32, vp, vp]),
    "bolap_star_groupby": (C.c_int, [vp, P(StarArgs), vp, vp, i64, P(C.c_int64),
                                     P(StarStats)]),
    "bolap_hll_i32": (C.c_int, [vp, vp, i64, P(C.c_uint8)]),
    "bolap_star_sample_counts": (C.c_int, [vp, P(StarArgs), C.c_int32, u64, P(C.c_int64)]),
    "bolap_gen_iota_i32": (C.c_int, [vp, i64, C.c_int32, vp]),
    "bolap_gen_uniform_i32": (C.c_int, [vp, u64, C.c_uint32, i64, C.c_int32, C.c_int32, vp]),
    "bolap_gen_uniform_u8": (C.c_int, [vp, u64, C.c_uint32, i64, C.c_int32, C.c_int32, vp]),
    "bolap_gen_zipf_i32": (C.c_int, [vp, u64, C.c_uint32, i64, vp, i64, u64, vp]),
    "bolap_gen_price_disc": (C.c_int, [vp, u64, C.c_uint32, i64, vp, vp]),
    "bolap_ceb_run": (C.c_int, [vp, P(CebDb), P(CebQuery), P(CebResult)]),
    "bolap_ceb_count": (C.c_int, [vp, P(CebCountArgs), P(CebCountResult)]),
    "bolap_ceb_run_batch": (C.c_int, [vp, P(CebDb), P(CebQuery), C.c_int32, P(CebResult)]),
    "bolap_q3_write_csv": (C.c_int, [C.c_char_p, P(GroupRow), i64, vp]),
    "bolap_group_minmax": (C.c_int, [vp, P(MinmaxArgs), vp, i64, P(C.c_int64),
                                     P(MinmaxStats)]),
    "bolap_string_pool_create": (C.c_int, [P(vp)]),
    "bolap_string_pool_destroy": (None, [vp]),
    "bolap_string_pool_intern": (C.c_int, [vp, vp, vp, i64, vp, vp]),
    "bolap_string_pool_size": (C.c_int, [vp, P(C.c_int64), P(C.c_int64)]),
    "bolap_string_pool_export": (C.c_int, [vp, C.c_int32, vp, vp]),
    "bolap_sort_order_i32": (C.c_int, [vp, vp, vp, i64, vp]),
    "bolap_gather_rows": (C.c_int, [vp, vp, i64, C.c_int32, vp, i64, vp]),
    "bolap_ceb_autotune": (C.c_int, [vp, P(CebDb), P(CebQuery), C.c_int32, P(C.c_double)]),
    "bolap_ilike_mask": (C.c_int, [C.c_char_p, P(C.c_int64), i64, C.c_char_p, P(C.c_uint8)]),
    "bolap_ilike_mask_device": (C.c_int, [vp, vp, vp, i64, C.c_char_p, vp]),
    "bolap_estimator_create": (C.c_int, [vp, vp, vp, vp, P(vp)]),
    "bolap_estimator_destroy": (C.c_int, [vp]),
    "bolap_estimator_save": (C.c_int, [vp, C.c_char_p]),
    "bolap_estimator_load": (C.c_int, [vp, C.c_char_p, P(vp)]),
    "bolap_estimator_hints": (C.c_int, [vp, P(EstColumn), C.c_int32, P(StarArgs), C.c_int32]),
    "bolap_ceb_q10a": (C.c_int, [vp, P(CebQ10aArgs), P(MinmaxRow), i64, P(C.c_int64),
                                 P(MinmaxStats)]),
    "bolap_mlp_create": (C.c_int, [vp, u64, P(vp)]),
    "bolap_mlp_destroy": (C.c_int, [vp]),
    "bolap_mlp_num_params": (C.c_int, []),
    "bolap_mlp_get_params": (C.c_int, [vp, P(C.c_float)]),
    "bolap_mlp_set_params": (C.c_int, [vp, P(C.c_float)]),
    "bolap_mlp_predict": (C.c_int, [vp, vp, i64, vp]),
    "bolap_mlp_train_step": (C.c_int, [vp, vp, vp, i64, C.c_float, P(C.c_float)]),
    "bolap_mlp_gradients": (C.c_int, [vp, vp, vp, i64, P(C.c_float), P(C.c_float)]),
    "bolap_q1_finalize": (C.c_int, [P(Q1Group), i32, i32, i32, P(C.c_char_p), P(C.c_char_p),
                                    P(Q1Row), P(i32)]),
    "bolap_q1_write_csv": (C.c_int, [C.c_char_p, P(Q1Row), i32]),
    "bolap_q6_write_csv": (C.c_int, [C.c_char_p, P(Q6Result)]),
    "bolap_q18_write_csv": (C.c_int, [C.c_char_p, P(Q18Row), i64, P(C.c_char_p), i64]),
    "bolap_q14_parse": (C.c_int, [C.c_char_p, P(i32), P(i32)]),
    "bolap_q14": (C.c_int, [vp, P(Q14Args), P(Q14Result)]),
    "bolap_q4": (C.c_int, [vp, P(Q4Args), P(i64)]),
    "bolap_q5": (C.c_int, [vp, P(Q5Args), P(i64)]),
    "bolap_q10": (C.c_int, [vp, P(Q10Args), P(Q10Row), i64, P(i64)]),
    "bolap_q12": (C.c_int, [vp, P(Q12Args), P(Q12Result)]),
    "bolap_q19": (C.c_int, [vp, P(Q19Args), P(i64)]),
    "bolap_q21": (C.c_int, [vp, P(Q21Args), P(i64), i64]),
    "bolap_q7": (C.c_int, [vp, P(Q7Args), P(i64)]),
    "bolap_string_masks": (C.c_int, [vp, vp, vp, i64, vp, vp]),
    "bolap_q13": (C.c_int, [vp, P(Q13Args), P(i64), i32, P(i32)]),
    "bolap_q16": (C.c_int, [vp, P(Q16Args), P(i64), i64]),
    "bolap_q2": (C.c_int, [vp, P(Q2Args), vp, i64, P(i64)]),
    "bolap_q11": (C.c_int, [vp, P(Q11Args), vp, vp, i64, P(i64), P(i64)]),
    "bolap_q15": (C.c_int, [vp, P(Q15Args), vp, vp, vp, i64, P(i64), P(i64)]),
    "bolap_q20": (C.c_int, [vp, P(Q20Args), vp, P(Q20Stats)]),
    "bolap_q8": (C.c_int, [vp, P(Q8Args), P(i64), P(i64)]),
    "bolap_q9": (C.c_int, [vp, P(Q9Args), P(i64), P(C.c_uint8), P(i32), P(i32)]),
    "bolap_q22": (C.c_int, [vp, P(Q22Args), P(Q22Result)]),
    "bolap_q14_write_csv": (C.c_int, [C.c_char_p, P(Q14Result)]),
    "bolap_q17": (C.c_int, [vp, P(Q17Args), P(Q17Result)]),
    "bolap_q17_write_csv": (C.c_int, [C.c_char_p, P(Q17Result)]),
    "bolap_q6_parse": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, P(Q6Params)]),
    "bolap_q6": (C.c_int, [vp, P(Q6Args), P(Q6Result)]),
    "bolap_q6_zones": (C.c_int, [vp, i64, vp, vp, vp, vp]),
    "bolap_q6_cluster": (C.c_int, [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "bolap_q18": (C.c_int, [vp, P(Q18Args), P(Q18Row), i64, P(Q18Stats)]),
    "bolap_derive_lineitem": (C.c_int, [vp, i64, vp, vp, vp, vp, i32, vp, vp, vp, P(i32)]),
    "bolap_date_order": (C.c_int, [vp, i64, vp, vp]),
    "bolap_order_ranges": (C.c_int, [vp, i64, vp, i64, vp, i64, vp, vp, vp]),
    "bolap_gen_lineitem": (C.c_int, [vp, u64, i64, P(GenLineitemCols)]),
    "bolap_gen_dim": (C.c_int, [vp, u64, i64, vp, vp, vp]),
    "bolap_gen_fact": (C.c_int, [vp, u64, i64, i64, vp, vp, vp]),
    "bolap_gen_customers": (C.c_int, [vp, u64, i64, vp, vp]),
    "bolap_gen_orders": (C.c_int, [vp, u64, i64, i64, vp, vp, vp, vp]),
    "bolap_gen_order_facts": (C.c_int, [vp, u64, i64, vp, i64, vp, vp, vp]),
}

_lib = None


class BolapError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"bolap status {status}: {message}")
        self.status = status


def load() -> C.CDLL:
    """Loads libbolap.so, raising when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(python -m bespokeolap_artifacts_amd.build); there is no CPU fallback")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != BOLAP_OK:
        msg = load().bolap_last_error()
        raise BolapError(status, msg.decode() if msg else "")

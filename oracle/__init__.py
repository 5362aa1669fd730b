"""ctypes wrapper of the C oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

_DIR = pathlib.Path(__file__).resolve().parent
# BOLAP_ORACLE_LIB names another build of the same sources (the CPU baseline's
# -march=native copy, compiled on the host that times it).
LIB_PATH = pathlib.Path(os.environ.get("BOLAP_ORACLE_LIB", _DIR / "liboracle.so"))
_lib = None

vp = C.c_void_p
i32, i64, u64 = C.c_int32, C.c_int64, C.c_uint64


class Q1Group(C.Structure):
    _fields_ = [("sum_qty", i64), ("sum_base_price", i64), ("sum_disc_price", i64),
                ("sum_charge_lo", u64), ("sum_charge_hi", i64), ("sum_disc", i64),
                ("count", i64), ("reserved", i64)]


class Pred(C.Structure):
    _fields_ = [("col", vp), ("dtype", i32), ("op", i32), ("lo", i64), ("hi", i64),
                ("inline_set", u64 * 4), ("set", vp), ("set_bits", i64)]


class JoinSumOut(C.Structure):
    _fields_ = [("revenue", i64), ("build_rows", i64), ("matches", i64)]


class GroupRow(C.Structure):
    _fields_ = [("key", i32), ("day", i32), ("revenue", i64)]


class ChainStats(C.Structure):
    _fields_ = [("cust_pass", i64), ("order_pass", i64), ("fact_matches", i64),
                ("groups", i64)]


GROUP_DTYPE = np.dtype([("key", "<i4"), ("day", "<i4"), ("revenue", "<i8")])


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run __graft_entry__.build()")
        L = C.CDLL(str(LIB_PATH))
        L.og_mix64.restype = u64
        L.og_mix64.argtypes = [u64]
        L.og_hash.restype = u64
        L.og_hash.argtypes = [u64, C.c_uint32, u64]
        L.og_uniform.restype = i64
        L.og_uniform.argtypes = [u64, i64, i64]
        L.og_lineitem.argtypes = [u64, i64, i64] + [vp] * 7
        L.og_dim.argtypes = [u64, i64, vp, vp, vp]
        L.og_fact.argtypes = [u64, i64, i64, i64, vp, vp, vp]
        L.og_customers.argtypes = [u64, i64, vp, vp]
        L.og_orders.restype = i64
        L.og_orders.argtypes = [u64, i64, i64, vp, vp, vp, vp]
        L.og_order_facts.argtypes = [u64, i64, vp, vp, vp, vp]
        L.oracle_q1_wide.argtypes = [i64, vp, i64, i64, vp, vp, vp, vp, vp, vp, i32, i32, vp]
        L.oracle_q1_ref.argtypes = [i64, vp, i64, i64, vp, vp, vp, vp, vp, vp, i32, vp]
        L.oracle_pred_row.restype = C.c_int
        L.oracle_pred_row.argtypes = [C.POINTER(Pred), i64]
        L.oracle_join_sum.argtypes = [i64, vp, i32, C.POINTER(Pred), i64, vp, vp, i32, vp,
                                      i32, C.POINTER(JoinSumOut)]
        L.oracle_q3_ranges.restype = i64
        L.oracle_q3_ranges.argtypes = [i64, vp, vp, i32, i64, vp, vp, vp, vp, vp, i32, vp, vp,
                                       vp, i64]
        L.oracle_chain_groups.restype = i64
        L.oracle_chain_groups.argtypes = [i64, vp, vp, i32, i64, vp, vp, vp, i32, i64, vp,
                                          vp, i32, vp, i32, vp, vp, i32, vp, i64,
                                          C.POINTER(ChainStats)]
        L.oracle_sort_groups.argtypes = [vp, i64]
        L.oracle_radix_sort_u64.argtypes = [vp, vp, i64, i32]
        L.oracle_topk_u64.argtypes = [vp, vp, i64, i32, vp, vp]
        L.oracle_set_mutation.argtypes = [i32]
        _lib = L
    return _lib


def _p(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


_NP_CODES = {np.dtype(np.uint8): 0, np.dtype(np.int16): 1, np.dtype(np.int32): 2,
             np.dtype(np.int64): 3}


def dtype_code(a: np.ndarray) -> int:
    return _NP_CODES[a.dtype]


def set_mutation(m: int) -> None:
    lib().oracle_set_mutation(m)


# ----------------------------------------------------------------- generators
def gen_lineitem(seed: int, rows: int, row0: int = 0) -> dict:
    c = {"ship": np.empty(rows, np.int32), "qty": np.empty(rows, np.int64),
         "price": np.empty(rows, np.int64), "disc": np.empty(rows, np.int64),
         "tax": np.empty(rows, np.int64), "flag_a": np.empty(rows, np.uint8),
         "flag_b": np.empty(rows, np.uint8)}
    lib().og_lineitem(seed, rows, row0, *(_p(c[k]) for k in
                                          ("ship", "qty", "price", "disc", "tax", "flag_a",
                                           "flag_b")))
    return c


def gen_dim(seed: int, rows: int) -> dict:
    c = {"key": np.empty(rows, np.int32), "cat": np.empty(rows, np.uint8),
         "cont": np.empty(rows, np.uint8)}
    lib().og_dim(seed, rows, _p(c["key"]), _p(c["cat"]), _p(c["cont"]))
    return c


def gen_fact(seed: int, rows: int, dim_rows: int, row0: int = 0) -> dict:
    c = {"fk": np.empty(rows, np.int32), "price": np.empty(rows, np.int64),
         "disc": np.empty(rows, np.int64)}
    lib().og_fact(seed, rows, row0, dim_rows, _p(c["fk"]), _p(c["price"]), _p(c["disc"]))
    return c


def gen_chain(seed: int, cust_rows: int, order_rows: int) -> dict:
    t = {"cust_key": np.empty(cust_rows, np.int32),
         "cust_segment": np.empty(cust_rows, np.uint8),
         "order_key": np.empty(order_rows, np.int32),
         "order_day": np.empty(order_rows, np.int32),
         "order_cust": np.empty(order_rows, np.int32),
         "fact_offsets": np.empty(order_rows + 1, np.int64)}
    L = lib()
    L.og_customers(seed, cust_rows, _p(t["cust_key"]), _p(t["cust_segment"]))
    n = L.og_orders(seed, order_rows, cust_rows, _p(t["order_key"]), _p(t["order_day"]),
                    _p(t["order_cust"]), _p(t["fact_offsets"]))
    t["fact_order"] = np.empty(n, np.int32)
    t["fact_price"] = np.empty(n, np.int64)
    t["fact_disc"] = np.empty(n, np.int64)
    L.og_order_facts(seed, order_rows, _p(t["fact_offsets"]), _p(t["fact_order"]),
                     _p(t["fact_price"]), _p(t["fact_disc"]))
    return t


# ------------------------------------------------------------------------ Q1
def _int128(lo: int, hi: int) -> int:
    return (hi << 64) + (lo & ((1 << 64) - 1))


def _groups(out) -> list[dict]:
    return [{"sum_qty": g.sum_qty, "sum_base_price": g.sum_base_price,
             "sum_disc_price": g.sum_disc_price,
             "sum_charge": _int128(g.sum_charge_lo, g.sum_charge_hi),
             "sum_disc": g.sum_disc, "count": g.count} for g in out]


def q1_wide(c: dict, ship_lo: int, ship_hi: int, n_a: int, n_b: int) -> list[dict]:
    out = (Q1Group * (n_a * n_b))()
    lib().oracle_q1_wide(len(c["ship"]), _p(c["ship"]), ship_lo, ship_hi, _p(c["qty"]),
                         _p(c["price"]), _p(c["disc"]), _p(c["tax"]), _p(c["flag_a"]),
                         _p(c["flag_b"]), n_a, n_b, out)
    return _groups(out)


def q1_ref(c: dict, ship_lo: int, ship_hi: int, n_groups: int) -> list[dict]:
    out = (Q1Group * n_groups)()
    lib().oracle_q1_ref(len(c["ship"]), _p(c["ship"]), ship_lo, ship_hi, _p(c["qty"]),
                        _p(c["price"]), _p(c["disc_price"]), _p(c["disc"]), _p(c["tax"]),
                        _p(c["code"]), n_groups, out)
    return _groups(out)


# --------------------------------------------------------------- predicates
def make_pred(col: np.ndarray, op: int, lo: int = 0, hi: int = 0, values=None,
              set_bitmap: np.ndarray | None = None) -> Pred:
    p = Pred()
    p.col = _p(col)
    p.dtype = dtype_code(col)
    p.op = op
    p.lo, p.hi = lo, hi
    if values is not None:
        words = [0, 0, 0, 0]
        for v in values:
            words[v >> 6] |= 1 << (v & 63)
        for i in range(4):
            p.inline_set[i] = words[i]
    if set_bitmap is not None:
        p.set = _p(set_bitmap)
        p.set_bits = len(set_bitmap) * 32
    return p


def join_sum(dim_key, preds: list, fact_fk, fact_price, fact_disc) -> dict:
    arr = (Pred * max(1, len(preds)))(*preds)
    out = JoinSumOut()
    lib().oracle_join_sum(len(dim_key), _p(dim_key), len(preds), arr, len(fact_fk),
                          _p(fact_fk), _p(fact_price), dtype_code(fact_price),
                          _p(fact_disc), dtype_code(fact_disc), C.byref(out))
    return {"revenue": out.revenue, "build_rows": out.build_rows, "matches": out.matches}


def chain_groups(t: dict, segment: int, day_lt: int, disc_price=None, fact_ship=None,
                 ship_gt: int = 0):
    n_orders = len(t["order_key"])
    out = np.empty(max(1, n_orders), GROUP_DTYPE)
    st = ChainStats()
    ship = None if fact_ship is None else np.ascontiguousarray(fact_ship, np.int16)
    n = lib().oracle_chain_groups(
        len(t["cust_key"]), _p(t["cust_key"]), _p(t["cust_segment"]), segment, n_orders,
        _p(t["order_key"]), _p(t["order_day"]), _p(t["order_cust"]), day_lt,
        len(t["fact_order"]), _p(t["fact_order"]), _p(t["fact_price"]),
        dtype_code(t["fact_price"]), _p(t["fact_disc"]), dtype_code(t["fact_disc"]),
        _p(disc_price), _p(ship), ship_gt, _p(out), len(out), C.byref(st))
    stats = {"cust_pass": st.cust_pass, "order_pass": st.order_pass,
             "fact_matches": st.fact_matches, "groups": st.groups}
    return out[:n].copy(), stats


def sort_groups(rows: np.ndarray) -> np.ndarray:
    r = np.ascontiguousarray(rows.copy())
    lib().oracle_sort_groups(_p(r), len(r))
    return r


def q3_ranges(t: dict, segment: int, day_lt: int, order_sorted: dict, start, end):
    """run_q3's own algorithm (oracle_q3_ranges): `order_sorted` holds the
    order columns sorted by day, `start`/`end` each sorted order's fact
    range.  Returns every group row, sorted."""
    n_orders = len(order_sorted["order_key"])
    out = np.empty(max(1, n_orders), GROUP_DTYPE)
    s = np.ascontiguousarray(start, np.uint32)
    e = np.ascontiguousarray(end, np.uint32)
    n = lib().oracle_q3_ranges(
        len(t["cust_key"]), _p(t["cust_key"]), _p(t["cust_segment"]), segment, n_orders,
        _p(order_sorted["order_key"]), _p(order_sorted["order_day"]),
        _p(order_sorted["order_cust"]), _p(s), _p(e), day_lt, _p(t["fact_price"]),
        _p(t["fact_disc"]), _p(out), len(out))
    return out[:n].copy()


def radix_sort_u64(keys: np.ndarray, vals: np.ndarray, key_bits: int = 64):
    k = np.ascontiguousarray(keys.astype(np.uint64))
    v = np.ascontiguousarray(vals.astype(np.uint32))
    lib().oracle_radix_sort_u64(_p(k), _p(v), len(k), key_bits)
    return k, v


def topk_u64(keys: np.ndarray, vals: np.ndarray, k: int):
    m = min(k, len(keys))
    ko = np.empty(max(1, m), np.uint64)
    vo = np.empty(max(1, m), np.uint32)
    kk = np.ascontiguousarray(keys.astype(np.uint64))
    vv = np.ascontiguousarray(vals.astype(np.uint32))
    lib().oracle_topk_u64(_p(kk), _p(vv), len(kk), k, _p(ko), _p(vo))
    return ko[:m], vo[:m]


# ------------------------------------------------------ star schema (star.c)
class StarDim(C.Structure):
    _fields_ = [("rows", i64), ("key", vp), ("key_min", i32), ("key_max", i32),
                ("n_preds", i32), ("preds", C.POINTER(Pred)), ("fact_fk", vp)]


class StarStats(C.Structure):
    _fields_ = [("dim_pass", i64 * 3), ("matches", i64), ("groups", i64), ("revenue", i64)]


def _star_protos(L) -> None:
    if getattr(L, "_star_ready", False):
        return
    L.og_uniform_i32.argtypes = [u64, C.c_uint32, i64, i32, i32, vp]
    L.og_uniform_u8.argtypes = [u64, C.c_uint32, i64, i32, i32, vp]
    L.og_zipf_i32.argtypes = [u64, C.c_uint32, i64, vp, i64, u64, vp]
    L.og_price_disc.argtypes = [u64, C.c_uint32, i64, vp, vp]
    L.oracle_hll_i32.argtypes = [vp, i64, vp]
    L.oracle_star_groupby.restype = i64
    L.oracle_star_groupby.argtypes = [i32, C.POINTER(StarDim), i64, vp, i32, i32, vp, i32, vp,
                                      i32, vp, vp, i64, C.POINTER(StarStats)]
    L.oracle_star_sample_counts.argtypes = [i32, C.POINTER(StarDim), i64, vp, u64, vp]
    L._star_ready = True


def _slib():
    L = lib()
    _star_protos(L)
    return L


def zipf_i32(seed: int, stream: int, rows: int, quantiles: np.ndarray, n_keys: int,
             mul: int) -> np.ndarray:
    q = np.ascontiguousarray(quantiles, np.int32)
    out = np.empty(rows, np.int32)
    _slib().og_zipf_i32(seed, stream, rows, _p(q), n_keys, mul, _p(out))
    return out


def uniform_i32(seed: int, stream: int, rows: int, lo: int, hi: int) -> np.ndarray:
    out = np.empty(rows, np.int32)
    _slib().og_uniform_i32(seed, stream, rows, lo, hi, _p(out))
    return out


def uniform_u8(seed: int, stream: int, rows: int, lo: int, hi: int) -> np.ndarray:
    out = np.empty(rows, np.uint8)
    _slib().og_uniform_u8(seed, stream, rows, lo, hi, _p(out))
    return out


def hll_i32(col: np.ndarray) -> np.ndarray:
    c = np.ascontiguousarray(col, np.int32)
    regs = np.zeros(4096, np.uint8)
    _slib().oracle_hll_i32(_p(c), len(c), _p(regs))
    return regs


def gen_star(seed: int, fact_rows: int, dim_rows, fk_s, n_groups: int, group_s: float) -> dict:
    """Host copy of the device star schema (same generators, same streams)."""
    from bespokeolap_artifacts_amd import workload as W
    dims = []
    for j, rows in enumerate(dim_rows):
        dims.append({"key": np.arange(1, rows + 1, dtype=np.int32),
                     "cat": uniform_u8(seed, W.dim_stream(j, 1), rows, 0, W.CAT_MAX),
                     "val": uniform_i32(seed, W.dim_stream(j, 2), rows, 0, W.VAL_MAX)})
    fact = {}
    for j, rows in enumerate(dim_rows):
        fact[f"fk{j}"] = zipf_i32(seed, W.FK_STREAM + j, fact_rows,
                                  W.zipf_quantiles(rows, fk_s[j]), rows, W.coprime_mul(rows))
    fact["grp"] = zipf_i32(seed, W.GRP_STREAM, fact_rows, W.zipf_quantiles(n_groups, group_s),
                           n_groups, W.coprime_mul(n_groups))
    price = np.empty(fact_rows, np.int64)
    disc = np.empty(fact_rows, np.int64)
    _slib().og_price_disc(seed, W.PRICE_STREAM, fact_rows, _p(price), _p(disc))
    fact["price"], fact["disc"] = price, disc
    return {"dims": dims, "fact": fact, "dim_rows": list(dim_rows), "n_groups": n_groups}


def _bind_star(t: dict, plan: dict):
    keep = []
    arr = (StarDim * max(1, len(plan["dims"])))()
    for i, d in enumerate(plan["dims"]):
        j = d["dim"]
        dt = t["dims"][j]
        preds = [make_pred(dt[p["col"]], p["op"], p.get("lo", 0), p.get("hi", 0),
                           values=p.get("values")) for p in d["preds"]]
        pa = (Pred * max(1, len(preds)))(*preds)
        keep.append(pa)
        arr[i].rows = len(dt["key"])
        arr[i].key = _p(dt["key"])
        arr[i].key_min, arr[i].key_max = 1, t["dim_rows"][j]
        arr[i].n_preds = len(preds)
        arr[i].preds = C.cast(pa, C.POINTER(Pred))
        arr[i].fact_fk = _p(t["fact"][f"fk{j}"])
    return arr, keep


def _group_range(t: dict, col: str):
    return (1, t["n_groups"]) if col == "grp" else (1, t["dim_rows"][int(col[2:])])


def star_groupby(t: dict, plan: dict):
    """(stats, keys, revenues) of a star plan over host tables, groups in key order."""
    arr, keep = _bind_star(t, plan)
    gmin, gmax = _group_range(t, plan["group"])
    gk = t["fact"][plan["group"]]
    cap = gmax - gmin + 1
    keys = np.empty(cap, np.int32)
    rev = np.empty(cap, np.int64)
    st = StarStats()
    n = _slib().oracle_star_groupby(len(plan["dims"]), arr, len(gk), _p(gk), gmin, gmax,
                                    _p(t["fact"]["price"]), 3, _p(t["fact"]["disc"]), 3,
                                    _p(keys), _p(rev), cap, C.byref(st))
    stats = {"dim_pass": list(st.dim_pass)[:len(plan["dims"])], "matches": st.matches,
             "groups": st.groups, "revenue": st.revenue}
    return stats, keys[:n].copy(), rev[:n].copy()


def star_sample_counts(t: dict, plan: dict, seed: int) -> np.ndarray:
    arr, keep = _bind_star(t, plan)
    gk = t["fact"][plan["group"]]
    out = np.zeros(16, np.int64)
    _slib().oracle_star_sample_counts(len(plan["dims"]), arr, len(gk), _p(gk), seed, _p(out))
    return out


# ------------------------------------------------------- CEB queries (ceb.c)
class CebDb(C.Structure):
    _fields_ = [("title_rows", i64), ("t_year", vp), ("t_kind", vp), ("t_code", vp),
                ("ci_rows", i64), ("ci_movie", vp), ("ci_role", vp),
                ("mi_rows", i64), ("mi_movie", vp), ("mi_type", vp), ("mi_info", vp),
                ("mk_rows", i64), ("mk_movie", vp), ("mk_keyword", vp)]


class CebQuery(C.Structure):
    _fields_ = [("tables", i32), ("year_lo", i32), ("year_hi", i32), ("kind_mask", C.c_uint32),
                ("role_mask", C.c_uint32), ("title_mask", vp), ("title_n", i64),
                ("info_mask", vp), ("info_n", i64), ("type_mask", vp), ("type_n", i64),
                ("kw_mask", vp), ("kw_n", i64)]


def gen_ceb(seed: int, scale: float) -> dict:
    """Host copy of the device CEB tables (same generators and streams; the
    fact tables' movie ids sorted ascending)."""
    from bespokeolap_artifacts_amd import ceb as CB
    from bespokeolap_artifacts_amd import workload as W
    n = {k: max(1, int(v * scale)) if k != "info_type" else v for k, v in CB.SIZES.items()}
    nt = n["title"]

    def zipf_sorted(rows, stream, keys):
        return np.sort(zipf_i32(seed, stream, rows, W.zipf_quantiles(keys, CB.ZIPF_S), keys,
                                W.coprime_mul(keys)))

    t = {"n": n,
         "title": {"year": uniform_i32(seed, CB.S_T_YEAR, nt, CB.YEAR_LO, CB.YEAR_HI),
                   "kind": uniform_i32(seed, CB.S_T_KIND, nt, 1, 7),
                   "code": uniform_i32(seed, CB.S_T_CODE, nt, 0, CB.DICT_SIZES["title"] - 1)},
         "ci": {"movie": zipf_sorted(n["cast_info"], CB.S_CI_MOVIE, nt),
                "role": uniform_i32(seed, CB.S_CI_ROLE, n["cast_info"], 1, 12)},
         "mi": {"movie": zipf_sorted(n["movie_info"], CB.S_MI_MOVIE, nt),
                "type": uniform_i32(seed, CB.S_MI_TYPE, n["movie_info"], 1, n["info_type"]),
                "info": uniform_i32(seed, CB.S_MI_INFO, n["movie_info"], 0,
                                    CB.DICT_SIZES["info"] - 1)},
         "mk": {"movie": zipf_sorted(n["movie_keyword"], CB.S_MK_MOVIE, nt),
                "keyword": zipf_i32(seed, CB.S_MK_KW, n["movie_keyword"],
                                    W.zipf_quantiles(n["keyword"], CB.ZIPF_S), n["keyword"],
                                    W.coprime_mul(n["keyword"]))}}
    return t


def ceb_run(t: dict, q: dict, masks: dict) -> dict:
    """One query over host tables.  masks: optional uint8 arrays 'title' (over
    title codes), 'info' (over info codes), 'info_type' and 'keyword' (over
    ids, index = id)."""
    L = lib()
    if not getattr(L, "_ceb_ready", False):
        L.oracle_ceb_run.argtypes = [C.POINTER(CebDb), C.POINTER(CebQuery), vp]
        L._ceb_ready = True
    db = CebDb(len(t["title"]["year"]), _p(t["title"]["year"]), _p(t["title"]["kind"]),
               _p(t["title"]["code"]), len(t["ci"]["movie"]), _p(t["ci"]["movie"]),
               _p(t["ci"]["role"]), len(t["mi"]["movie"]), _p(t["mi"]["movie"]),
               _p(t["mi"]["type"]), _p(t["mi"]["info"]), len(t["mk"]["movie"]),
               _p(t["mk"]["movie"]), _p(t["mk"]["keyword"]))
    cq = CebQuery(q["tables"], q["year_lo"], q["year_hi"], q["kind_mask"], q["role_mask"])
    keep = []
    for field, key in (("title", "title"), ("info", "info"), ("type", "info_type"),
                       ("kw", "keyword")):
        m = masks.get(key)
        if m is None:
            continue
        m = np.ascontiguousarray(m, np.uint8)
        keep.append(m)
        setattr(cq, f"{field}_mask", _p(m))
        setattr(cq, f"{field}_n", len(m))
    out = np.zeros(5, np.uint64)
    L.oracle_ceb_run(C.byref(db), C.byref(cq), _p(out))
    return {"count": (int(out[1]) << 64) | int(out[0]), "min_year": int(out[2].view(np.int64)),
            "title_pass": int(out[3]), "joined_movies": int(out[4])}


# --------------------------------------------- TPC-H Q6 / Q18, build (tpch.c)
class Q6Out(C.Structure):
    _fields_ = [("revenue", i64), ("rows_emitted", i64), ("rows_scanned", i64)]


class Q18Row(C.Structure):
    _fields_ = [("custkey", i32), ("orderkey", i32), ("orderdate", i32), ("totalprice", i32),
                ("sum_qty", i64), ("cust_row", i32), ("order_row", i32)]


Q18_DTYPE = np.dtype([("custkey", "<i4"), ("orderkey", "<i4"), ("orderdate", "<i4"),
                      ("totalprice", "<i4"), ("sum_qty", "<i8"), ("cust_row", "<i4"),
                      ("order_row", "<i4")])


class Q14Out(C.Structure):
    _fields_ = [("promo_sum", i64), ("total_sum", i64), ("rows_emitted", i64),
                ("promo_revenue", C.c_double)]


class Q17Out(C.Structure):
    _fields_ = [("sum_extendedprice", i64), ("parts", i64), ("matches", i64),
                ("rows_emitted", i64)]


def _tpch_lib():
    L = lib()
    if not getattr(L, "_tpch_ready", False):
        L.oracle_days_from_civil.restype = i64
        L.oracle_days_from_civil.argtypes = [i32, i32, i32]
        L.oracle_civil_from_days.argtypes = [i32, vp, vp, vp]
        L.oracle_q6_flat.argtypes = [i64, vp, vp, vp, vp, i32, i32, i32, i32, i32,
                                     C.POINTER(Q6Out)]
        L.oracle_q6_sharded.argtypes = [i64, vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, i32,
                                        C.POINTER(Q6Out)]
        L.oracle_q18.restype = i64
        L.oracle_q18.argtypes = [i64, vp, vp, i64, vp, vp, vp, vp, i64, vp, i32, vp, i64, vp]
        L.oracle_derive_lineitem.restype = i32
        L.oracle_derive_lineitem.argtypes = [i64, vp, vp, vp, vp, i32, vp, vp, vp]
        L.oracle_date_order.argtypes = [i64, vp, vp]
        L.oracle_order_ranges.argtypes = [i64, vp, i64, vp, i64, vp, i32, vp, vp]
        L.oracle_q14.argtypes = [i64, vp, vp, vp, i32, i32, i64, vp, vp, C.POINTER(Q14Out)]
        L.oracle_q17.argtypes = [i64, vp, i32, C.POINTER(Pred), i64, vp, vp, vp,
                                 C.POINTER(Q17Out)]
        L._tpch_ready = True
    return L


def days_from_civil(y: int, m: int, d: int) -> int:
    return int(_tpch_lib().oracle_days_from_civil(y, m, d))


def civil_from_days(z: int) -> tuple[int, int, int]:
    y, m, d = C.c_int32(), C.c_int32(), C.c_int32()
    _tpch_lib().oracle_civil_from_days(z, C.byref(y), C.byref(m), C.byref(d))
    return y.value, m.value, d.value


def _q6_cols(c: dict):
    cols = [np.ascontiguousarray(c[k], dt) for k, dt in
            (("ship", np.int16), ("disc", np.uint8), ("qty", np.int16), ("price", np.int32))]
    return cols, [_p(x) for x in cols]


def q6_flat(c: dict, p: dict) -> dict:
    cols, ptrs = _q6_cols(c)
    out = Q6Out()
    _tpch_lib().oracle_q6_flat(len(cols[0]), *ptrs, p["ship_lo"], p["ship_hi"], p["disc_lo"],
                               p["disc_hi"], p["qty_lt"], C.byref(out))
    return {"revenue": out.revenue, "rows_emitted": out.rows_emitted,
            "rows_scanned": out.rows_scanned}


def q6_sharded(c: dict, p: dict) -> dict:
    cols, ptrs = _q6_cols(c)
    out = Q6Out()
    _tpch_lib().oracle_q6_sharded(len(cols[0]), *ptrs, p["ship_lo"], p["ship_hi"],
                                  p["disc_lo"], p["disc_hi"], p["qty_lt"], p["month_lo"],
                                  p["month_hi"], C.byref(out))
    return {"revenue": out.revenue, "rows_emitted": out.rows_emitted,
            "rows_scanned": out.rows_scanned}


def q18(t: dict, threshold: int, max_out: int = 1 << 20):
    """Rows (structured array, Q18_DTYPE) in output order and the dense flag."""
    a = {k: np.ascontiguousarray(t[k], dt) for k, dt in
         (("l_orderkey", np.int32), ("l_qty", np.int16), ("o_orderkey", np.int32),
          ("o_custkey", np.int32), ("o_orderdate", np.int16), ("o_totalprice", np.int32),
          ("c_custkey", np.int32))}
    out = np.zeros(max(1, max_out), Q18_DTYPE)
    dense = C.c_int32()
    n = _tpch_lib().oracle_q18(len(a["l_orderkey"]), _p(a["l_orderkey"]), _p(a["l_qty"]),
                               len(a["o_orderkey"]), _p(a["o_orderkey"]), _p(a["o_custkey"]),
                               _p(a["o_orderdate"]), _p(a["o_totalprice"]),
                               len(a["c_custkey"]), _p(a["c_custkey"]), threshold, _p(out),
                               max_out, C.byref(dense))
    assert n <= max_out, (n, max_out)
    return out[:n], bool(dense.value)


def derive_lineitem(price, disc, rf, ls, n_ls: int, orderkey):
    price = np.ascontiguousarray(price, np.int32)
    disc, rf, ls = (np.ascontiguousarray(x, np.uint8) for x in (disc, rf, ls))
    orderkey = np.ascontiguousarray(orderkey, np.int32)
    dp = np.empty(len(price), np.int32)
    code = np.empty(len(price), np.uint8)
    s = _tpch_lib().oracle_derive_lineitem(len(price), _p(price), _p(disc), _p(rf), _p(ls),
                                           n_ls, _p(orderkey), _p(dp), _p(code))
    return dp, code, bool(s)


def date_order(date) -> np.ndarray:
    date = np.ascontiguousarray(date, np.int16)
    out = np.empty(max(1, len(date)), np.uint32)
    _tpch_lib().oracle_date_order(len(date), _p(date), _p(out))
    return out[:len(date)]


def order_ranges(o_orderkey, l_orderkey, li_sorted: bool):
    ok = np.ascontiguousarray(o_orderkey, np.int32)
    lk = np.ascontiguousarray(l_orderkey, np.int32)
    map_size = int(ok.max()) + 1 if len(ok) and ok.max() >= 0 else 0
    m = np.empty(max(1, map_size), np.int32)
    start = np.empty(max(1, len(ok)), np.uint32)
    end = np.empty(max(1, len(ok)), np.uint32)
    _tpch_lib().oracle_order_ranges(len(ok), _p(ok), map_size, _p(m), len(lk), _p(lk),
                                    int(li_sorted), _p(start), _p(end))
    return m[:map_size], start[:len(ok)], end[:len(ok)]


def q14(li: dict, part: dict, ship_lo: int, ship_hi: int) -> dict:
    a = [np.ascontiguousarray(li[k], dt) for k, dt in
         (("ship", np.int16), ("partkey", np.int32), ("disc_price", np.int32))]
    pk = np.ascontiguousarray(part["partkey"], np.int32)
    pr = np.ascontiguousarray(part["promo"], np.uint8)
    out = Q14Out()
    _tpch_lib().oracle_q14(len(a[0]), *[_p(x) for x in a], ship_lo, ship_hi, len(pk), _p(pk),
                           _p(pr), C.byref(out))
    return {"promo_sum": out.promo_sum, "total_sum": out.total_sum,
            "rows_emitted": out.rows_emitted, "promo_revenue": out.promo_revenue}


def q17(part_key, preds: list, li: dict) -> dict:
    pk = np.ascontiguousarray(part_key, np.int32)
    a = [np.ascontiguousarray(li[k], dt) for k, dt in
         (("partkey", np.int32), ("qty", np.int16), ("price", np.int32))]
    arr = (Pred * max(1, len(preds)))(*preds)
    out = Q17Out()
    _tpch_lib().oracle_q17(len(pk), _p(pk), len(preds), arr, len(a[0]), *[_p(x) for x in a],
                           C.byref(out))
    return {"sum_extendedprice": out.sum_extendedprice, "parts": out.parts,
            "matches": out.matches, "rows_emitted": out.rows_emitted}


# ------------------------------------------------ grouped MIN/MAX (ceb.c)
MINMAX_DTYPE = np.dtype([("group", "<i4"), ("key", "<i4"), ("has", "<i4"), ("min", "<i4"),
                         ("max", "<i4")])


def group_minmax(pair_group, pair_item, item_offset, item_count, item_has, item_value,
                 range_value, groups: int, keys: int, dense_max: int = 0) -> np.ndarray:
    """Grouped MIN/MAX rows in (group, key) order (item_has may be None)."""
    L = lib()
    if not getattr(L, "_mm_ready", False):
        L.oracle_group_minmax.restype = i64
        L.oracle_group_minmax.argtypes = [i64, vp, vp, i64, vp, vp, vp, vp, i64, vp, i32, i32,
                                          i64, vp, i64]
        L._mm_ready = True
    c32 = lambda a: np.ascontiguousarray(a, np.int32)
    pg, pi, off, cnt = c32(pair_group), c32(pair_item), c32(item_offset), c32(item_count)
    val, rv = c32(item_value), c32(range_value)
    hv = None if item_has is None else np.ascontiguousarray(item_has, np.uint8)
    cap = max(1, min(int(cnt[pi[pg >= 0]].astype(np.int64).sum()) if len(pg) else 0,
                     groups * keys))
    out = np.zeros(cap, MINMAX_DTYPE)
    n = L.oracle_group_minmax(len(pg), _p(pg), _p(pi), len(off), _p(off), _p(cnt),
                              None if hv is None else _p(hv), _p(val), len(rv), _p(rv),
                              groups, keys, dense_max, _p(out), cap)
    if n < 0:
        raise ValueError("oracle_group_minmax: out of bounds" if n == -1 else "short output")
    return out[:n]


def sort_order_i32(primary, secondary=None) -> np.ndarray:
    """Row order sorting by (primary, secondary) ascending, ties in input order."""
    L = lib()
    if not getattr(L, "_so_ready", False):
        L.oracle_sort_order_i32.restype = None
        L.oracle_sort_order_i32.argtypes = [vp, vp, i64, vp]
        L._so_ready = True
    p = np.ascontiguousarray(primary, np.int32)
    s = None if secondary is None else np.ascontiguousarray(secondary, np.int32)
    out = np.zeros(len(p), np.uint32)
    L.oracle_sort_order_i32(_p(p), None if s is None else _p(s), len(p), _p(out))
    return out


# ------------------------------------------- Q4, Q5, Q10, Q12, Q19 (tpch2.c)
Q10_DTYPE = np.dtype([("custkey", "<i4"), ("cust_row", "<i4"), ("revenue", "<i8")])


class Q19Group(C.Structure):
    _fields_ = [("enabled", i32), ("brand", i32), ("containers", i32 * 4), ("size_hi", i32),
                ("qty_lo", i32)]


def _tpch2():
    L = lib()
    if not getattr(L, "_tpch2_ready", False):
        L.oracle_q4.restype = None
        L.oracle_q4.argtypes = [i64, vp, vp, vp, i32, vp, vp, vp, i32, vp, i32, i32, vp]
        L.oracle_q5.restype = None
        L.oracle_q5.argtypes = [i64, vp, vp, vp, vp, vp, vp, i32, vp, vp, vp, i32, i32, i32, vp]
        L.oracle_q10.restype = i64
        L.oracle_q10.argtypes = [i64, vp, vp, vp, vp, vp, vp, i32, vp, i32, i32, vp, i32, i32,
                                 i64, vp, vp, i32, vp, i64]
        L.oracle_q12.restype = None
        L.oracle_q12.argtypes = [i64, vp, vp, vp, vp, i32, i32, i32, i32, vp, i64, vp, i32, i32,
                                 vp, vp]
        L.oracle_q19.restype = i64
        L.oracle_q19.argtypes = [i64, vp, vp, vp, vp, i64, vp, vp, vp, vp, vp, vp,
                                 C.POINTER(Q19Group), i32, i32, i32]
        L._tpch2_ready = True
    return L


def _c(a, dt):
    return np.ascontiguousarray(a, dt)


def q4(db: dict, date_lo: int, date_hi: int, n_priority: int) -> np.ndarray:
    """Per-priority-code counts of run_q4 (oracle_q4)."""
    out = np.zeros(max(1, n_priority), np.int64)
    a = [_c(db["o_orderkey"], np.int32), _c(db["o_orderdate"], np.int16),
         _c(db["o_orderpriority"], np.uint16)]
    r = [_c(db["r_start"], np.uint32), _c(db["r_end"], np.uint32)]
    lk, cr = _c(db["l_orderkey"], np.int32), _c(db["l_commit_receipt"], np.uint32)
    _tpch2().oracle_q4(len(a[0]), _p(a[0]), _p(a[1]), _p(a[2]), n_priority, _p(r[0]), _p(r[1]),
                       _p(lk), int(db["l_sorted"]), _p(cr), date_lo, date_hi, _p(out))
    return out[:n_priority]


def q5(db: dict, nation_in_region: np.ndarray, date_lo: int, date_hi: int) -> np.ndarray:
    """Revenue per nation key of run_q5 (oracle_q5)."""
    nr = _c(nation_in_region, np.uint8)
    out = np.zeros(max(1, len(nr)), np.int64)
    a = [_c(db["o_orderkey"], np.int32), _c(db["o_orderdate"], np.int16),
         _c(db["o_cust_nation"], np.int16), _c(db["r_start"], np.uint32),
         _c(db["r_end"], np.uint32), _c(db["l_orderkey"], np.int32)]
    sn, dp = _c(db["l_supp_nation"], np.int16), _c(db["l_disc_price"], np.int32)
    _tpch2().oracle_q5(len(a[0]), *[_p(x) for x in a], int(db["l_sorted"]), _p(sn), _p(dp),
                       _p(nr), len(nr), date_lo, date_hi, _p(out))
    return out[:len(nr)]


def q10(db: dict, code_lo: int, code_span: int, date_lo: int, date_hi: int,
        n_nation_names: int) -> np.ndarray:
    """(custkey, cust_row, revenue) rows of run_q10 in its order (oracle_q10)."""
    a = [_c(db["o_orderkey"], np.int32), _c(db["o_orderdate"], np.int16),
         _c(db["o_custkey"], np.int32), _c(db["r_start"], np.uint32),
         _c(db["r_end"], np.uint32), _c(db["l_orderkey"], np.int32)]
    code, dp = _c(db["l_code"], np.uint8), _c(db["l_disc_price"], np.int32)
    ck, nk = _c(db["c_custkey"], np.int32), _c(db["c_nationkey"], np.int32)
    cap = max(1, len(ck))
    out = np.zeros(cap, Q10_DTYPE)
    n = _tpch2().oracle_q10(len(a[0]), *[_p(x) for x in a], int(db["l_sorted"]), _p(code),
                            code_lo, code_span, _p(dp), date_lo, date_hi, len(ck), _p(ck),
                            _p(nk), n_nation_names, _p(out), cap)
    assert n <= cap
    return out[:n]


def q12(db: dict, mode1: int, mode2: int, date_lo: int, date_hi: int, high1: int,
        high2: int):
    """((high, low) for mode1, (high, low) for mode2) of run_q12 (oracle_q12)."""
    hi = np.zeros(2, np.int64)
    lo = np.zeros(2, np.int64)
    a = [_c(db["l_shipmode"], np.uint16), _c(db["l_commit_receipt"], np.uint32),
         _c(db["l_shipdate"], np.int16), _c(db["l_orderkey"], np.int32)]
    m, pr = _c(db["order_map"], np.int32), _c(db["o_orderpriority"], np.uint16)
    _tpch2().oracle_q12(len(a[0]), *[_p(x) for x in a], mode1, mode2, date_lo, date_hi, _p(m),
                        len(m), _p(pr), high1, high2, _p(hi), _p(lo))
    return (int(hi[0]), int(lo[0])), (int(hi[1]), int(lo[1]))


def q19_groups(groups) -> "C.Array":
    """groups: 3 tuples (enabled, brand, [4 container codes], size_hi, qty_lo)."""
    arr = (Q19Group * 3)()
    for g, (en, br, cont, sh, ql) in enumerate(groups):
        arr[g].enabled, arr[g].brand, arr[g].size_hi, arr[g].qty_lo = int(en), br, sh, ql
        for c in range(4):
            arr[g].containers[c] = int(cont[c])
    return arr


def q19(db: dict, groups, mode1: int, mode2: int, instruct: int) -> int:
    """Revenue of run_q19 (oracle_q19)."""
    p = [_c(db["p_partkey"], np.int32), _c(db["p_brand"], np.uint16),
         _c(db["p_container"], np.uint16), _c(db["p_size"], np.int32)]
    li = [_c(db["l_partkey"], np.int32), _c(db["l_quantity"], np.int16),
          _c(db["l_extendedprice"], np.int32), _c(db["l_discount"], np.uint8),
          _c(db["l_shipmode"], np.uint16), _c(db["l_shipinstruct"], np.uint16)]
    return int(_tpch2().oracle_q19(len(p[0]), *[_p(x) for x in p], len(li[0]),
                                   *[_p(x) for x in li], q19_groups(groups), mode1, mode2,
                                   instruct))


def q21(db: dict, final_code: int, target_nation: int, n_suppkeys: int) -> np.ndarray:
    """Late-line counts per supplier key of run_q21 (oracle_q21)."""
    L = lib()
    if not getattr(L, "_q21_ready", False):
        L.oracle_q21.restype = None
        L.oracle_q21.argtypes = [i64, vp, vp, i32, vp, vp, vp, i32, vp, vp, vp, i64, vp]
        L.oracle_q22.restype = None
        L.oracle_q22.argtypes = [i64, vp, vp, vp, vp, vp, i32, vp, i64, vp, vp, vp]
        L._q21_ready = True
    tgt = np.zeros(max(1, n_suppkeys), np.uint8)
    sk, sn = db["s_suppkey"], db["s_nationkey"]
    for k, nk in zip(sk.tolist(), sn.tolist()):
        if 0 <= k < n_suppkeys:
            tgt[k] = 1 if nk == target_nation else 0   # last row of a key wins
    out = np.zeros(max(1, n_suppkeys), np.int64)
    a = [_c(db["o_orderkey"], np.int32), _c(db["o_orderstatus"], np.uint16)]
    r = [_c(db["r_start"], np.uint32), _c(db["r_end"], np.uint32), _c(db["l_orderkey"], np.int32)]
    li = [_c(db["l_suppkey"], np.int32), _c(db["l_commit_receipt"], np.uint32)]
    L.oracle_q21(len(a[0]), _p(a[0]), _p(a[1]), final_code, _p(r[0]), _p(r[1]), _p(r[2]),
                 int(db["l_sorted"]), _p(li[0]), _p(li[1]), _p(tgt), n_suppkeys, _p(out))
    return out[:n_suppkeys]


def q22(db: dict, numeric_codes, extra_pairs):
    """(numeric[100, 2], extra[n, 2]) = (customers, acctbal sum) per prefix of
    run_q22 (oracle_q22)."""
    q21  # noqa: B018 (argtypes are registered together)
    L = lib()
    if not getattr(L, "_q21_ready", False):
        q21({"s_suppkey": np.zeros(0, np.int32), "s_nationkey": np.zeros(0, np.int32),
             "o_orderkey": np.zeros(0, np.int32), "o_orderstatus": np.zeros(0, np.uint16),
             "r_start": np.zeros(0, np.uint32), "r_end": np.zeros(0, np.uint32),
             "l_orderkey": np.zeros(0, np.int32), "l_sorted": 1,
             "l_suppkey": np.zeros(0, np.int32), "l_commit_receipt": np.zeros(0, np.uint32)},
            0, 0, 0)
    allowed = np.zeros(100, np.uint8)
    for c in numeric_codes:
        allowed[c] = 1
    ex = _c(np.asarray(list(extra_pairs) or [0], np.int32), np.int32)
    num = np.zeros(200, np.int64)
    ext = np.zeros(max(2, 2 * len(extra_pairs)), np.int64)
    c = [_c(db["c_custkey"], np.int32), _c(db["c_phone_code"], np.uint8),
         _c(db["c_phone_pair"], np.int32), _c(db["c_acctbal"], np.int32)]
    ok = _c(db["o_custkey"], np.int32)
    L.oracle_q22(len(c[0]), *[_p(x) for x in c], _p(allowed), len(extra_pairs), _p(ex), len(ok),
                 _p(ok), _p(num), _p(ext))
    return num.reshape(100, 2), ext[:2 * len(extra_pairs)].reshape(-1, 2)


def _tpch3():
    L = lib()
    if not getattr(L, "_tpch3_ready", False):
        L.oracle_q7.restype = None
        L.oracle_q7.argtypes = [i64, vp, vp, vp, vp, vp, i32, vp, vp, vp, i64, vp, vp, i32, i32,
                                i32, i32, i32, vp]
        L.oracle_q8.restype = None
        L.oracle_q8.argtypes = [i64, vp, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp, i64, vp, vp, vp,
                                i32, i64, vp, vp, i32, i64, vp, vp, i32, i32, i32, i32, vp, vp]
        L.oracle_q9.restype = i32
        L.oracle_q9.argtypes = [i64, vp, vp, i64, vp, vp, vp, i64, vp, vp, i64, vp, vp, vp, vp,
                                vp, vp, i32, vp, vp, vp]
        L._tpch3_ready = True
    return L


def _orders_side(db):
    return [_c(db["o_orderkey"], np.int32), _c(db["r_start"], np.uint32),
            _c(db["r_end"], np.uint32), _c(db["l_orderkey"], np.int32)]


def q7(db: dict, n1: int, n2: int, ship_lo: int, ship_hi: int, split: int) -> np.ndarray:
    """rev[supp, cust, year] of run_q7 (oracle_q7), shape (2, 2, 2)."""
    ok, rs, re_, lk = _orders_side(db)
    cn = _c(db["o_cust_nation"], np.int16)
    li = [_c(db["l_shipdate"], np.int16), _c(db["l_suppkey"], np.int32),
          _c(db["l_disc_price"], np.int32)]
    s = [_c(db["s_suppkey"], np.int32), _c(db["s_nationkey"], np.int32)]
    out = np.zeros(8, np.int64)
    _tpch3().oracle_q7(len(ok), _p(ok), _p(cn), _p(rs), _p(re_), _p(lk), int(db["l_sorted"]),
                       *[_p(x) for x in li], len(s[0]), _p(s[0]), _p(s[1]), n1, n2, ship_lo,
                       ship_hi, split, _p(out))
    return out.reshape(2, 2, 2)


def q8(db: dict, nation_in_region, target: int, type_code: int, date_lo: int, date_hi: int,
       split: int):
    """(total[2], nation[2]) volumes of run_q8 (oracle_q8)."""
    ok, rs, re_, lk = _orders_side(db)
    od, oc = _c(db["o_orderdate"], np.int16), _c(db["o_custkey"], np.int32)
    li = [_c(db["l_partkey"], np.int32), _c(db["l_suppkey"], np.int32),
          _c(db["l_extendedprice"], np.int32), _c(db["l_discount"], np.uint8)]
    c = [_c(db["c_custkey"], np.int32), _c(db["c_nationkey"], np.int32)]
    nr = _c(nation_in_region, np.uint8)
    s = [_c(db["s_suppkey"], np.int32), _c(db["s_nationkey"], np.int32)]
    p = [_c(db["p_partkey"], np.int32), _c(db["p_type"], np.uint16)]
    tot, nat = np.zeros(2, np.int64), np.zeros(2, np.int64)
    _tpch3().oracle_q8(len(ok), _p(ok), _p(od), _p(oc), _p(rs), _p(re_), _p(lk),
                       int(db["l_sorted"]), *[_p(x) for x in li], len(c[0]), _p(c[0]), _p(c[1]),
                       _p(nr), len(nr), len(s[0]), _p(s[0]), _p(s[1]), target, len(p[0]),
                       _p(p[0]), _p(p[1]), type_code, date_lo, date_hi, split, _p(tot), _p(nat))
    return tot, nat


def q9(db: dict, p_ok_row, n_nation: int):
    """(min_year, totals[n_nation, span], used[n_nation, span]) of run_q9."""
    po = _c(p_ok_row, np.uint8)
    p = _c(db["p_partkey"], np.int32)
    ps = [_c(db["ps_partkey"], np.int32), _c(db["ps_suppkey"], np.int32),
          _c(db["ps_supplycost"], np.int32)]
    o = [_c(db["o_orderkey"], np.int32), _c(db["o_orderdate"], np.int16)]
    li = [_c(db["l_partkey"], np.int32), _c(db["l_suppkey"], np.int32),
          _c(db["l_orderkey"], np.int32), _c(db["l_quantity"], np.int16),
          _c(db["l_disc_price"], np.int32), _c(db["l_supp_nation"], np.int16)]
    cap = n_nation * 64
    tot, used = np.zeros(cap, np.int64), np.zeros(cap, np.uint8)
    y0 = C.c_int32(0)
    span = _tpch3().oracle_q9(len(p), _p(p), _p(po), len(ps[0]), *[_p(x) for x in ps], len(o[0]),
                              _p(o[0]), _p(o[1]), len(li[0]), *[_p(x) for x in li], n_nation,
                              _p(tot), _p(used), C.byref(y0))
    assert span <= 64
    return y0.value, tot[:n_nation * span].reshape(n_nation, span), \
        used[:n_nation * span].reshape(n_nation, span)


def q13(db: dict, w1: str, w2: str):
    """(max count, dist[max count + 1]) of run_q13 (oracle_q13) over the raw
    order rows (o_custkey_raw with their comments)."""
    L = lib()
    if not getattr(L, "_q13_ready", False):
        L.oracle_q13.restype = i32
        L.oracle_q13.argtypes = [i64, vp, vp, vp, C.c_char_p, C.c_char_p, i64, vp, vp, i64]
        L.oracle_q16.restype = i64
        L.oracle_q16.argtypes = [i64, vp, vp, vp, vp, i32, i32, vp, u64, i64, i64, vp, vp, vp,
                                 i64, vp, i64]
        L._q13_ready = True
    ok = _c(db["o_custkey_raw"], np.int32)
    ch, off = _c(db["o_comment_chars"], np.uint8), _c(db["o_comment_offsets"], np.uint32)
    ck = _c(db["c_custkey"], np.int32)
    cap = 4096
    dist = np.zeros(cap, np.int64)
    mx = L.oracle_q13(len(ok), _p(ok), _p(ch), _p(off), w1.encode(), w2.encode(), len(ck), _p(ck),
                      _p(dist), cap)
    assert mx < cap
    return mx, dist[:mx + 1]


def q16(db: dict, banned_brand: int, type_banned, size_mask: int, n_part_keys: int,
        n_supp_keys: int):
    """counts[group] (group = (brand * n_types + type) * 51 + size) of run_q16."""
    q13  # noqa: B018
    L = lib()
    if not getattr(L, "_q13_ready", False):
        q13({"o_custkey_raw": np.zeros(0, np.int32), "o_comment_chars": np.zeros(1, np.uint8),
             "o_comment_offsets": np.zeros(1, np.uint32), "c_custkey": np.zeros(0, np.int32)},
            "a", "b")
    tb = _c(type_banned, np.uint8)
    n_types = len(tb)
    n_counts = len(db["brand_dict"]) * n_types * 51
    counts = np.zeros(max(1, n_counts), np.int64)
    p = [_c(db["p_partkey"], np.int32), _c(db["p_brand"], np.uint16), _c(db["p_type"], np.uint16),
         _c(db["p_size"], np.int32)]
    ps = [_c(db["ps_partkey"], np.int32), _c(db["ps_suppkey"], np.int32)]
    sb = np.zeros(max(1, n_supp_keys), np.uint8)
    for k, b in zip(db["s_suppkey"].tolist(), db["s_banned"].tolist()):
        if 0 <= k < n_supp_keys and b:
            sb[k] = 1
    L.oracle_q16(len(p[0]), *[_p(x) for x in p], n_types, banned_brand, _p(tb), size_mask,
                 n_part_keys, len(ps[0]), _p(ps[0]), _p(ps[1]), _p(sb), n_supp_keys, _p(counts),
                 n_counts)
    return counts[:n_counts]


# --------------------------------------------------- Q2, Q11, Q15, Q20 (tpch3.c)
def _tpch4():
    L = lib()
    if not getattr(L, "_tpch4_ready", False):
        f64 = C.c_double
        L.oracle_q2.restype = i64
        L.oracle_q2.argtypes = [i64, vp, vp, vp, i32, i64, vp, vp, vp, i64, vp, vp, vp, i64]
        L.oracle_q11.restype = i64
        L.oracle_q11.argtypes = [i64, vp, vp, i32, i64, vp, vp, vp, f64, vp, vp, i64, vp, vp]
        L.oracle_q15.restype = i64
        L.oracle_q15.argtypes = [i64, vp, vp, vp, i32, i32, i64, vp, vp, vp, vp, i64, vp]
        L.oracle_q20.restype = i64
        L.oracle_q20.argtypes = [i64, vp, vp, i64, vp, vp, vp, vp, vp, i32, i32, i32, i64, vp, vp,
                                 vp, vp, i64, i32, vp, vp]
        L._tpch4_ready = True
    return L


def q2(db: dict, type_ok, size: int, s_ok) -> np.ndarray:
    """(part row, supplier row) pairs of run_q2 before its sort, in the
    reference's emission order (oracle_q2)."""
    to = _c(type_ok, np.uint8) if len(type_ok) else np.zeros(1, np.uint8)
    so = _c(s_ok, np.uint8) if len(s_ok) else np.zeros(1, np.uint8)
    p = [_c(db["p_size"], np.int32), _c(db["p_type"], np.uint16)]
    ps = [_c(db["ps_partkey"], np.int32), _c(db["ps_suppkey"], np.int32),
          _c(db["ps_supplycost"], np.int32)]
    cap = max(1, len(ps[0]))
    prow, srow = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    n = _tpch4().oracle_q2(len(p[0]), _p(p[0]), _p(p[1]), _p(to), size, len(ps[0]),
                           *[_p(x) for x in ps], len(s_ok), _p(so), _p(prow), _p(srow), cap)
    assert n <= cap
    return np.stack([prow[:n], srow[:n]], axis=1)


def q11(db: dict, nation: int, fraction: float):
    """(partkeys, values, total, groups) of run_q11 (oracle_q11), sorted by
    value descending then partkey."""
    s = [_c(db["s_suppkey"], np.int32), _c(db["s_nationkey"], np.int32)]
    ps = [_c(db["ps_partkey"], np.int32), _c(db["ps_suppkey"], np.int32),
          _c(db["ps_value"], np.int32)]
    cap = max(1, len(ps[0]))
    pk, val = np.zeros(cap, np.int32), np.zeros(cap, np.int64)
    tot, grp = C.c_int64(0), C.c_int64(0)
    n = _tpch4().oracle_q11(len(s[0]), _p(s[0]), _p(s[1]), nation, len(ps[0]),
                            *[_p(x) for x in ps], fraction, _p(pk), _p(val), cap, C.byref(tot),
                            C.byref(grp))
    return pk[:n], val[:n], tot.value, grp.value


def q15(db: dict, date_lo: int, date_hi: int):
    """[(suppkey, supplier row, revenue)] of run_q15 (oracle_q15) and the
    number of lineitems in the window."""
    li = [_c(db["l_shipdate"], np.int16), _c(db["l_suppkey"], np.int32),
          _c(db["l_disc_price"], np.int32)]
    sk = _c(db["s_suppkey"], np.int32)
    cap = max(1, len(sk))
    a, b, c = np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(cap, np.int64)
    em = C.c_int64(0)
    n = _tpch4().oracle_q15(len(li[0]), *[_p(x) for x in li], date_lo, date_hi, len(sk), _p(sk),
                            _p(a), _p(b), _p(c), cap, C.byref(em))
    assert n <= cap
    return [(int(a[i]), int(b[i]), int(c[i])) for i in range(n)], em.value


def q20(db: dict, p_ok_row, date_lo: int, date_hi: int, nation: int):
    """(qualified[max suppkey + 1], count, pairs) of run_q20 (oracle_q20)."""
    po = _c(p_ok_row, np.uint8)
    p = _c(db["p_partkey"], np.int32)
    li = [_c(db["l_shipdate"], np.int16), _c(db["l_partkey"], np.int32),
          _c(db["l_suppkey"], np.int32), _c(db["l_supp_nation"], np.int16),
          _c(db["l_quantity"], np.int16)]
    ps = [_c(db["ps_partkey"], np.int32), _c(db["ps_suppkey"], np.int32),
          _c(db["ps_availqty"], np.int32)]
    sk = db["s_suppkey"]
    smax = int(sk.max()) if len(sk) else -1
    m = np.full(smax + 1 if smax >= 0 else 0, -1, np.int16)
    for k, nk in zip(sk.tolist(), db["s_nationkey"].tolist()):
        if k >= 0:
            m[k] = nk
    m = _c(m, np.int16) if len(m) else np.zeros(1, np.int16)
    q = np.zeros(max(1, smax + 1), np.uint8)
    pairs = C.c_int64(0)
    n = _tpch4().oracle_q20(len(p), _p(p), _p(po), len(li[0]), *[_p(x) for x in li], date_lo,
                            date_hi, nation, len(ps[0]), *[_p(x) for x in ps], _p(m),
                            smax + 1 if smax >= 0 else 0, smax, _p(q), C.byref(pairs))
    return q[:smax + 1] if smax >= 0 else q[:0], n, pairs.value

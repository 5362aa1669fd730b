"""The reference's own TPC-H pipelines as a checker (oracle/_ref/libref_tpch.so).

`RefTpch(raw)` fills the reference's `Database` (bespoke_tpch/builder_impl.hpp:
11-171) from a `tests.tpch_data.raw()` draw, laid out as the reference's
build() lays it out (builder_impl.cpp:447-999): first-appearance dictionaries,
orders stably sorted by date with their comments, the orderkey -> row map
(last row wins), orders_by_customer CSR, lineitem ranges, truncated
discounted_price, returnflag_linestatus codes, packed commit_receipt,
denormalised nation keys and phone prefix codes.  `run(lines)` then answers
request lines with the reference's run_qN + write_qN_csv and returns the CSV
texts.  Columns the generator does not draw (clerk, mfgr, retail price,
available quantity, supplier account balance / address / phone, part and
partsupp comments) are zeros or empty strings; none of the pipelines run
here reads them.
"""
from __future__ import annotations

import ctypes as C
import datetime
import pathlib
import tempfile

import numpy as np

from tests import tpch_data as T

ROOT = pathlib.Path(__file__).resolve().parent.parent
LIB = ROOT / "oracle" / "_ref" / "libref_tpch.so"
# the same query unit with the GPU drop-in bodies of Q1, Q3, Q6, Q14, Q17, Q18
DROPIN = ROOT / "oracle" / "_ref" / "libdropin_tpch.so"
BASE_DAYS = (T.BASE - datetime.date(1970, 1, 1)).days


def available() -> bool:
    return LIB.exists()


def _load(path: pathlib.Path):
    lib = C.CDLL(str(path))
    vp, cp, u64 = C.c_void_p, C.c_char_p, C.c_uint64
    lib.ref_tpch_db_new.restype = vp
    lib.ref_tpch_db_new.argtypes = [C.c_int32]
    lib.ref_tpch_db_free.argtypes = [vp]
    lib.ref_tpch_set.argtypes = [vp, cp, vp, u64, u64]
    lib.ref_tpch_set_ranges.argtypes = [vp, vp, vp, u64]
    lib.ref_tpch_set_dict.argtypes = [vp, cp, vp, u64, cp, u64]
    lib.ref_tpch_set_strings.argtypes = [vp, cp, cp, vp, u64]
    lib.ref_tpch_add_nation.argtypes = [vp, C.c_int32, C.c_int32, cp]
    lib.ref_tpch_add_region.argtypes = [vp, C.c_int32, cp]
    lib.ref_tpch_finish.argtypes = [vp, C.c_int]
    lib.ref_tpch_run.argtypes = [vp, cp, cp, cp, u64]
    return lib


def _strings(values: list[str]) -> tuple[bytes, np.ndarray]:
    enc = [v.encode() for v in values]
    off = np.zeros(len(enc) + 1, np.uint32)
    np.cumsum([len(e) for e in enc], out=off[1:]) if enc else None
    return b"".join(enc), off


class RefTpch:
    def __init__(self, r: dict, lib: pathlib.Path = LIB):
        self.lib = _load(lib)
        self.db = self.lib.ref_tpch_db_new(BASE_DAYS)
        self._keep = []
        self._fill(r)

    def close(self):
        if self.db:
            self.lib.ref_tpch_db_free(self.db)
            self.db = None

    __del__ = close

    # -- setters -------------------------------------------------------------
    def _set(self, name: str, a: np.ndarray):
        a = np.ascontiguousarray(a)
        if self.lib.ref_tpch_set(self.db, name.encode(), a.ctypes.data, a.size, a.itemsize):
            raise ValueError(f"shim rejected {name} ({a.dtype})")

    def _dict(self, name: str, codes: np.ndarray, values: list[str], keep_codes=True):
        codes = np.ascontiguousarray(codes if keep_codes else np.zeros(0), np.uint16)
        blob = b"".join(v.encode() + b"\0" for v in values)
        if self.lib.ref_tpch_set_dict(self.db, name.encode(), codes.ctypes.data, codes.size,
                                      blob, len(values)):
            raise ValueError(name)

    def _str(self, name: str, values: list[str]):
        data, off = _strings(values)
        self._str_raw(name, data, off)

    def _str_raw(self, name: str, data: bytes, off: np.ndarray):
        off = np.ascontiguousarray(off, np.uint32)
        if self.lib.ref_tpch_set_strings(self.db, name.encode(), data, off.ctypes.data,
                                         len(off) - 1):
            raise ValueError(name)

    # -- the reference build() layout ------------------------------------------
    def _fill(self, r: dict):
        db = T.build(r)
        n_li, n_o = len(r["l_orderkey"]), len(r["o_orderkey"])
        # lineitem (builder_impl.cpp:447-590, supplier nation :838-853)
        for k, v in [("orderkey", db["l_orderkey"]), ("partkey", db["l_partkey"]),
                     ("suppkey", db["l_suppkey"]), ("supp_nationkey", db["l_supp_nation"]),
                     ("shipdate", db["l_shipdate"]), ("commit_receipt", db["l_commit_receipt"]),
                     ("extendedprice", db["l_extendedprice"]), ("discount", db["l_discount"]),
                     ("tax", r["l_tax"].astype(np.uint8)), ("quantity", db["l_quantity"]),
                     ("discounted_price", db["l_disc_price"]),
                     ("returnflag_linestatus", db["l_code"])]:
            self._set("lineitem." + k, v)
        # line numbers 1.. within each order's run of rows
        ok = db["l_orderkey"]
        first = np.ones(n_li, bool)
        first[1:] = ok[1:] != ok[:-1]
        start = np.maximum.accumulate(np.where(first, np.arange(n_li), 0)) if n_li else first
        self._set("lineitem.linenumber", (np.arange(n_li) - start + 1).astype(np.int32))
        # returnflag / linestatus keep only their dictionaries (codes are folded
        # into returnflag_linestatus and released, :556-568)
        self._dict("lineitem.returnflag", np.zeros(0), db["rf_dict"], keep_codes=False)
        self._dict("lineitem.linestatus", np.zeros(0), db["ls_dict"], keep_codes=False)
        self._dict("lineitem.shipinstruct", db["l_shipinstruct"], db["instruct_dict"])
        self._dict("lineitem.shipmode", db["l_shipmode"], db["shipmode_dict"])
        self._str("lineitem.comment", [""] * n_li)
        # orders, stably sorted by date (:578-608)
        perm = np.argsort(r["o_orderdate"], kind="stable")
        for k, v in [("orderkey", db["o_orderkey"]), ("custkey", db["o_custkey"]),
                     ("cust_nationkey", db["o_cust_nation"]), ("orderdate", db["o_orderdate"]),
                     ("shippriority", r["o_shippriority"][perm].astype(np.int32)),
                     ("totalprice", r["o_totalprice"][perm].astype(np.int32)),
                     ("orderkey_to_row", db["order_map"])]:
            self._set("orders." + k, v)
        self._dict("orders.orderstatus", db["o_orderstatus"], db["status_dict"])
        self._dict("orders.orderpriority", db["o_orderpriority"], db["priority_dict"])
        self._str("orders.clerk", [""] * n_o)
        self._str("orders.comment", [r["o_comment"][i] for i in perm.tolist()])
        # orders_by_customer CSR over order rows (:610-665)
        ck = db["o_custkey"].astype(np.int64)
        cmax = int(ck.max()) if n_o else -1
        offs = np.zeros(cmax + 2 if cmax >= 0 else 0, np.uint32)
        if cmax >= 0:
            np.cumsum(np.bincount(ck[ck >= 0], minlength=cmax + 1), out=offs[1:])
        rows = np.argsort(ck, kind="stable").astype(np.uint32)
        self._set("orders.orders_by_customer_offsets", offs)
        self._set("orders.orders_by_customer_rows", rows)
        s, e = db["r_start"], db["r_end"]
        self.lib.ref_tpch_set_ranges(self.db, s.ctypes.data, e.ctypes.data, n_o)
        # customer (:668-700)
        ckeys = db["c_custkey"]
        nb = np.full(int(ckeys.max()) + 1 if len(ckeys) else 0, -1, np.int16)
        nb[ckeys] = db["c_nationkey"]
        self._set("customer.custkey", ckeys)
        self._set("customer.nationkey", db["c_nationkey"])
        self._set("customer.nationkey_by_custkey", nb)
        self._set("customer.acctbal", db["c_acctbal"])
        seg, seg_dict = T.first_appearance(np.asarray(r["c_mktsegment"]), T.SEGMENTS)
        self._dict("customer.mktsegment", seg, seg_dict)
        for k in ("name", "address", "phone", "comment"):
            self._str("customer." + k, r["c_" + k])
        code = np.array([int(p[:2]) if len(p) >= 2 and p[:2].isdigit() else 255
                         for p in r["c_phone"]], np.uint8)
        self._set("customer.phone_prefix_code", code)
        # part (:702-716)
        n_p = len(db["p_partkey"])
        self._set("part.partkey", db["p_partkey"])
        self._set("part.size", db["p_size"])
        self._set("part.retailprice", np.zeros(n_p, np.int32))
        self._dict("part.mfgr", db["p_mfgr"], db["mfgr_dict"])
        self._dict("part.brand", db["p_brand"], db["brand_dict"])
        self._dict("part.container", db["p_container"], db["container_dict"])
        self._dict("part.type", db["p_type"], db["type_dict"])
        self._str("part.name", r["p_name"])
        self._str("part.comment", [""] * n_p)
        # supplier (:718-748)
        skeys = db["s_suppkey"]
        sb = np.full(int(skeys.max()) + 1 if len(skeys) else 0, -1, np.int16)
        sb[skeys] = db["s_nationkey"]
        n_s = len(skeys)
        self._set("supplier.suppkey", skeys)
        self._set("supplier.nationkey", db["s_nationkey"])
        self._set("supplier.nationkey_by_suppkey", sb)
        self._set("supplier.acctbal", db["s_acctbal"])
        self._str("supplier.name", r["s_name"])
        self._str("supplier.comment", r["s_comment"])
        self._str("supplier.address", r["s_address"])
        self._str("supplier.phone", r["s_phone"])
        # partsupp, stably sorted by part key (:750-778)
        order = np.argsort(r["ps_partkey"], kind="stable")
        n_ps = len(order)
        cost = r["ps_supplycost"][order].astype(np.int32)
        self._set("partsupp.partkey", r["ps_partkey"][order].astype(np.int32))
        self._set("partsupp.suppkey", r["ps_suppkey"][order].astype(np.int32))
        self._set("partsupp.supplycost", cost)
        avail = r["ps_availqty"][order].astype(np.int32)
        self._set("partsupp.availqty", avail)
        self._set("partsupp.supplycost_availqty", (cost.astype(np.int64) * avail).astype(np.int32))
        self._str("partsupp.comment", [""] * n_ps)
        # nation / region rows and maps (:780-822)
        for k, reg, name in zip(r["n_nationkey"].tolist(), r["n_regionkey"].tolist(),
                                r["n_name"]):
            self.lib.ref_tpch_add_nation(self.db, k, reg, name.encode())
        for k, name in zip(r["r_regionkey"].tolist(), r["r_name"]):
            self.lib.ref_tpch_add_region(self.db, k, name.encode())
        self.lib.ref_tpch_finish(self.db, int(db["l_sorted"]))

    # -- queries ---------------------------------------------------------------
    def run(self, lines: list[str]) -> list[str | None]:
        """CSV text per request line; None for an id the reference does not serve."""
        out = []
        err = C.create_string_buffer(4096)
        with tempfile.TemporaryDirectory() as d:
            for i, line in enumerate(lines):
                path = pathlib.Path(d) / f"result{i + 1}.csv"
                rc = self.lib.ref_tpch_run(self.db, line.encode(), str(path).encode(), err,
                                           len(err))
                if rc == 1:
                    raise RuntimeError(f"reference threw on {line!r}: {err.value.decode()}")
                out.append(path.read_text() if rc == 0 else None)
        return out

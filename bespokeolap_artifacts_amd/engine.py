"""Python handle on libbolap for tests and benchmarks.

Device columns are torch tensors on a ROCm device (torch is used only to
own HBM and to name the stream); every computation runs in libbolap's HIP
kernels through the C ABI.  Results come back as plain Python values.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from . import native as N

_DTYPE_CODES = {}


def _torch():
    import torch
    return torch


def dtype_code(t) -> int:
    torch = _torch()
    if not _DTYPE_CODES:
        _DTYPE_CODES.update({torch.uint8: N.U8, torch.int16: N.I16, torch.int32: N.I32,
                             torch.int64: N.I64})
    return _DTYPE_CODES[t.dtype]


def ptr(t) -> int | None:
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("libbolap needs device-resident columns")
    if not t.is_contiguous():
        raise ValueError("columns must be contiguous")
    return t.data_ptr()


def int128(lo: int, hi: int) -> int:
    """Two's-complement 128-bit value from (lo u64, hi i64)."""
    return (hi << 64) + (lo & ((1 << 64) - 1))


@dataclass
class Q1Group:
    sum_qty: int
    sum_base_price: int
    sum_disc_price: int
    sum_charge: int
    sum_disc: int
    count: int


class Engine:
    """One libbolap context on one device, issuing on torch's current stream."""

    def __init__(self, device: int = 0, stream=None):
        torch = _torch()
        self.lib = N.load()
        self.device = device
        if stream is None:
            stream = torch.cuda.current_stream(device)
        self.stream = stream
        h = C.c_void_p()
        N.check(self.lib.bolap_ctx_create(device, C.c_void_p(stream.cuda_stream), C.byref(h)))
        self.ctx = h

    def close(self) -> None:
        if self.ctx:
            self.lib.bolap_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self) -> None:
        N.check(self.lib.bolap_ctx_sync(self.ctx))

    def device_info(self) -> dict:
        info = N.DeviceInfo()
        N.check(self.lib.bolap_ctx_device_info(self.ctx, C.byref(info)))
        return {"device": info.device, "compute_units": info.compute_units,
                "lds_bytes_per_cu": info.lds_bytes_per_cu, "wavefront": info.wavefront,
                "hbm_bytes": info.hbm_bytes, "arch": info.arch.decode()}

    def tune_set(self, knob: str, value: int) -> None:
        N.check(self.lib.bolap_tune_set(self.ctx, knob.encode(), int(value)))

    def tune_get(self, knob: str) -> int:
        v = C.c_int64(0)
        N.check(self.lib.bolap_tune_get(self.ctx, knob.encode(), C.byref(v)))
        return v.value

    def set_trace(self, on: bool) -> None:
        N.check(self.lib.bolap_ctx_set_trace(self.ctx, int(bool(on))))

    def read_trace(self, max_entries: int = 4096) -> list[tuple[str, float, int]]:
        """(name, ms, count) per traced stage or count, in issue order."""
        buf = (N.TraceEntry * max_entries)()
        n = C.c_int32(0)
        N.check(self.lib.bolap_ctx_read_trace(self.ctx, buf, max_entries, C.byref(n)))
        return [(buf[i].name.decode(), buf[i].ms, buf[i].count)
                for i in range(min(n.value, max_entries))]

    # ------------------------------------------------------------------ Q1
    @staticmethod
    def q1_args(cols: dict, ship_lo: int, ship_hi: int, n_a: int, n_b: int,
                layout: int = N.Q1_LAYOUT_WIDE) -> N.Q1Args:
        a = N.Q1Args()
        a.layout = layout
        a.rows = int(cols["ship"].numel())
        a.ship = ptr(cols["ship"])
        a.ship_lo, a.ship_hi = int(ship_lo), int(ship_hi)
        a.qty = ptr(cols["qty"])
        a.price = ptr(cols["price"])
        a.disc_price = ptr(cols.get("disc_price"))
        a.disc = ptr(cols["disc"])
        a.tax = ptr(cols["tax"])
        a.flag_a = ptr(cols["flag_a"])
        a.flag_b = ptr(cols.get("flag_b"))
        a.n_a, a.n_b = int(n_a), int(n_b)
        return a

    def q1_scan_agg(self, cols: dict, ship_lo: int, ship_hi: int, n_a: int, n_b: int,
                    layout: int = N.Q1_LAYOUT_WIDE) -> list[Q1Group]:
        a = self.q1_args(cols, ship_lo, ship_hi, n_a, n_b, layout)
        out = (N.Q1Group * (n_a * n_b))()
        N.check(self.lib.bolap_q1_scan_agg(self.ctx, C.byref(a), out))
        return [Q1Group(g.sum_qty, g.sum_base_price, g.sum_disc_price,
                        int128(g.sum_charge_lo, g.sum_charge_hi), g.sum_disc, g.count)
                for g in out]

    def q1_scan_agg_async(self, a: N.Q1Args, out_dev) -> None:
        N.check(self.lib.bolap_q1_scan_agg_async(self.ctx, C.byref(a),
                                                 C.c_void_p(out_dev.data_ptr())))

    # ------------------------------------------------------------ join + sum
    @staticmethod
    def pred(col, op: int = N.OP_RANGE, lo: int = 0, hi: int = 0, values=None,
             set_bitmap=None) -> N.Pred:
        """Predicate on a device column: lo <= v <= hi, or v in a set given as
        inline values (0..255) or as a device uint32 bitmap."""
        p = N.Pred()
        p.col = ptr(col)
        p.dtype = dtype_code(col)
        p.op = op
        p.lo, p.hi = int(lo), int(hi)
        if values is not None:
            words = [0, 0, 0, 0]
            for v in values:
                words[v >> 6] |= 1 << (v & 63)
            for i in range(4):
                p.inline_set[i] = words[i]
        if set_bitmap is not None:
            p.set = ptr(set_bitmap)
            p.set_bits = int(set_bitmap.numel()) * 32
        return p

    def join_sum_args(self, dim_key, preds: list, fact_fk, fact_price, fact_disc,
                      capacity_hint: int = 0, placement: int = N.PLACE_AUTO) -> N.JoinSumArgs:
        a = N.JoinSumArgs()
        a.dim_rows = int(dim_key.numel())
        a.dim_key = ptr(dim_key)
        a.n_dim_preds = len(preds)
        for i, p in enumerate(preds):
            a.dim_preds[i] = p
        a.fact_rows = int(fact_fk.numel())
        a.fact_fk = ptr(fact_fk)
        a.fact_price = ptr(fact_price)
        a.price_dtype = dtype_code(fact_price)
        a.fact_disc = ptr(fact_disc)
        a.disc_dtype = dtype_code(fact_disc)
        a.capacity_hint = int(capacity_hint)
        a.placement = int(placement)
        return a

    def join_sum(self, *args, **kw) -> dict:
        a = self.join_sum_args(*args, **kw)
        return self.join_sum_run(a)

    def join_sum_run(self, a: N.JoinSumArgs) -> dict:
        r = N.JoinSumResult()
        N.check(self.lib.bolap_join_sum(self.ctx, C.byref(a), C.byref(r)))
        return {"revenue": r.revenue, "build_rows": r.build_rows, "matches": r.matches,
                "capacity": r.capacity, "placement": r.placement, "fallbacks": r.fallbacks}

    # ------------------------------------------------- chain join + group-by
    def chain_args(self, t: dict, segment: int, day_lt: int, capacity_hint: int = 0,
                   disc_price=None) -> N.ChainArgs:
        a = N.ChainArgs()
        a.cust_rows = int(t["cust_key"].numel())
        a.cust_key, a.cust_segment = ptr(t["cust_key"]), ptr(t["cust_segment"])
        a.segment = int(segment)
        a.order_rows = int(t["order_key"].numel())
        a.order_key, a.order_day = ptr(t["order_key"]), ptr(t["order_day"])
        a.order_cust = ptr(t["order_cust"])
        a.day_lt = int(day_lt)
        a.fact_rows = int(t["fact_order"].numel())
        a.fact_order = ptr(t["fact_order"])
        a.fact_price = ptr(t.get("fact_price"))
        a.price_dtype = dtype_code(t["fact_price"]) if t.get("fact_price") is not None else N.I64
        a.fact_disc = ptr(t.get("fact_disc"))
        a.disc_dtype = dtype_code(t["fact_disc"]) if t.get("fact_disc") is not None else N.I64
        a.fact_disc_price = ptr(disc_price)
        a.capacity_hint = int(capacity_hint)
        return a

    @staticmethod
    def _chain_stats(st: N.ChainStats) -> dict:
        return {"cust_pass": st.cust_pass, "order_pass": st.order_pass,
                "fact_matches": st.fact_matches, "groups": st.groups, "capacity": st.capacity,
                "fallbacks": st.fallbacks, "sort_passes": st.sort_passes}

    def chain_topk(self, a: N.ChainArgs, k: int):
        out = (N.GroupRow * max(1, k))()
        n = C.c_int32(0)
        st = N.ChainStats()
        N.check(self.lib.bolap_chain_topk(self.ctx, C.byref(a), k, out, C.byref(n),
                                          C.byref(st)))
        return [(out[i].key, out[i].day, out[i].revenue) for i in range(n.value)], \
            self._chain_stats(st)

    def chain_groups_sorted(self, a: N.ChainArgs, max_rows: int):
        out = (N.GroupRow * max(1, max_rows))()
        n = C.c_int64(0)
        st = N.ChainStats()
        N.check(self.lib.bolap_chain_groups_sorted(self.ctx, C.byref(a), out, max_rows,
                                                   C.byref(n), C.byref(st)))
        import numpy as np
        arr = np.frombuffer(out, dtype=np.dtype([("key", "<i4"), ("day", "<i4"),
                                                 ("revenue", "<i8")]), count=n.value).copy()
        return arr, self._chain_stats(st)

    # ------------------------------------------------------------ radix sort
    def radix_sort(self, keys, vals, key_bits: int = 64) -> None:
        """Sorts device uint64 keys (int64 tensor) with int32 payloads in place."""
        N.check(self.lib.bolap_radix_sort_u64(self.ctx, ptr(keys), ptr(vals), keys.numel(),
                                              key_bits))

    def topk(self, keys, vals, k: int):
        torch = _torch()
        ko = torch.empty(max(1, k), dtype=torch.int64, device=keys.device)
        vo = torch.empty(max(1, k), dtype=torch.int32, device=keys.device)
        N.check(self.lib.bolap_topk_u64(self.ctx, ptr(keys), ptr(vals), keys.numel(), k,
                                        ptr(ko), ptr(vo)))
        m = min(k, keys.numel())
        return ko[:m], vo[:m]

    # ------------------------------------------------- star join + group-by
    def star_groupby(self, a: N.StarArgs, want_rows: bool = False, max_rows: int = 0):
        """Runs bolap_star_groupby; returns (stats dict, keys, revenues) with
        device tensors of the groups in key order when want_rows."""
        torch = _torch()
        dev = f"cuda:{self.device}"
        keys = rev = None
        if want_rows:
            cap = max_rows or (int(a.group_max) - int(a.group_min) + 1)
            keys = torch.empty(max(1, cap), dtype=torch.int32, device=dev)
            rev = torch.empty(max(1, cap), dtype=torch.int64, device=dev)
        n = C.c_int64(0)
        st = N.StarStats()
        N.check(self.lib.bolap_star_groupby(self.ctx, C.byref(a), ptr(keys), ptr(rev),
                                            keys.numel() if want_rows else 0, C.byref(n),
                                            C.byref(st)))
        k = int(a.n_dims)
        stats = {"dim_pass": list(st.dim_pass)[:k], "dim_placement": list(st.dim_placement)[:k],
                 "group_placement": st.group_placement, "matches": st.matches,
                 "groups": st.groups, "group_capacity": st.group_capacity,
                 "revenue": st.revenue, "fallbacks": st.fallbacks,
                 "count_passes": st.count_passes}
        if want_rows:
            return stats, keys[:n.value], rev[:n.value]
        return stats, None, None

    def hll(self, col):
        """HyperLogLog registers (4096, uint8 numpy) of a device int32 column."""
        import numpy as np
        regs = np.zeros(N.HLL_REGS, np.uint8)
        N.check(self.lib.bolap_hll_i32(self.ctx, ptr(col), int(col.numel()),
                                       regs.ctypes.data_as(C.POINTER(C.c_uint8))))
        return regs

    def sample_counts(self, args: list, seed: int):
        """Sample counts (n_plans x 16 int64 numpy) of a batch of star plans."""
        import numpy as np
        n = len(args)
        out = np.zeros((max(1, n), N.SAMPLE_COUNTS), np.int64)
        if n == 0:
            return out[:0]
        arr = (N.StarArgs * n)(*args)
        N.check(self.lib.bolap_star_sample_counts(self.ctx, arr, n, seed,
                                                  out.ctypes.data_as(C.POINTER(C.c_int64))))
        return out

    # ------------------------------------------------------- generators
    def gen_lineitem(self, seed: int, rows: int) -> dict:
        torch = _torch()
        dev = f"cuda:{self.device}"
        cols = {
            "ship": torch.empty(rows, dtype=torch.int32, device=dev),
            "qty": torch.empty(rows, dtype=torch.int64, device=dev),
            "price": torch.empty(rows, dtype=torch.int64, device=dev),
            "disc": torch.empty(rows, dtype=torch.int64, device=dev),
            "tax": torch.empty(rows, dtype=torch.int64, device=dev),
            "flag_a": torch.empty(rows, dtype=torch.uint8, device=dev),
            "flag_b": torch.empty(rows, dtype=torch.uint8, device=dev),
        }
        g = N.GenLineitemCols(*(ptr(cols[k]) for k in
                                ("ship", "qty", "price", "disc", "tax", "flag_a", "flag_b")))
        N.check(self.lib.bolap_gen_lineitem(self.ctx, seed, rows, C.byref(g)))
        return cols

    def gen_dim(self, seed: int, rows: int) -> dict:
        torch = _torch()
        dev = f"cuda:{self.device}"
        cols = {"key": torch.empty(rows, dtype=torch.int32, device=dev),
                "cat": torch.empty(rows, dtype=torch.uint8, device=dev),
                "cont": torch.empty(rows, dtype=torch.uint8, device=dev)}
        N.check(self.lib.bolap_gen_dim(self.ctx, seed, rows, ptr(cols["key"]),
                                       ptr(cols["cat"]), ptr(cols["cont"])))
        return cols

    def gen_fact(self, seed: int, rows: int, dim_rows: int) -> dict:
        torch = _torch()
        dev = f"cuda:{self.device}"
        cols = {"fk": torch.empty(rows, dtype=torch.int32, device=dev),
                "price": torch.empty(rows, dtype=torch.int64, device=dev),
                "disc": torch.empty(rows, dtype=torch.int64, device=dev)}
        N.check(self.lib.bolap_gen_fact(self.ctx, seed, rows, dim_rows, ptr(cols["fk"]),
                                        ptr(cols["price"]), ptr(cols["disc"])))
        return cols

    def gen_chain(self, seed: int, cust_rows: int, order_rows: int) -> dict:
        """Customers, orders and their clustered fact rows (1..7 per order)."""
        torch = _torch()
        dev = f"cuda:{self.device}"
        t = {"cust_key": torch.empty(cust_rows, dtype=torch.int32, device=dev),
             "cust_segment": torch.empty(cust_rows, dtype=torch.uint8, device=dev),
             "order_key": torch.empty(order_rows, dtype=torch.int32, device=dev),
             "order_day": torch.empty(order_rows, dtype=torch.int32, device=dev),
             "order_cust": torch.empty(order_rows, dtype=torch.int32, device=dev),
             "fact_offsets": torch.empty(order_rows + 1, dtype=torch.int64, device=dev)}
        N.check(self.lib.bolap_gen_customers(self.ctx, seed, cust_rows, ptr(t["cust_key"]),
                                             ptr(t["cust_segment"])))
        N.check(self.lib.bolap_gen_orders(self.ctx, seed, order_rows, cust_rows,
                                          ptr(t["order_key"]), ptr(t["order_day"]),
                                          ptr(t["order_cust"]), ptr(t["fact_offsets"])))
        fact_rows = int(t["fact_offsets"][-1].item())
        t["fact_order"] = torch.empty(fact_rows, dtype=torch.int32, device=dev)
        t["fact_price"] = torch.empty(fact_rows, dtype=torch.int64, device=dev)
        t["fact_disc"] = torch.empty(fact_rows, dtype=torch.int64, device=dev)
        N.check(self.lib.bolap_gen_order_facts(self.ctx, seed, order_rows,
                                               ptr(t["fact_offsets"]), fact_rows,
                                               ptr(t["fact_order"]), ptr(t["fact_price"]),
                                               ptr(t["fact_disc"])))
        return t

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

    def tune_save(self, path: str) -> None:
        N.check(self.lib.bolap_tune_save(self.ctx, str(path).encode()))

    def tune_load(self, path: str) -> None:
        N.check(self.lib.bolap_tune_load(self.ctx, str(path).encode()))

    def join_autotune(self, a: N.JoinSumArgs, reps: int = 3) -> float:
        """Keeps the fastest probe configuration for these arguments; returns
        its median call time in ms."""
        ms = C.c_double(0.0)
        N.check(self.lib.bolap_join_autotune(self.ctx, C.byref(a), int(reps), C.byref(ms)))
        return ms.value

    def sort_autotune(self, keys, vals, key_bits: int, reps: int = 3) -> float:
        """Keeps the fastest radix digit width for sorting these pairs (left
        unchanged); returns its median sort time in ms."""
        ms = C.c_double(0.0)
        N.check(self.lib.bolap_sort_autotune(self.ctx, ptr(keys), ptr(vals), keys.numel(),
                                             int(key_bits), int(reps), C.byref(ms)))
        return ms.value

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

    @staticmethod
    def q1_finalize(groups: list, rf_dict: list[str], ls_dict: list[str],
                    layout: int = N.Q1_LAYOUT_WIDE) -> list[dict]:
        """run_q1's result rows from the exact group accumulators (host C++);
        groups are indexed rf * len(ls_dict) + ls."""
        n_rf, n_ls = len(rf_dict), len(ls_dict)
        gs = (N.Q1Group * (n_rf * n_ls))()
        for i, g in enumerate(groups):
            charge = g.sum_charge & ((1 << 128) - 1)
            lo, hi = charge & ((1 << 64) - 1), charge >> 64
            gs[i] = N.Q1Group(g.sum_qty, g.sum_base_price, g.sum_disc_price, lo,
                              hi - (1 << 64) if hi >= (1 << 63) else hi, g.sum_disc, g.count, 0)
        rf = (C.c_char_p * n_rf)(*(x.encode() for x in rf_dict))
        ls = (C.c_char_p * n_ls)(*(x.encode() for x in ls_dict))
        rows = (N.Q1Row * (n_rf * n_ls))()
        n = C.c_int32()
        N.check(N.load().bolap_q1_finalize(gs, layout, n_rf, n_ls, rf, ls, rows, C.byref(n)))
        out = []
        for r in rows[:n.value]:
            d = {k: getattr(r, k) for k, _ in N.Q1Row._fields_}
            d["returnflag"], d["linestatus"] = d["returnflag"].decode(), d["linestatus"].decode()
            out.append(d)
        return out

    @staticmethod
    def q1_write_csv(path: str, rows: list[dict]) -> None:
        arr = (N.Q1Row * max(1, len(rows)))()
        for i, r in enumerate(rows):
            arr[i] = N.Q1Row(r["returnflag"].encode(), r["linestatus"].encode(), r["sum_qty"],
                             r["sum_base_price"], r["sum_disc_price"], r["sum_charge"],
                             r["avg_qty"], r["avg_price"], r["avg_disc"], r["count_order"])
        N.check(N.load().bolap_q1_write_csv(str(path).encode(), arr, len(rows)))

    @staticmethod
    def q6_write_csv(path: str, result: dict) -> None:
        r = N.Q6Result(result["revenue"], result["rows_emitted"], result["rows_scanned"])
        N.check(N.load().bolap_q6_write_csv(str(path).encode(), C.byref(r)))

    @staticmethod
    def q18_write_csv(path: str, rows, c_name: list[str]) -> None:
        """rows: the structured array returned by q18_run."""
        n = len(rows)
        arr = (N.Q18Row * max(1, n))()
        for i, r in enumerate(rows):
            arr[i] = N.Q18Row(*(int(r[k]) for k, _ in N.Q18Row._fields_))
        names = (C.c_char_p * max(1, len(c_name)))(*(x.encode() for x in c_name))
        N.check(N.load().bolap_q18_write_csv(str(path).encode(), arr, n, names, len(c_name)))

    @staticmethod
    def q3_write_csv(path: str, rows, shippriority=None) -> None:
        """rows: (key, day, revenue) tuples or a GROUP_DTYPE array, in output order."""
        n = len(rows)
        arr = (N.GroupRow * max(1, n))()
        for i, r in enumerate(rows):
            arr[i] = N.GroupRow(int(r[0]), int(r[1]), int(r[2]))
        sp = None
        if shippriority is not None:
            sp = (C.c_int32 * max(1, n))(*(int(x) for x in shippriority))
        N.check(N.load().bolap_q3_write_csv(str(path).encode(), arr, n, sp))

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

    def join_sum_run(self, a: N.JoinSumArgs, noalloc: bool = False) -> dict:
        """noalloc: the allocation-free entry point (bolap_join_sum_noalloc)."""
        r = N.JoinSumResult()
        fn = self.lib.bolap_join_sum_noalloc if noalloc else self.lib.bolap_join_sum
        N.check(fn(self.ctx, C.byref(a), C.byref(r)))
        return {"revenue": r.revenue, "build_rows": r.build_rows, "matches": r.matches,
                "capacity": r.capacity, "placement": r.placement, "fallbacks": r.fallbacks}

    # ------------------------------------------------- chain join + group-by
    def chain_args(self, t: dict, segment: int, day_lt: int, capacity_hint: int = 0,
                   disc_price=None, ranges=None, key_ranges=None, fact_ship=None,
                   ship_gt: int = 0, fact_order_sorted: bool = False) -> N.ChainArgs:
        """`ranges` = (start, end) device uint32/int32 arrays per order row
        (order_ranges): the fact side is then read by ranges;
        `fact_order_sorted` says the fact rows are sorted by order key (else
        each row's key is checked against its order's).  `key_ranges` =
        (cust_min, cust_max, order_min, order_max) from the tables' statistics:
        the call then skips its own min/max pass."""
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
        if ranges is not None:
            a.order_fact_start, a.order_fact_end = ptr(ranges[0]), ptr(ranges[1])
            a.fact_order_sorted = int(bool(fact_order_sorted))
        if fact_ship is not None:  # sum only fact rows with fact_ship > ship_gt
            a.fact_ship, a.ship_gt = ptr(fact_ship), int(ship_gt)
        if key_ranges is not None:
            a.key_ranges_known = 1
            (a.cust_key_min, a.cust_key_max, a.order_key_min,
             a.order_key_max) = (int(v) for v in key_ranges)
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

    def group_minmax(self, pair_group, pair_item, item_offset, item_count, item_has,
                     item_value, range_value, groups: int, keys: int, dense_max: int = 0,
                     max_rows: int | None = None):
        """Runs bolap_group_minmax over device int32 columns (item_has: device
        uint8 or None); returns (stats dict, rows as a device int32 tensor of
        shape (n, 5): group, key, has, min, max in (group, key) order)."""
        torch = _torch()
        a = N.MinmaxArgs(int(pair_group.numel()), ptr(pair_group), ptr(pair_item),
                         int(item_offset.numel()), ptr(item_offset), ptr(item_count),
                         ptr(item_has), ptr(item_value), int(range_value.numel()),
                         ptr(range_value), int(groups), int(keys), int(dense_max))
        st = N.MinmaxStats()
        N.check(self.lib.bolap_group_minmax(self.ctx, C.byref(a), None, 0, None, C.byref(st)))
        cap = st.groups if max_rows is None else max_rows
        out = torch.empty((max(1, cap), 5), dtype=torch.int32, device=f"cuda:{self.device}")
        n = C.c_int64(0)
        N.check(self.lib.bolap_group_minmax(self.ctx, C.byref(a), ptr(out), cap, C.byref(n),
                                            C.byref(st)))
        stats = {"expanded_rows": st.expanded_rows, "groups": st.groups,
                 "placement": st.placement, "slots": st.slots}
        return stats, out[:n.value]

    def sort_order(self, primary, secondary=None):
        """Device uint32 row order sorting by (primary, secondary) ascending."""
        torch = _torch()
        order = torch.empty(max(1, primary.numel()), dtype=torch.int32,
                            device=f"cuda:{self.device}")
        N.check(self.lib.bolap_sort_order_i32(self.ctx, ptr(primary), ptr(secondary),
                                              int(primary.numel()), ptr(order)))
        return order[:primary.numel()]

    def gather_rows(self, src, order):
        """src[order] on the device through bolap_gather_rows."""
        torch = _torch()
        dst = torch.empty((max(1, order.numel()),) + tuple(src.shape[1:]), dtype=src.dtype,
                          device=src.device)
        N.check(self.lib.bolap_gather_rows(self.ctx, ptr(src), int(src.shape[0]),
                                           int(src.element_size() * (src[0].numel() if
                                                                     src.dim() > 1 else 1)),
                                           ptr(order), int(order.numel()), ptr(dst)))
        return dst[:order.numel()]

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

    # --------------------------------------------------------------- Q6
    @staticmethod
    def q6_params(date: str, discount: str, quantity: str) -> dict:
        """run_q6's arguments from the query strings (host-side C++)."""
        p = N.Q6Params()
        lib = N.load()
        N.check(lib.bolap_q6_parse(date.encode(), discount.encode(), quantity.encode(),
                                   C.byref(p)))
        return {k: getattr(p, k) for k, _ in N.Q6Params._fields_}

    @staticmethod
    def q6_args(cols: dict, p: dict, zones=None) -> N.Q6Args:
        a = N.Q6Args()
        a.rows = int(cols["ship"].numel())
        a.ship, a.disc, a.qty, a.price = (ptr(cols[k]) for k in ("ship", "disc", "qty", "price"))
        a.ship_lo, a.ship_hi = p["ship_lo"], p["ship_hi"]
        a.disc_lo, a.disc_hi = p["disc_lo"], p["disc_hi"]
        a.qty_lt = p["qty_lt"]
        a.zones = ptr(zones) if zones is not None else None
        return a

    def q6_run(self, a: N.Q6Args) -> dict:
        r = N.Q6Result()
        N.check(self.lib.bolap_q6(self.ctx, C.byref(a), C.byref(r)))
        return {"revenue": r.revenue, "rows_emitted": r.rows_emitted,
                "rows_scanned": r.rows_scanned}

    def q6(self, cols: dict, p: dict, zones=None) -> dict:
        return self.q6_run(self.q6_args(cols, p, zones))

    def q6_zones(self, cols: dict):
        """Zone map (uint8 tensor of 16-byte zones) of the Q6 columns."""
        torch = _torch()
        rows = int(cols["ship"].numel())
        nz = (rows + N.ZONE_ROWS - 1) // N.ZONE_ROWS
        z = torch.empty(max(1, nz) * 16, dtype=torch.uint8, device=f"cuda:{self.device}")
        N.check(self.lib.bolap_q6_zones(self.ctx, rows, ptr(cols["ship"]), ptr(cols["disc"]),
                                        ptr(cols["qty"]), ptr(z)))
        return z

    def q6_cluster(self, cols: dict, with_perm: bool = False) -> dict:
        """Copies of the Q6 columns ordered by the shard key."""
        torch = _torch()
        rows = int(cols["ship"].numel())
        out = {k: torch.empty_like(cols[k]) for k in ("ship", "disc", "qty", "price")}
        perm = torch.empty(max(1, rows), dtype=torch.int32, device=f"cuda:{self.device}") \
            if with_perm else None
        N.check(self.lib.bolap_q6_cluster(self.ctx, rows, *(ptr(cols[k]) for k in
                                                            ("ship", "disc", "qty", "price")),
                                          *(ptr(out[k]) for k in ("ship", "disc", "qty", "price")),
                                          ptr(perm) if perm is not None else None))
        if perm is not None:
            out["perm"] = perm[:rows]
        return out

    # ---------------------------------------------------------- Q14 / Q17
    @staticmethod
    def q14_window(date: str) -> tuple[int, int]:
        lo, hi = C.c_int32(), C.c_int32()
        N.check(N.load().bolap_q14_parse(date.encode(), C.byref(lo), C.byref(hi)))
        return lo.value, hi.value

    def q14(self, t: dict, ship_lo: int, ship_hi: int, zones=None) -> dict:
        return self.q14_run(self.q14_args(t, ship_lo, ship_hi, zones))

    def q14_args(self, t: dict, ship_lo: int, ship_hi: int, zones=None) -> N.Q14Args:
        a = N.Q14Args()
        a.li_rows = int(t["l_ship"].numel())
        a.l_ship, a.l_partkey = ptr(t["l_ship"]), ptr(t["l_partkey"])
        a.l_disc_price = ptr(t["l_disc_price"])
        a.ship_lo, a.ship_hi = ship_lo, ship_hi
        a.zones = ptr(zones) if zones is not None else None
        a.part_rows = int(t["p_partkey"].numel())
        a.p_partkey, a.p_promo = ptr(t["p_partkey"]), ptr(t["p_promo"])
        return a

    def q14_run(self, a: N.Q14Args) -> dict:
        r = N.Q14Result()
        N.check(self.lib.bolap_q14(self.ctx, C.byref(a), C.byref(r)))
        return {"promo_sum": r.promo_sum, "total_sum": r.total_sum,
                "rows_emitted": r.rows_emitted, "promo_revenue": r.promo_revenue}

    # ------------------------------------- Q4, Q5, Q10, Q12, Q19 (tpch_orders)
    @staticmethod
    def order_side(t: dict) -> N.OrderSide:
        """t: reference-layout device columns o_orderdate (sorted), o_orderkey,
        r_start, r_end, l_orderkey and the flag l_sorted."""
        o = N.OrderSide()
        o.order_rows = int(t["o_orderdate"].numel())
        o.o_orderdate, o.o_orderkey = ptr(t["o_orderdate"]), ptr(t["o_orderkey"])
        o.range_start, o.range_end = ptr(t["r_start"]), ptr(t["r_end"])
        o.l_orderkey, o.l_orderkey_sorted = ptr(t["l_orderkey"]), int(t["l_sorted"])
        return o

    def q4(self, t: dict, n_priority: int, date_lo: int, date_hi: int) -> list:
        a = N.Q4Args()
        a.orders = self.order_side(t)
        a.o_orderpriority, a.n_priority = ptr(t["o_orderpriority"]), int(n_priority)
        a.l_commit_receipt = ptr(t["l_commit_receipt"])
        a.date_lo, a.date_hi = int(date_lo), int(date_hi)
        out = (C.c_int64 * max(1, n_priority))()
        N.check(self.lib.bolap_q4(self.ctx, C.byref(a), out))
        return list(out)[:n_priority]

    def q5(self, t: dict, nation_in_region, date_lo: int, date_hi: int) -> list:
        import numpy as np
        nr = np.ascontiguousarray(nation_in_region, np.uint8)
        a = N.Q5Args()
        a.orders = self.order_side(t)
        a.o_cust_nation, a.l_supp_nation = ptr(t["o_cust_nation"]), ptr(t["l_supp_nation"])
        a.l_disc_price = ptr(t["l_disc_price"])
        a.nation_in_region, a.n_nation = nr.ctypes.data, len(nr)
        a.date_lo, a.date_hi = int(date_lo), int(date_hi)
        out = (C.c_int64 * max(1, len(nr)))()
        N.check(self.lib.bolap_q5(self.ctx, C.byref(a), out))
        return list(out)[:len(nr)]

    def q10(self, t: dict, code_lo: int, code_span: int, date_lo: int, date_hi: int,
            n_nation_names: int, max_rows: int | None = None, raw: bool = False):
        """Rows (custkey, cust_row, revenue) in the reference's order; raw=True
        returns them as a numpy record array (a view of a reused buffer)."""
        a = N.Q10Args()
        a.orders = self.order_side(t)
        a.o_custkey, a.l_code = ptr(t["o_custkey"]), ptr(t["l_code"])
        a.code_lo, a.code_span = int(code_lo), int(code_span)
        a.l_disc_price = ptr(t["l_disc_price"])
        a.date_lo, a.date_hi = int(date_lo), int(date_hi)
        a.cust_rows = int(t["c_custkey"].numel())
        a.c_custkey, a.c_nationkey = ptr(t["c_custkey"]), ptr(t["c_nationkey"])
        a.n_nation_names = int(n_nation_names)
        import numpy as np
        cap = max(1, a.cust_rows if max_rows is None else max_rows)
        if getattr(self, "_q10_rows_cap", 0) < cap:   # reused host buffer
            self._q10_rows = (N.Q10Row * cap)()
            self._q10_rows_cap = cap
        rows = self._q10_rows
        n = C.c_int64(0)
        N.check(self.lib.bolap_q10(self.ctx, C.byref(a), rows, cap, C.byref(n)))
        rec = np.frombuffer(rows, dtype=np.dtype([("custkey", "<i4"), ("cust_row", "<i4"),
                                                  ("revenue", "<i8")]), count=n.value)
        if raw:
            return rec
        return list(zip(rec["custkey"].tolist(), rec["cust_row"].tolist(),
                        rec["revenue"].tolist()))

    def q12(self, t: dict, mode1: int, mode2: int, date_lo: int, date_hi: int, high1: int,
            high2: int):
        a = N.Q12Args()
        a.li_rows = int(t["l_shipmode"].numel())
        a.l_shipmode, a.l_commit_receipt = ptr(t["l_shipmode"]), ptr(t["l_commit_receipt"])
        a.l_shipdate, a.l_orderkey = ptr(t["l_shipdate"]), ptr(t["l_orderkey"])
        a.mode1, a.mode2, a.date_lo, a.date_hi = int(mode1), int(mode2), int(date_lo), int(date_hi)
        a.order_map, a.map_size = ptr(t["order_map"]), int(t["order_map"].numel())
        a.o_orderpriority, a.high1, a.high2 = ptr(t["o_orderpriority"]), int(high1), int(high2)
        r = N.Q12Result()
        N.check(self.lib.bolap_q12(self.ctx, C.byref(a), C.byref(r)))
        return (r.high[0], r.low[0]), (r.high[1], r.low[1])

    def q19(self, t: dict, groups, mode1: int, mode2: int, instruct: int) -> int:
        """groups: 3 tuples (enabled, brand, [4 container codes], size_hi, qty_lo)."""
        a = N.Q19Args()
        a.part_rows = int(t["p_partkey"].numel())
        a.p_partkey, a.p_size = ptr(t["p_partkey"]), ptr(t["p_size"])
        a.p_brand, a.p_brand_dtype = ptr(t["p_brand"]), dtype_code(t["p_brand"])
        a.p_container, a.p_container_dtype = ptr(t["p_container"]), dtype_code(t["p_container"])
        a.li_rows = int(t["l_partkey"].numel())
        a.l_partkey, a.l_quantity = ptr(t["l_partkey"]), ptr(t["l_quantity"])
        a.l_price, a.l_disc = ptr(t["l_extendedprice"]), ptr(t["l_discount"])
        a.l_shipmode, a.l_shipinstruct = ptr(t["l_shipmode"]), ptr(t["l_shipinstruct"])
        for g, (en, br, cont, sh, ql) in enumerate(groups):
            a.groups[g].enabled, a.groups[g].brand = int(en), int(br)
            a.groups[g].size_hi, a.groups[g].qty_lo = int(sh), int(ql)
            for c in range(4):
                a.groups[g].containers[c] = int(cont[c])
        a.mode1, a.mode2, a.shipinstruct = int(mode1), int(mode2), int(instruct)
        r = C.c_int64(0)
        N.check(self.lib.bolap_q19(self.ctx, C.byref(a), C.byref(r)))
        return r.value

    def q21(self, t: dict, final_code: int, nation: int, n_suppkeys: int) -> list:
        a = N.Q21Args()
        a.order_rows = int(t["o_orderstatus"].numel())
        a.o_orderstatus, a.final_code = ptr(t["o_orderstatus"]), int(final_code)
        a.o_orderkey, a.range_start, a.range_end = (ptr(t["o_orderkey"]), ptr(t["r_start"]),
                                                    ptr(t["r_end"]))
        a.l_orderkey, a.l_orderkey_sorted = ptr(t["l_orderkey"]), int(t["l_sorted"])
        a.l_suppkey, a.l_commit_receipt = ptr(t["l_suppkey"]), ptr(t["l_commit_receipt"])
        a.supp_rows = int(t["s_suppkey"].numel())
        a.s_suppkey, a.s_nationkey, a.nation = ptr(t["s_suppkey"]), ptr(t["s_nationkey"]), nation
        import numpy as np
        out = np.zeros(max(1, n_suppkeys), np.int64)
        N.check(self.lib.bolap_q21(self.ctx, C.byref(a), out.ctypes.data_as(C.POINTER(C.c_int64)),
                                   int(n_suppkeys)))
        return out[:n_suppkeys].tolist()

    def q22(self, t: dict, numeric_codes, extra_pairs):
        a = N.Q22Args()
        a.cust_rows = int(t["c_custkey"].numel())
        a.c_custkey, a.c_phone_code = ptr(t["c_custkey"]), ptr(t["c_phone_code"])
        a.c_phone_pair, a.c_acctbal = ptr(t["c_phone_pair"]), ptr(t["c_acctbal"])
        for c in numeric_codes:
            a.numeric_allowed[c] = 1
        a.n_extra = len(extra_pairs)
        for i, p in enumerate(extra_pairs):
            a.extra_pairs[i] = p
        a.order_rows = int(t["o_custkey"].numel())
        a.o_custkey = ptr(t["o_custkey"])
        r = N.Q22Result()
        N.check(self.lib.bolap_q22(self.ctx, C.byref(a), C.byref(r)))
        num = [[r.numcust[c], r.totacctbal[c]] for c in range(100)]
        ext = [[r.extra_numcust[i], r.extra_totacctbal[i]] for i in range(len(extra_pairs))]
        return num, ext

    def q7(self, t: dict, n1: int, n2: int, ship_lo: int, ship_hi: int, split: int) -> list:
        a = N.Q7Args()
        a.orders = self.order_side(t)
        a.o_cust_nation, a.l_shipdate = ptr(t["o_cust_nation"]), ptr(t["l_shipdate"])
        a.l_suppkey, a.l_disc_price = ptr(t["l_suppkey"]), ptr(t["l_disc_price"])
        a.supp_rows = int(t["s_suppkey"].numel())
        a.s_suppkey, a.s_nationkey = ptr(t["s_suppkey"]), ptr(t["s_nationkey"])
        a.n_suppkeys = int(t["s_suppkey"].max().item()) + 1 if a.supp_rows else 0
        a.nation1, a.nation2, a.ship_lo, a.ship_hi, a.split = n1, n2, ship_lo, ship_hi, split
        out = (C.c_int64 * 8)()
        N.check(self.lib.bolap_q7(self.ctx, C.byref(a), out))
        return list(out)

    def q8(self, t: dict, nation_in_region, target: int, type_code: int, date_lo: int,
           date_hi: int, split: int):
        import numpy as np
        nr = np.ascontiguousarray(nation_in_region, np.uint8)
        a = N.Q8Args()
        a.orders = self.order_side(t)
        a.o_custkey, a.l_partkey, a.l_suppkey = (ptr(t["o_custkey"]), ptr(t["l_partkey"]),
                                                 ptr(t["l_suppkey"]))
        a.l_price, a.l_disc = ptr(t["l_extendedprice"]), ptr(t["l_discount"])
        a.cust_rows = int(t["c_custkey"].numel())
        a.c_custkey, a.c_nationkey = ptr(t["c_custkey"]), ptr(t["c_nationkey"])
        a.n_cust_keys = int(t["c_custkey"].max().item()) + 1
        a.nation_in_region, a.n_nation = nr.ctypes.data, len(nr)
        a.supp_rows = int(t["s_suppkey"].numel())
        a.s_suppkey, a.s_nationkey = ptr(t["s_suppkey"]), ptr(t["s_nationkey"])
        a.n_supp_keys = int(t["s_suppkey"].max().item()) + 1
        a.target_nation = target
        a.part_rows = int(t["p_partkey"].numel())
        a.p_partkey, a.p_type = ptr(t["p_partkey"]), ptr(t["p_type"])
        a.n_part_keys = int(t["p_partkey"].max().item()) + 1
        a.type_code, a.date_lo, a.date_hi, a.split = type_code, date_lo, date_hi, split
        tot, nat = (C.c_int64 * 2)(), (C.c_int64 * 2)()
        N.check(self.lib.bolap_q8(self.ctx, C.byref(a), tot, nat))
        return list(tot), list(nat)

    def q9(self, t: dict, p_ok, n_nation: int):
        import numpy as np
        a = N.Q9Args()
        a.part_rows = int(t["p_partkey"].numel())
        a.p_partkey, a.p_ok = ptr(t["p_partkey"]), ptr(p_ok)
        a.n_part_keys = int(t["p_partkey"].max().item()) + 1
        a.ps_rows = int(t["ps_partkey"].numel())
        a.ps_partkey, a.ps_suppkey = ptr(t["ps_partkey"]), ptr(t["ps_suppkey"])
        a.ps_supplycost = ptr(t["ps_supplycost"])
        a.order_rows = int(t["o_orderdate"].numel())
        a.o_orderdate, a.order_map = ptr(t["o_orderdate"]), ptr(t["order_map"])
        a.map_size = int(t["order_map"].numel())
        a.li_rows = int(t["l_partkey"].numel())
        a.l_partkey, a.l_suppkey, a.l_orderkey = (ptr(t["l_partkey"]), ptr(t["l_suppkey"]),
                                                  ptr(t["l_orderkey"]))
        a.l_quantity, a.l_disc_price = ptr(t["l_quantity"]), ptr(t["l_disc_price"])
        a.l_supp_nation, a.n_nation = ptr(t["l_supp_nation"]), n_nation
        tot = np.zeros(2048, np.int64)
        used = np.zeros(2048, np.uint8)
        y0, span = C.c_int32(0), C.c_int32(0)
        N.check(self.lib.bolap_q9(self.ctx, C.byref(a), tot.ctypes.data_as(C.POINTER(C.c_int64)),
                                  used.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(y0),
                                  C.byref(span)))
        g = n_nation * span.value
        return y0.value, tot[:g].reshape(n_nation, span.value), used[:g].reshape(n_nation,
                                                                                 span.value)

    def string_masks(self, chars, offsets):
        torch = _torch()
        n = int(offsets.numel()) - 1
        alpha = torch.empty(max(1, n), dtype=torch.int32, device=chars.device)
        bigram = torch.empty(max(1, n), dtype=torch.int64, device=chars.device)
        N.check(self.lib.bolap_string_masks(self.ctx, ptr(chars), ptr(offsets), n, ptr(alpha),
                                            ptr(bigram)))
        return alpha[:n], bigram[:n]

    def q13(self, t: dict, w1: str, w2: str, masks=None, max_dist: int = 4096):
        a = N.Q13Args()
        a.order_rows = int(t["o_custkey_raw"].numel())
        a.o_custkey, a.chars = ptr(t["o_custkey_raw"]), ptr(t["o_comment_chars"])
        a.offsets = ptr(t["o_comment_offsets"])
        if masks is not None:
            a.alpha, a.bigram = ptr(masks[0]), ptr(masks[1])
        a.word1, a.word2 = w1.encode(), w2.encode()
        a.cust_rows = int(t["c_custkey"].numel())
        a.c_custkey = ptr(t["c_custkey"])
        dist = (C.c_int64 * max_dist)()
        mx = C.c_int32(0)
        N.check(self.lib.bolap_q13(self.ctx, C.byref(a), dist, max_dist, C.byref(mx)))
        return mx.value, list(dist)[:mx.value + 1]

    def q16(self, t: dict, n_brands: int, banned_brand: int, type_banned, size_mask: int,
            n_part_keys: int, n_supp_keys: int):
        import numpy as np
        tb = np.ascontiguousarray(type_banned, np.uint8)
        a = N.Q16Args()
        a.part_rows = int(t["p_partkey"].numel())
        a.p_partkey, a.p_brand, a.p_type = ptr(t["p_partkey"]), ptr(t["p_brand"]), ptr(t["p_type"])
        a.p_size, a.n_brands, a.n_types = ptr(t["p_size"]), n_brands, len(tb)
        a.banned_brand, a.type_banned, a.size_mask = banned_brand, tb.ctypes.data, size_mask
        a.n_part_keys = n_part_keys
        a.ps_rows = int(t["ps_partkey"].numel())
        a.ps_partkey, a.ps_suppkey = ptr(t["ps_partkey"]), ptr(t["ps_suppkey"])
        a.supp_rows = int(t["s_suppkey"].numel())
        a.s_suppkey, a.s_banned, a.n_supp_keys = ptr(t["s_suppkey"]), ptr(t["s_banned"]), n_supp_keys
        n = n_brands * len(tb) * 51
        out = np.zeros(n, np.int64)
        N.check(self.lib.bolap_q16(self.ctx, C.byref(a), out.ctypes.data_as(C.POINTER(C.c_int64)),
                                   n))
        return out

    # ----------------------------------------------- Q2, Q11, Q15, Q20 (tpch_supp)
    def q2(self, t: dict, type_ok, size: int, nation_ok, max_rows: int | None = None):
        """(part row, supplier row) pairs of run_q2 in its emission order."""
        import numpy as np
        to = np.ascontiguousarray(type_ok, np.uint8)
        no = np.ascontiguousarray(nation_ok, np.uint8)
        a = N.Q2Args()
        a.part_rows = int(t["p_size"].numel())
        a.p_size, a.p_type = ptr(t["p_size"]), ptr(t["p_type"])
        a.type_ok, a.n_types = to.ctypes.data if len(to) else None, len(to)
        a.size = size
        a.ps_rows = int(t["ps_partkey"].numel())
        a.ps_partkey, a.ps_suppkey = ptr(t["ps_partkey"]), ptr(t["ps_suppkey"])
        a.ps_supplycost = ptr(t["ps_supplycost"])
        a.supp_rows = int(t["s_nationkey"].numel())
        a.s_nationkey = ptr(t["s_nationkey"])
        a.nation_ok, a.n_nation = no.ctypes.data if len(no) else None, len(no)
        cap = max_rows if max_rows is not None else max(1, a.ps_rows)
        out = np.empty((max(cap, 1), 2), np.int32)
        n = C.c_int64(0)
        N.check(self.lib.bolap_q2(self.ctx, C.byref(a), out.ctypes.data, cap, C.byref(n)))
        return out[:n.value]

    def q11(self, t: dict, nation: int, n_supp_keys: int, fraction: float,
            max_rows: int | None = None):
        """(partkeys, values, total) of run_q11, sorted as the reference."""
        import numpy as np
        a = N.Q11Args()
        a.supp_rows = int(t["s_suppkey"].numel())
        a.s_suppkey, a.s_nationkey = ptr(t["s_suppkey"]), ptr(t["s_nationkey"])
        a.nation, a.n_supp_keys = nation, n_supp_keys
        a.ps_rows = int(t["ps_partkey"].numel())
        a.ps_partkey, a.ps_suppkey, a.ps_value = (ptr(t["ps_partkey"]), ptr(t["ps_suppkey"]),
                                                  ptr(t["ps_value"]))
        a.fraction = fraction
        cap = max_rows if max_rows is not None else max(1, a.ps_rows)
        pk, val = np.empty(max(cap, 1), np.int32), np.empty(max(cap, 1), np.int64)
        n, tot = C.c_int64(0), C.c_int64(0)
        N.check(self.lib.bolap_q11(self.ctx, C.byref(a), pk.ctypes.data, val.ctypes.data, cap,
                                   C.byref(n), C.byref(tot)))
        return pk[:n.value], val[:n.value], tot.value

    def q15(self, t: dict, date_lo: int, date_hi: int, n_supp_keys: int):
        """([(suppkey, supplier row, revenue)], lineitems in the window) of run_q15."""
        import numpy as np
        a = N.Q15Args()
        a.li_rows = int(t["l_shipdate"].numel())
        a.l_shipdate, a.l_suppkey = ptr(t["l_shipdate"]), ptr(t["l_suppkey"])
        a.l_disc_price = ptr(t["l_disc_price"])
        a.date_lo, a.date_hi = date_lo, date_hi
        a.supp_rows = int(t["s_suppkey"].numel())
        a.s_suppkey, a.n_supp_keys = ptr(t["s_suppkey"]), n_supp_keys
        cap = max(1, n_supp_keys)
        k, r, v = np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(cap, np.int64)
        n, em = C.c_int64(0), C.c_int64(0)
        N.check(self.lib.bolap_q15(self.ctx, C.byref(a), k.ctypes.data, r.ctypes.data,
                                   v.ctypes.data, cap, C.byref(n), C.byref(em)))
        return [(int(k[i]), int(r[i]), int(v[i])) for i in range(n.value)], em.value

    def q20(self, t: dict, p_ok_row, max_partkey: int, date_lo: int, date_hi: int, nation: int,
            supp_nation_by_key, max_suppkey: int, capacity_hint: int = 0):
        """(qualified[max_suppkey + 1], stats) of run_q20; p_ok_row and
        supp_nation_by_key are device tensors."""
        import numpy as np
        a = N.Q20Args()
        a.part_rows = int(t["p_partkey"].numel())
        a.p_partkey, a.p_ok_row, a.max_partkey = ptr(t["p_partkey"]), ptr(p_ok_row), max_partkey
        a.li_rows = int(t["l_shipdate"].numel())
        a.l_shipdate, a.l_partkey, a.l_suppkey = (ptr(t["l_shipdate"]), ptr(t["l_partkey"]),
                                                  ptr(t["l_suppkey"]))
        a.l_supp_nation, a.l_quantity = ptr(t["l_supp_nation"]), ptr(t["l_quantity"])
        a.date_lo, a.date_hi, a.nation = date_lo, date_hi, nation
        a.ps_rows = int(t["ps_partkey"].numel())
        a.ps_partkey, a.ps_suppkey, a.ps_availqty = (ptr(t["ps_partkey"]), ptr(t["ps_suppkey"]),
                                                     ptr(t["ps_availqty"]))
        a.supp_nation_by_key, a.n_supp_map = ptr(supp_nation_by_key), int(supp_nation_by_key.numel())
        a.max_suppkey, a.capacity_hint = max_suppkey, capacity_hint
        q = np.zeros(max(1, max_suppkey + 1), np.uint8)
        st = N.Q20Stats()
        N.check(self.lib.bolap_q20(self.ctx, C.byref(a), q.ctypes.data, C.byref(st)))
        return q[:max(0, max_suppkey + 1)], st

    @staticmethod
    def ceb_pred(col, mask, is_null=None) -> "N.CebPred":
        """A predicate over a device int32 column: value v passes when it is
        non-null and mask[v] != 0 (mask: a device uint8 tensor)."""
        return N.CebPred(ptr(col), ptr(is_null), ptr(mask), int(mask.numel()))

    def ceb_count_args(self, title: dict, terms: list, year_lo: int, year_hi: int,
                       movie_cap: int) -> "N.CebCountArgs":
        """title: {"id", "id_null"?, "year"?, "year_null"?, "preds": [CebPred]};
        terms: [{"movie", "movie_null"?, "preds": [CebPred], "weight"?}]."""
        a = N.CebCountArgs()
        a.title_rows = int(title["id"].numel())
        a.t_id, a.t_id_null = ptr(title["id"]), ptr(title.get("id_null"))
        a.t_year, a.t_year_null = ptr(title.get("year")), ptr(title.get("year_null"))
        a.year_lo, a.year_hi, a.movie_cap = int(year_lo), int(year_hi), int(movie_cap)
        a.n_title_preds = len(title.get("preds", []))
        for i, pr in enumerate(title.get("preds", [])):
            a.title_preds[i] = pr
        a.n_terms = len(terms)
        for k, tm in enumerate(terms):
            dst = a.terms[k]
            dst.rows = int(tm["movie"].numel())
            dst.movie, dst.movie_null = ptr(tm["movie"]), ptr(tm.get("movie_null"))
            dst.weight = int(tm.get("weight", 1))
            dst.n_preds = len(tm.get("preds", []))
            for i, pr in enumerate(tm.get("preds", [])):
                dst.preds[i] = pr
        return a

    def ceb_count(self, a: "N.CebCountArgs") -> dict:
        r = N.CebCountResult()
        N.check(self.lib.bolap_ceb_count(self.ctx, C.byref(a), C.byref(r)))
        return {"count": r.count, "title_pass": r.title_pass, "joined": r.joined,
                "term_rows": list(r.term_rows)[:a.n_terms]}

    def q17_args(self, t: dict, preds: list) -> N.Q17Args:
        a = N.Q17Args()
        a.part_rows = int(t["p_partkey"].numel())
        a.p_partkey = ptr(t["p_partkey"])
        a.n_part_preds = len(preds)
        for i, p in enumerate(preds):
            a.part_preds[i] = p
        a.li_rows = int(t["l_partkey"].numel())
        a.l_partkey, a.l_quantity, a.l_price = (ptr(t[k]) for k in
                                                ("l_partkey", "l_qty", "l_price"))
        return a

    def q17_run(self, a: N.Q17Args) -> dict:
        r = N.Q17Result()
        N.check(self.lib.bolap_q17(self.ctx, C.byref(a), C.byref(r)))
        return {"sum_extendedprice": r.sum_extendedprice, "parts": r.parts,
                "matches": r.matches, "rows_emitted": r.rows_emitted}

    def q17(self, t: dict, preds: list) -> dict:
        return self.q17_run(self.q17_args(t, preds))

    def gen_part_lineitem(self, seed: int, parts: int, rows: int) -> dict:
        """Part (keys 1..parts; brand U{0..24}, container U{0..39}, type
        U{0..149} with the first 25 types PROMO) and lineitem rows in the
        reference layout (ship i16, partkey U{1..parts}, quantity x100 i16,
        extendedprice i32, truncated discounted_price i32)."""
        torch = _torch()
        li = self.gen_q6_lineitem(seed, rows)
        dev = f"cuda:{self.device}"
        pk = torch.empty(max(1, parts), dtype=torch.int32, device=dev)[:parts]
        N.check(self.lib.bolap_gen_iota_i32(self.ctx, parts, 1, ptr(pk) if parts else None))
        disc = li["disc"].to(torch.int64)
        dp = (li["price"].to(torch.int64) * (100 - disc)).div(100, rounding_mode="trunc")
        return {"p_partkey": pk,
                "p_brand": self.gen_uniform_i32(seed, 0x800, parts, 0, 24).to(torch.uint8),
                "p_container": self.gen_uniform_i32(seed, 0x801, parts, 0, 39).to(torch.uint8),
                "p_promo": (self.gen_uniform_i32(seed, 0x802, parts, 0, 149) < 25).to(torch.uint8),
                "l_ship": li["ship"], "l_qty": li["qty"], "l_price": li["price"],
                "l_disc": li["disc"],
                "l_partkey": self.gen_uniform_i32(seed, 0x803, rows, 1, max(1, parts)),
                "l_disc_price": dp.to(torch.int32)}

    # -------------------------------------------------------------- Q18
    @staticmethod
    def q18_args(t: dict, threshold: int, orderkey_sorted: bool = True) -> N.Q18Args:
        a = N.Q18Args()
        a.li_rows = int(t["l_orderkey"].numel())
        a.l_orderkey, a.l_quantity = ptr(t["l_orderkey"]), ptr(t["l_qty"])
        a.orderkey_sorted = int(orderkey_sorted)
        a.order_rows = int(t["o_orderkey"].numel())
        a.o_orderkey, a.o_custkey = ptr(t["o_orderkey"]), ptr(t["o_custkey"])
        a.o_orderdate, a.o_totalprice = ptr(t["o_orderdate"]), ptr(t["o_totalprice"])
        a.cust_rows = int(t["c_custkey"].numel())
        a.c_custkey = ptr(t["c_custkey"])
        a.quantity_threshold = int(threshold)
        return a

    def q18_run(self, a: N.Q18Args, max_rows: int = 1 << 16):
        """Result rows as a structured numpy array (oracle Q18_DTYPE layout)
        and the stats dict."""
        import numpy as np
        rows = (N.Q18Row * max(1, max_rows))()
        st = N.Q18Stats()
        N.check(self.lib.bolap_q18(self.ctx, C.byref(a), rows, max_rows, C.byref(st)))
        dt = np.dtype([("custkey", "<i4"), ("orderkey", "<i4"), ("orderdate", "<i4"),
                       ("totalprice", "<i4"), ("sum_qty", "<i8"), ("cust_row", "<i4"),
                       ("order_row", "<i4")])
        arr = np.frombuffer(bytes(rows), dtype=dt)[:st.rows].copy()
        return arr, {"rows": st.rows, "dense_path": st.dense_path,
                     "sorted_pass": st.sorted_pass, "fallbacks": st.fallbacks}

    def q18(self, t: dict, threshold: int, orderkey_sorted: bool = True, max_rows: int = 1 << 16):
        return self.q18_run(self.q18_args(t, threshold, orderkey_sorted), max_rows)

    # ---------------------------------------------------- build derivations
    def derive_lineitem(self, price, disc, rf, ls, n_ls: int, orderkey=None):
        torch = _torch()
        rows = int(price.numel())
        dp = torch.empty(rows, dtype=torch.int32, device=price.device)
        code = torch.empty(rows, dtype=torch.uint8, device=price.device)
        s = C.c_int32()
        N.check(self.lib.bolap_derive_lineitem(self.ctx, rows, ptr(price), ptr(disc), ptr(rf),
                                               ptr(ls), n_ls, ptr(orderkey), ptr(dp), ptr(code),
                                               C.byref(s)))
        return dp, code, bool(s.value)

    def date_order(self, date):
        torch = _torch()
        rows = int(date.numel())
        out = torch.empty(max(1, rows), dtype=torch.int32, device=date.device)
        N.check(self.lib.bolap_date_order(self.ctx, rows, ptr(date), ptr(out)))
        return out[:rows]

    def order_ranges(self, o_orderkey, l_orderkey, map_size: int):
        torch = _torch()
        o_rows = int(o_orderkey.numel())
        dev = o_orderkey.device
        m = torch.empty(max(1, map_size), dtype=torch.int32, device=dev)
        start = torch.empty(max(1, o_rows), dtype=torch.int32, device=dev)
        end = torch.empty(max(1, o_rows), dtype=torch.int32, device=dev)
        N.check(self.lib.bolap_order_ranges(self.ctx, o_rows, ptr(o_orderkey), map_size, ptr(m),
                                            int(l_orderkey.numel()), ptr(l_orderkey),
                                            ptr(start), ptr(end)))
        return m[:map_size], start[:o_rows], end[:o_rows]

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

    def gen_uniform_i32(self, seed: int, stream: int, rows: int, lo: int, hi: int):
        torch = _torch()
        out = torch.empty(max(1, rows), dtype=torch.int32, device=f"cuda:{self.device}")
        N.check(self.lib.bolap_gen_uniform_i32(self.ctx, seed, stream, rows, lo, hi, ptr(out)))
        return out[:rows]

    def gen_q6_lineitem(self, seed: int, rows: int) -> dict:
        """Q6 columns in the reference layout (builder_impl.hpp:39-66): ship
        i16 day offset, discount u8 hundredths, quantity i16 x 100, price i32
        cents, from the seeded lineitem generator."""
        torch = _torch()
        w = self.gen_lineitem(seed, rows)
        return {"ship": w["ship"].to(torch.int16), "disc": w["disc"].to(torch.uint8),
                "qty": (w["qty"] * 100).to(torch.int16), "price": w["price"].to(torch.int32)}

    def gen_q18_tables(self, seed: int, cust_rows: int, order_rows: int) -> dict:
        """Customers, orders and lineitem for Q18 in the reference layout:
        lineitem sorted by order key with 1..7 rows per order (the chain
        generator), quantity U{1..50} x 100 (i16), order date i16, total price
        U{1000.00..500000.00} in cents (i32); keys 1..N."""
        torch = _torch()
        t = self.gen_chain(seed, cust_rows, order_rows)
        n = int(t["fact_order"].numel())
        qty = self.gen_uniform_i32(seed, 0x700, n, 1, 50)
        price = self.gen_uniform_i32(seed, 0x701, order_rows, 100_000, 50_000_000)
        return {"l_orderkey": t["fact_order"], "l_qty": (qty * 100).to(torch.int16),
                "o_orderkey": t["order_key"], "o_custkey": t["order_cust"],
                "o_orderdate": t["order_day"].to(torch.int16), "o_totalprice": price,
                "c_custkey": t["cust_key"]}

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

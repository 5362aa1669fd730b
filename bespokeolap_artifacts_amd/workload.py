"""Star-schema workloads: seeded star tables in HBM and random star plans.

A star schema here is one fact table with three foreign-key columns
(fk0..fk2, Zipf-distributed over the keys of dimension 0..2), a group column
`grp` (Zipf over 1..n_groups) and price/disc, plus three dimension tables
with dense keys 1..rows, a uint8 column `cat` U{0..255} and an int32 column
`val` U{0..999,999}.  A plan joins the fact table with 1..3 filtered
dimensions and groups on `grp` or on one of its foreign keys; the C ABI
entry is bolap_star_groupby.

Plans are pure data (column names and constants), so the same plan binds to
device tensors (`bind`) and, in the tests, to host arrays.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import native as N

ZIPF_STEPS = 65536
VAL_MAX = 999_999
CAT_MAX = 255
N_DIMS = 3

# Generator stream ids (one independent sequence per column).
def dim_stream(j: int, col: int) -> int:
    return 0x100 + 16 * j + col


FK_STREAM = 0x200
GRP_STREAM = 0x210
PRICE_STREAM = 0x220


def coprime_mul(n: int) -> int:
    """Multiplier that scatters Zipf ranks over 1..n: the first odd number
    from 1,000,003 up that is coprime to n."""
    m = 1_000_003
    while math.gcd(m, n) != 1:
        m += 2
    return m


def zipf_quantiles(n_keys: int, s: float) -> np.ndarray:
    """Inverse CDF of Zipf(s) over ranks 1..n_keys at the 65,537 points
    i / 65536: q[i] is the smallest rank whose cumulative probability reaches
    i / 65536.  Ranks up to 2^22 use exact cumulative sums; beyond that the
    tail uses the integral of x^-s (Euler-Maclaurin leading term)."""
    n = int(n_keys)
    t = np.arange(ZIPF_STEPS + 1, dtype=np.float64) / ZIPF_STEPS
    if s == 0.0:
        q = np.ceil(t * n)
    else:
        m = min(n, 1 << 22)
        w = np.arange(1, m + 1, dtype=np.float64) ** (-s)
        head = np.cumsum(w)
        hm = head[-1]

        def tail_mass(x):
            if abs(s - 1.0) < 1e-12:
                return np.log(x / m)
            return (x ** (1.0 - s) - m ** (1.0 - s)) / (1.0 - s)

        total = hm + (tail_mass(float(n)) if n > m else 0.0)
        target = t * total
        q = np.empty_like(t)
        in_head = target <= hm
        q[in_head] = np.searchsorted(head, target[in_head], side="left") + 1
        rest = target[~in_head] - hm
        if rest.size:
            if abs(s - 1.0) < 1e-12:
                x = m * np.exp(rest)
            else:
                x = (m ** (1.0 - s) + (1.0 - s) * rest) ** (1.0 / (1.0 - s))
            q[~in_head] = np.ceil(x)
    q = np.clip(q, 1, n)
    q[0] = 1
    q[-1] = n
    q = np.maximum.accumulate(q)
    return q.astype(np.int32)


class StarSchema:
    """Seeded star tables generated in HBM (see the module docstring)."""

    def __init__(self, engine, seed: int, fact_rows: int, dim_rows, fk_s, n_groups: int,
                 group_s: float):
        import torch
        self.engine = engine
        self.seed = int(seed)
        self.fact_rows = int(fact_rows)
        self.dim_rows = [int(r) for r in dim_rows]
        self.fk_s = [float(x) for x in fk_s]
        self.n_groups = int(n_groups)
        self.group_s = float(group_s)
        lib, ctx = engine.lib, engine.ctx
        dev = f"cuda:{engine.device}"
        self.dims = []
        for j, rows in enumerate(self.dim_rows):
            d = {"key": torch.empty(rows, dtype=torch.int32, device=dev),
                 "cat": torch.empty(rows, dtype=torch.uint8, device=dev),
                 "val": torch.empty(rows, dtype=torch.int32, device=dev)}
            N.check(lib.bolap_gen_iota_i32(ctx, rows, 1, _ptr(d["key"])))
            N.check(lib.bolap_gen_uniform_u8(ctx, self.seed, dim_stream(j, 1), rows, 0, CAT_MAX,
                                             _ptr(d["cat"])))
            N.check(lib.bolap_gen_uniform_i32(ctx, self.seed, dim_stream(j, 2), rows, 0,
                                              VAL_MAX, _ptr(d["val"])))
            self.dims.append(d)
        f = {}
        n = self.fact_rows
        for j, rows in enumerate(self.dim_rows):
            q = torch.from_numpy(zipf_quantiles(rows, self.fk_s[j])).to(dev)
            f[f"fk{j}"] = torch.empty(n, dtype=torch.int32, device=dev)
            N.check(lib.bolap_gen_zipf_i32(ctx, self.seed, FK_STREAM + j, n, _ptr(q), rows,
                                           coprime_mul(rows), _ptr(f[f"fk{j}"])))
        q = torch.from_numpy(zipf_quantiles(self.n_groups, self.group_s)).to(dev)
        f["grp"] = torch.empty(n, dtype=torch.int32, device=dev)
        N.check(lib.bolap_gen_zipf_i32(ctx, self.seed, GRP_STREAM, n, _ptr(q), self.n_groups,
                                       coprime_mul(self.n_groups), _ptr(f["grp"])))
        f["price"] = torch.empty(n, dtype=torch.int64, device=dev)
        f["disc"] = torch.empty(n, dtype=torch.int64, device=dev)
        N.check(lib.bolap_gen_price_disc(ctx, self.seed, PRICE_STREAM, n, _ptr(f["price"]),
                                         _ptr(f["disc"])))
        self.fact = f
        engine.sync()

    def key_range(self, col: str) -> tuple[int, int]:
        """[min, max] of a fact key column (the generator's domain)."""
        if col == "grp":
            return 1, self.n_groups
        return 1, self.dim_rows[int(col[2:])]

    def describe(self) -> dict:
        return {"seed": self.seed, "fact_rows": self.fact_rows, "dim_rows": self.dim_rows,
                "fk_s": self.fk_s, "n_groups": self.n_groups, "group_s": self.group_s}


def _ptr(t):
    return C.c_void_p(t.data_ptr())


# ------------------------------------------------------------------ plans
def random_plan(rng: np.random.Generator) -> dict:
    """A star plan with 1..3 filtered dimensions and a group column.

    Dimension predicates: a range on `val` with a log-uniform width, an
    equality, IN-list or range on `cat`, or both; selectivities span about
    1e-6 .. 1."""
    n_dims = int(rng.integers(1, N_DIMS + 1))
    which = sorted(rng.choice(N_DIMS, n_dims, replace=False).tolist())
    dims = []
    for j in which:
        preds = []
        kind = int(rng.integers(0, 4))
        if kind in (0, 3):
            frac = 10 ** rng.uniform(-5.5, 0.0)
            width = max(1, int(frac * (VAL_MAX + 1)))
            lo = int(rng.integers(0, VAL_MAX + 2 - width))
            preds.append({"col": "val", "op": N.OP_RANGE, "lo": lo, "hi": lo + width - 1})
        if kind in (1, 3):
            m = int(rng.integers(1, 17))
            values = sorted(rng.choice(CAT_MAX + 1, m, replace=False).tolist())
            preds.append({"col": "cat", "op": N.OP_IN_SET, "values": values})
        if kind == 2:
            w = int(rng.integers(1, 129))
            lo = int(rng.integers(0, CAT_MAX + 2 - w))
            preds.append({"col": "cat", "op": N.OP_RANGE, "lo": lo, "hi": lo + w - 1})
        dims.append({"dim": j, "preds": preds})
    if rng.random() < 0.6:
        group = "grp"
    else:
        group = f"fk{which[int(rng.integers(0, n_dims))]}"
    return {"dims": dims, "group": group}


def rotate(plan: dict, r: int) -> dict:
    """The same plan with its dimension list rotated left by r (slot 0 is the
    dimension whose selectivity the estimator's first output describes)."""
    d = plan["dims"]
    r %= len(d)
    return {"dims": d[r:] + d[:r], "group": plan["group"]}


def bind(schema: StarSchema, plan: dict, pass_hints=None, group_hint: int = 0) -> N.StarArgs:
    """C ABI arguments of a plan over device tables."""
    from .engine import Engine
    a = N.StarArgs()
    a.n_dims = len(plan["dims"])
    for i, d in enumerate(plan["dims"]):
        j = d["dim"]
        t = schema.dims[j]
        sd = a.dims[i]
        sd.rows = schema.dim_rows[j]
        sd.key = _ptr(t["key"]).value
        sd.key_min, sd.key_max = 1, schema.dim_rows[j]
        sd.n_preds = len(d["preds"])
        for k, p in enumerate(d["preds"]):
            sd.preds[k] = Engine.pred(t[p["col"]], p["op"], p.get("lo", 0), p.get("hi", 0),
                                      values=p.get("values"))
        sd.fact_fk = _ptr(schema.fact[f"fk{j}"]).value
        sd.pass_hint = int(pass_hints[i]) if pass_hints is not None else 0
    a.fact_rows = schema.fact_rows
    a.group_key = _ptr(schema.fact[plan["group"]]).value
    a.group_min, a.group_max = schema.key_range(plan["group"])
    a.fact_price = _ptr(schema.fact["price"]).value
    a.price_dtype = N.I64
    a.fact_disc = _ptr(schema.fact["disc"]).value
    a.disc_dtype = N.I64
    a.group_hint = int(group_hint)
    return a

"""Learned cardinality estimator for star plans.

No reference counterpart: the reference sizes hash tables from hard-coded
selectivities (bespoke_tpch/query_q20.cpp:16, 405-417) and grows them on
overflow (:158-160); its `TRACE` COUNT lines (e.g. bespoke_tpch/query_q1.cpp:
78-92) are the cardinalities a plain engine records.  Here the plain engine
(bolap_star_groupby with every hint 0, which counts before it builds) records
those counts on seeded synthetic star schemas; an MLP 256-128-64-3 (BF16
MFMA, FP32 accumulation, trained on the device by bolap_mlp_train_step)
learns

    y0 = log10((pass0 + 1) / (dim rows0 + 1))          filter selectivity of slot 0
    y1 = log10((join rows + 1) / (fact rows + 1))       join size
    y2 = log10((groups + 1) / (min(fact rows, key span) + 1))   group count

from 256 features: 64-bin histograms of the HyperLogLog registers of the
slot-0 foreign key, the group key and the slot-1 foreign key, the sample
counts of the plan on 4096 sampled rows (bolap_star_sample_counts), and the
normalised predicate constants.  A plan with k dimensions is fed k times,
rotated so that each dimension takes slot 0 once; y1 and y2 are averaged
over the rotations.

The sizing policy turns predictions into the pipeline's hints (pass_hint per
dimension, group_hint); a hint that proves too small costs one exact rerun,
never a wrong result.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import tempfile
import time

import numpy as np

from . import native as N
from . import workload as W

N_FEAT = 256
N_OUT = 3
HIST_BINS = 64
SAMPLE = N.SAMPLE_ROWS
ALPHA_M = 0.7213 / (1 + 1.079 / N.HLL_REGS)


# ------------------------------------------------------------- statistics
def hll_hist(regs: np.ndarray) -> np.ndarray:
    """Histogram of register values 0..63, as fractions of the 4096 registers."""
    return np.bincount(np.minimum(regs, HIST_BINS - 1), minlength=HIST_BINS)[:HIST_BINS] \
        .astype(np.float32) / N.HLL_REGS


def hll_estimate(regs: np.ndarray) -> float:
    """HyperLogLog cardinality (Flajolet et al. 2007) with linear counting
    for small ranges, m = 4096.  The harmonic sum runs over register values
    in ascending order (exact terms), as the library's estimator sums it."""
    m = N.HLL_REGS
    cnt = np.bincount(regs.astype(np.int64), minlength=256)
    s = 0.0
    for v in range(len(cnt)):
        if cnt[v]:
            s += math.ldexp(float(cnt[v]), -v)
    est = ALPHA_M * m * m / s
    zeros = int(cnt[0])
    if est <= 2.5 * m and zeros:
        est = m * math.log(m / zeros)
    return est


class SchemaStats:
    """Per-column statistics of one star schema, computed once at load."""

    def __init__(self, fact_rows: int, dim_rows, n_groups: int, regs: dict):
        self.fact_rows = int(fact_rows)
        self.dim_rows = [int(r) for r in dim_rows]
        self.n_groups = int(n_groups)
        self.regs = regs
        self.hist = {c: hll_hist(r) for c, r in regs.items()}
        self.distinct = {c: hll_estimate(r) for c, r in regs.items()}

    @classmethod
    def compute(cls, engine, schema: W.StarSchema) -> "SchemaStats":
        regs = {c: engine.hll(schema.fact[c]) for c in ("fk0", "fk1", "fk2", "grp")}
        return cls(schema.fact_rows, schema.dim_rows, schema.n_groups, regs)

    def span(self, col: str) -> int:
        return self.n_groups if col == "grp" else self.dim_rows[int(col[2:])]


# --------------------------------------------------------------- features
def _l(x: float, scale: float) -> float:
    return math.log10(max(x, 1e-12)) / scale


def features(ss: SchemaStats, plan: dict, counts: np.ndarray) -> np.ndarray:
    """256 features of a plan whose dimension slots are in plan order;
    `counts` are the plan's sample counts in the same slot order."""
    f = np.zeros(N_FEAT, np.float32)
    dims = plan["dims"]
    fk = [f"fk{d['dim']}" for d in dims]
    f[0:64] = ss.hist[fk[0]]
    f[64:128] = ss.hist[plan["group"]]
    if len(dims) > 1:
        f[128:192] = ss.hist[fk[1]]
    o = 192
    for j in range(W.N_DIMS):
        if j < len(dims):
            d = dims[j]
            rows = ss.dim_rows[d["dim"]]
            f[o + 0] = 1.0
            f[o + 1] = _l(rows, 8)
            for p in d["preds"]:
                if p["col"] == "val":
                    w = (p["hi"] - p["lo"] + 1) / (W.VAL_MAX + 1)
                    f[o + 2] = 1.0
                    f[o + 3] = p["lo"] / (W.VAL_MAX + 1)
                    f[o + 4] = w
                    f[o + 5] = _l(w, 6)
                else:
                    n = len(p["values"]) if p["op"] == N.OP_IN_SET else p["hi"] - p["lo"] + 1
                    f[o + 6] = 1.0
                    f[o + 7] = n / (W.CAT_MAX + 1)
            f[o + 8] = counts[j] / SAMPLE
            f[o + 9] = _l((counts[j] + 0.5) / SAMPLE, 4)
            f[o + 10] = counts[3 + j] / SAMPLE
            f[o + 11] = _l((counts[3 + j] + 0.5) / SAMPLE, 4)
        o += 12
    nf = ss.fact_rows
    gspan = ss.span(plan["group"])
    joined = nf * counts[6] / SAMPLE
    f[228] = counts[6] / SAMPLE
    f[229] = _l((counts[6] + 0.5) / SAMPLE, 4)
    f[230] = _l(nf, 8)
    f[231] = _l(ss.distinct[plan["group"]] + 1, 8)
    f[232] = _l(gspan, 8)
    if plan["group"] != "grp":
        f[233] = 1.0
        slot = fk.index(plan["group"]) if plan["group"] in fk else -1
        if slot >= 0:
            f[234 + slot] = 1.0
    f[237] = counts[7] / SAMPLE
    f[238] = counts[8] / SAMPLE
    f[239] = counts[9] / SAMPLE
    f[240] = _l(counts[7] + 1, 4)
    f[241] = _l(gee_groups(counts, joined) + 1, 8)
    f[242] = _l(joined + 1, 8)
    f[243] = len(dims) / 3
    for j in range(min(3, len(dims))):
        f[244 + j] = _l(ss.distinct[fk[j]] + 1, 8)
    sel = 1.0
    for j in range(len(dims)):
        sel *= (counts[j] + 0.5) / SAMPLE
    f[247] = _l(sel, 8)
    f[255] = 1.0
    return f


def gee_groups(counts: np.ndarray, joined: float) -> float:
    """Guaranteed-error estimator of distinct values (Charikar et al. 2000):
    sqrt(N / n) f1 + sum_{i >= 2} f_i from the sampled joined rows."""
    n = counts[6]
    if n <= 0:
        return 0.0
    d, f1 = counts[7], counts[8]
    return math.sqrt(max(joined, n) / n) * f1 + (d - f1)


def rotated_counts(counts: np.ndarray, r: int, k: int) -> np.ndarray:
    c = counts.copy()
    idx = [(i + r) % k for i in range(k)]
    c[0:k] = counts[0:k][idx]
    c[3:3 + k] = counts[3:3 + k][idx]
    return c


def targets(ss: SchemaStats, plan: dict, stats: dict) -> np.ndarray:
    rows0 = ss.dim_rows[plan["dims"][0]["dim"]]
    nf = ss.fact_rows
    bound = min(nf, ss.span(plan["group"]))
    return np.array([math.log10((stats["dim_pass"][0] + 1) / (rows0 + 1)),
                     math.log10((stats["matches"] + 1) / (nf + 1)),
                     math.log10((stats["groups"] + 1) / (bound + 1))], np.float32)


def from_targets(ss: SchemaStats, plan: dict, y: np.ndarray) -> tuple[float, float, float]:
    """(passing rows of slot 0, join rows, groups) from the three outputs."""
    rows0 = ss.dim_rows[plan["dims"][0]["dim"]]
    nf = ss.fact_rows
    bound = min(nf, ss.span(plan["group"]))
    return (max(0.0, 10 ** float(y[0]) * (rows0 + 1) - 1),
            max(0.0, 10 ** float(y[1]) * (nf + 1) - 1),
            max(0.0, 10 ** float(y[2]) * (bound + 1) - 1))


# ---------------------------------------------------------- records
class Record:
    """A plan with its rotated features and the plain engine's counts."""
    __slots__ = ("plan", "ss", "counts", "stats", "x", "y")

    def __init__(self, plan, ss, counts, stats):
        self.plan, self.ss, self.counts, self.stats = plan, ss, counts, stats
        k = len(plan["dims"])
        xs, ys = [], []
        for r in range(k):
            p = W.rotate(plan, r)
            xs.append(features(ss, p, rotated_counts(counts, r, k)))
            if stats is not None:
                st = dict(stats)
                st["dim_pass"] = [stats["dim_pass"][(i + r) % k] for i in range(k)]
                ys.append(targets(ss, p, st))
        self.x = np.stack(xs)
        self.y = np.stack(ys) if ys else None


SAMPLE_SEED = N.EST_SAMPLE_SEED  # the library estimator samples with the same seed


def record(engine, schema: W.StarSchema, ss: SchemaStats, plans: list[dict],
           label: bool = True) -> list[Record]:
    """Sample counts for every plan (one launch) and, when `label`, the plain
    engine's exact counts (every hint 0: count, then build)."""
    args = [W.bind(schema, p) for p in plans]
    counts = engine.sample_counts(args, SAMPLE_SEED)
    out = []
    for p, a, c in zip(plans, args, counts):
        stats = engine.star_groupby(a)[0] if label else None
        out.append(Record(p, ss, c, stats))
    return out


# ------------------------------------------------------------- the model
class Estimator:
    """The MLP, its target standardisation and the sizing policy."""

    def __init__(self, engine, seed: int = 1):
        self.engine = engine
        h = C.c_void_p()
        N.check(engine.lib.bolap_mlp_create(engine.ctx, seed, C.byref(h)))
        self.h = h
        self.mu = np.zeros(N_OUT, np.float32)
        self.sd = np.ones(N_OUT, np.float32)
        self.margin = {"dim": 2.0, "group": 2.0}
        self.train_seconds = 0.0
        self.losses = []

    def close(self):
        if self.h:
            self.engine.lib.bolap_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _dev(self, a):
        import torch
        return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(
            f"cuda:{self.engine.device}")

    def fit(self, recs: list[Record], epochs: int = 40, batch: int = 512, lr: float = 2e-3,
            seed: int = 0) -> None:
        import torch
        x = np.concatenate([r.x for r in recs])
        y = np.concatenate([r.y for r in recs])
        self.mu = y.mean(axis=0)
        self.sd = y.std(axis=0) + 1e-6
        yn = (y - self.mu) / self.sd
        xd, yd = self._dev(x), self._dev(yn)
        n = len(x)
        rng = np.random.default_rng(seed)
        lib = self.engine.lib
        loss = C.c_float(0)
        t0 = time.perf_counter()
        steps = 0
        for ep in range(epochs):
            perm = torch.from_numpy(rng.permutation(n)).to(xd.device)
            xs, ys = xd[perm].contiguous(), yd[perm].contiguous()
            step_lr = lr * (0.5 * (1 + math.cos(math.pi * ep / epochs)))
            tot = 0.0
            for b0 in range(0, n - batch + 1, batch):
                xb, yb = xs[b0:b0 + batch], ys[b0:b0 + batch]
                N.check(lib.bolap_mlp_train_step(self.h, C.c_void_p(xb.data_ptr()),
                                                 C.c_void_p(yb.data_ptr()), batch,
                                                 step_lr, None if b0 else C.byref(loss)))
                steps += 1
                if b0 == 0:
                    tot = loss.value
            self.losses.append(tot)
        self.engine.sync()
        self.train_seconds = time.perf_counter() - t0
        self.train_steps = steps

    def predict_rows(self, x: np.ndarray) -> np.ndarray:
        import torch
        xd = self._dev(x)
        yd = torch.empty((len(x), N_OUT), dtype=torch.float32, device=xd.device)
        N.check(self.engine.lib.bolap_mlp_predict(self.h, C.c_void_p(xd.data_ptr()), len(x),
                                                  C.c_void_p(yd.data_ptr())))
        return yd.cpu().numpy() * self.sd + self.mu

    def predict(self, recs: list[Record]) -> list[dict]:
        """Per record: predicted passing rows per dimension (plan order),
        join rows and groups (averaged over the rotations)."""
        x = np.concatenate([r.x for r in recs])
        y = self.predict_rows(x)
        out, i = [], 0
        for r in recs:
            k = len(r.plan["dims"])
            yy = y[i:i + k]
            i += k
            passes, joins, groups = [], [], []
            for rot in range(k):
                p = W.rotate(r.plan, rot)
                a, b, c = from_targets(r.ss, p, yy[rot])
                passes.append(a)
                joins.append(b)
                groups.append(c)
            # mean in log1p space, summed in rotation order (as the library)
            out.append({"dim_pass": passes,
                        "matches": math.exp(sum(math.log1p(j) for j in joins) / k) - 1,
                        "groups": math.exp(sum(math.log1p(g) for g in groups) / k) - 1})
        return out

    def hints(self, rec: Record, pred: dict) -> tuple[list[int], int]:
        return sizing_hints(rec, pred, self.margin)

    def save(self, path: str) -> None:
        """The trained estimator as a file bolap_estimator_load reads: MLP
        parameters, output standardisation and margins."""
        lib = self.engine.lib
        params = np.zeros(lib.bolap_mlp_num_params(), np.float32)
        N.check(lib.bolap_mlp_get_params(self.h, params.ctypes.data_as(C.POINTER(C.c_float))))
        musd = np.concatenate([self.mu, self.sd]).astype(np.float32)
        margins = np.array([self.margin["dim"], self.margin["group"]], np.float64)
        est = C.c_void_p()
        N.check(lib.bolap_estimator_create(self.engine.ctx, params.ctypes.data,
                                           musd.ctypes.data, margins.ctypes.data, C.byref(est)))
        try:
            N.check(lib.bolap_estimator_save(est, str(path).encode()))
        finally:
            lib.bolap_estimator_destroy(est)


def sample_estimate(rec: Record) -> dict:
    """The estimate without the network: independent sample selectivities,
    the sampled join fraction and the GEE distinct estimator."""
    c = rec.counts
    ss = rec.ss
    passes = []
    for j, d in enumerate(rec.plan["dims"]):
        passes.append(max(c[j], 0.5) / SAMPLE * ss.dim_rows[d["dim"]])
    joined = ss.fact_rows * max(c[6], 0.5) / SAMPLE
    return {"dim_pass": passes, "matches": joined, "groups": max(gee_groups(c, joined), 1.0)}


def sizing_hints(rec: Record, pred: dict, margin: dict) -> tuple[list[int], int]:
    """Pipeline hints from an estimate: each is the estimate times a margin,
    at least 1; the group hint is capped by the key span and the join
    estimate and is never below the distinct keys already seen in the
    sample (a true lower bound)."""
    ss = rec.ss
    dims = [max(1, int(math.ceil(p * margin["dim"]))) for p in pred["dim_pass"]]
    g = pred["groups"] * margin["group"]
    g = min(g, ss.span(rec.plan["group"]), max(pred["matches"] * margin["group"], 1.0))
    g = max(g, float(rec.counts[7]), 1.0)
    return dims, int(math.ceil(g))


# --------------------------------------------------- the library estimator
def est_columns(schema: W.StarSchema, ss: SchemaStats):
    """Column statistics of a star schema as bolap_estimator_hints takes
    them: each fact key (FK) and the group column (GROUP) with their
    HyperLogLog registers, each dimension's value (VALUE) and category
    (CATEGORY) columns with their domains."""
    cols = []

    def add(t, kind, span, regs=None):
        c = N.EstColumn()
        c.col = t.data_ptr()
        c.kind = kind
        c.span = int(span)
        if regs is not None:
            C.memmove(c.hll, np.ascontiguousarray(regs, np.uint8).ctypes.data, N.HLL_REGS)
        cols.append(c)

    for j, rows in enumerate(schema.dim_rows):
        add(schema.fact[f"fk{j}"], N.EST_FK, rows, ss.regs[f"fk{j}"])
        add(schema.dims[j]["val"], N.EST_VALUE, W.VAL_MAX + 1)
        add(schema.dims[j]["cat"], N.EST_CATEGORY, W.CAT_MAX + 1)
    add(schema.fact["grp"], N.EST_GROUP, schema.n_groups, ss.regs["grp"])
    return (N.EstColumn * len(cols))(*cols)


class LibraryEstimator:
    """A saved estimator loaded into the library (bolap_estimator_load)."""

    def __init__(self, engine, path: str):
        self.lib = engine.lib
        self.h = C.c_void_p()
        N.check(self.lib.bolap_estimator_load(engine.ctx, str(path).encode(), C.byref(self.h)))

    def hints(self, cols, args: list) -> list[tuple[list[int], int]]:
        """(pass hints in plan order, group hint) per bound plan."""
        arr = (N.StarArgs * len(args))(*args)
        N.check(self.lib.bolap_estimator_hints(self.h, cols, len(cols), arr, len(args)))
        return [([int(a.dims[i].pass_hint) for i in range(a.n_dims)], int(a.group_hint))
                for a in arr]

    def close(self):
        if self.h:
            self.lib.bolap_estimator_destroy(self.h)
            self.h = None


def q_error(est: float, true: float) -> float:
    e, t = max(est, 1.0), max(true, 1.0)
    return max(e / t, t / e)


def q_errors(recs: list[Record], preds: list[dict]) -> dict:
    qs = {"dim_pass": [], "matches": [], "groups": []}
    for r, p in zip(recs, preds):
        for a, b in zip(p["dim_pass"], r.stats["dim_pass"]):
            qs["dim_pass"].append(q_error(a, b))
        qs["matches"].append(q_error(p["matches"], r.stats["matches"]))
        qs["groups"].append(q_error(p["groups"], r.stats["groups"]))
    return {k: np.array(v) for k, v in qs.items()}


def percentiles(q: np.ndarray) -> dict:
    return {f"p{p}": float(np.percentile(q, p)) for p in (50, 90, 95, 99)} | \
        {"max": float(q.max())}


def would_fall_back(rec: Record, dims: list[int], group_hint: int) -> bool:
    """Whether the pipeline would rerun exactly with these hints (mirrors the
    overflow rules of bolap_star_groupby: an LDS key set over
    BOLAP_STAR_LDS_KEYS keys, or a group table over half full)."""
    for h, t in zip(dims, rec.stats["dim_pass"]):
        if h <= N.STAR_LDS_KEYS and t > N.STAR_LDS_KEYS:
            return True
    cap = 1 << max(10, math.ceil(math.log2(2 * max(group_hint, 1))))
    return rec.stats["groups"] > cap // 2


def tune_margins(recs: list[Record], preds: list[dict], target: float = 0.02) -> dict:
    """Smallest margins (from a fixed ladder) whose fallback rate on the
    training plans is at most `target`: the dimension margin first (with the
    largest group margin), then the group margin."""
    ladder = [1.25, 1.5, 2.0, 3.0, 4.0, 6.0, 8.0, 16.0]

    def rate(margin):
        fb = sum(would_fall_back(r, *sizing_hints(r, p, margin)) for r, p in zip(recs, preds))
        return fb / max(1, len(recs))

    best = {"dim": ladder[-1], "group": ladder[-1]}
    for m in ladder:
        if rate({"dim": m, "group": ladder[-1]}) <= target / 2:
            best["dim"] = m
            break
    for m in ladder:
        if rate({"dim": best["dim"], "group": m}) <= target:
            best["group"] = m
            break
    return best


# ------------------------------------------------------------ experiment
S_VALUES = (0.0, 0.5, 1.0, 1.5)


def schema_specs(n: int, seed: int, fact_rows=None) -> list[dict]:
    """Seeded star-schema shapes: fact tables of 1M-60M rows, dimensions of
    100k-4M rows, foreign keys and group keys Zipf with s in {0, .5, 1, 1.5}."""
    rng = np.random.default_rng(seed)
    sizes = [1_000_000, 4_000_000, 15_000_000, 60_000_000]
    out = []
    for i in range(n):
        out.append(dict(seed=int(rng.integers(1, 2**31)),
                        fact_rows=int(fact_rows or sizes[i % len(sizes)]),
                        dim_rows=[int(x) for x in rng.choice([100_000, 1_000_000, 4_000_000], 3)],
                        fk_s=[float(x) for x in rng.choice(S_VALUES, 3)],
                        n_groups=int(rng.choice([1_000, 100_000, 1_000_000, 10_000_000])),
                        group_s=float(rng.choice(S_VALUES))))
    return out


def collect(engine, specs: list[dict], plans_per_schema: int, seed: int,
            keep_last: bool = False, log=None):
    """Records (features + plain-engine labels) of random plans over each
    schema; returns the records and, when keep_last, the last schema."""
    rng = np.random.default_rng(seed)
    recs, schema = [], None
    for i, spec in enumerate(specs):
        schema = W.StarSchema(engine, **spec)
        ss = SchemaStats.compute(engine, schema)
        plans = [W.random_plan(rng) for _ in range(plans_per_schema)]
        recs.extend(record(engine, schema, ss, plans))
        if log:
            log(f"schema {i + 1}/{len(specs)}: {spec['fact_rows']:,} fact rows, "
                f"{len(recs)} plans recorded")
        if not (keep_last and i == len(specs) - 1):
            del schema
            schema = None
    return recs, schema


def _timed(engine, a, reps: int):
    best, out = None, None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = engine.star_groupby(a, want_rows=True)
        engine.sync()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return best, out


def evaluate(engine, est: "Estimator", schema: W.StarSchema, recs: list[Record],
             reps: int = 3) -> dict:
    """Learned-sized vs exact-sized runs of held-out plans over one schema:
    raw q-errors of the network and of the sample-only estimate, fallbacks,
    times, and whether every learned-sized result is bit-identical to the
    exact one.  `summarize` folds several of these."""
    import torch
    t0 = time.perf_counter()
    args = [W.bind(schema, r.plan) for r in recs]
    counts = engine.sample_counts(args, SAMPLE_SEED)
    for r, c in zip(recs, counts):
        assert np.array_equal(r.counts, c)
    preds = est.predict(recs)
    engine.sync()
    infer_s = time.perf_counter() - t0
    # the same hints from the library (saved weights, C featurisation)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "estimator.bin")
        est.save(path)
        lib_est = LibraryEstimator(engine, path)
    cols = est_columns(schema, recs[0].ss)
    t1 = time.perf_counter()
    lib_hints = lib_est.hints(cols, [W.bind(schema, r.plan) for r in recs])
    lib_s = time.perf_counter() - t1
    lib_est.close()
    base = [sample_estimate(r) for r in recs]
    out = {"t_exact": 0.0, "t_learned": 0.0, "t_base": 0.0, "infer_s": infer_s,
           "library_s": lib_s, "library_same": 0,
           "fallbacks": 0, "base_fallbacks": 0, "identical": True, "n": len(recs),
           "qn": q_errors(recs, preds), "qb": q_errors(recs, base)}
    for r, p, b, lh in zip(recs, preds, base, lib_hints):
        te, (st_e, ke, re_) = _timed(engine, W.bind(schema, r.plan), reps)
        dims, g = est.hints(r, p)
        out["library_same"] += (list(dims), g) == lh
        tl, (st_l, kl, rl) = _timed(engine, W.bind(schema, r.plan, dims, g), reps)
        bd, bg = sizing_hints(r, b, est.margin)
        tb, (st_b, kb, rb) = _timed(engine, W.bind(schema, r.plan, bd, bg), reps)
        same = (torch.equal(ke, kl) and torch.equal(re_, rl) and torch.equal(ke, kb)
                and torch.equal(re_, rb) and st_e["matches"] == st_l["matches"]
                and st_e["groups"] == st_l["groups"] and st_e["revenue"] == st_l["revenue"]
                and st_e["matches"] == r.stats["matches"])
        out["identical"] = out["identical"] and same
        out["t_exact"] += te
        out["t_learned"] += tl
        out["t_base"] += tb
        out["fallbacks"] += st_l["fallbacks"]
        out["base_fallbacks"] += st_b["fallbacks"]
    return out


def summarize(parts: list[dict]) -> dict:
    n = sum(p["n"] for p in parts)
    cat = lambda key, k: np.concatenate([p[key][k] for p in parts])
    t_e = sum(p["t_exact"] for p in parts)
    t_l = sum(p["t_learned"] for p in parts)
    t_b = sum(p["t_base"] for p in parts)
    inf = sum(p["infer_s"] for p in parts)
    return {
        "plans": n,
        "q_error": {k: percentiles(cat("qn", k)) for k in ("dim_pass", "matches", "groups")},
        "q_error_without_network": {k: percentiles(cat("qb", k))
                                    for k in ("dim_pass", "matches", "groups")},
        "fallback_rate": sum(p["fallbacks"] for p in parts) / n,
        "fallback_rate_without_network": sum(p["base_fallbacks"] for p in parts) / n,
        "exact_ms_per_plan": 1e3 * t_e / n,
        "learned_ms_per_plan": 1e3 * t_l / n,
        "without_network_ms_per_plan": 1e3 * t_b / n,
        "estimator_ms_per_plan": 1e3 * inf / n,
        "library_estimator_ms_per_plan": 1e3 * sum(p["library_s"] for p in parts) / n,
        "library_hints_same": sum(p["library_same"] for p in parts) / n,
        "exact_over_learned": t_e / t_l,
        "exact_over_learned_incl_estimator": t_e / (t_l + inf),
        "exact_over_without_network": t_e / t_b,
        "bit_identical": all(p["identical"] for p in parts),
    }


def experiment(engine, n_train_schemas: int, train_plans: int, n_eval_schemas: int,
               eval_plans: int, eval_fact_rows: int, seed: int, epochs: int = 40,
               reps: int = 3, log=None) -> dict:
    """Trains on plans over `n_train_schemas` seeded schemas and evaluates on
    held-out plans over fresh schemas of `eval_fact_rows` fact rows."""
    t0 = time.perf_counter()
    recs, _ = collect(engine, schema_specs(n_train_schemas, seed), train_plans, seed + 1,
                      log=log)
    t_rec = time.perf_counter() - t0
    est = Estimator(engine, seed=seed)
    est.fit(recs, epochs=epochs)
    est.margin = tune_margins(recs, est.predict(recs))
    if log:
        log(f"trained on {len(recs)} plans in {est.train_seconds:.2f} s "
            f"({est.train_steps} Adam steps), margins {est.margin}")
    parts = []
    for i, spec in enumerate(schema_specs(n_eval_schemas, seed + 7, eval_fact_rows)):
        held, schema = collect(engine, [spec], eval_plans, seed + 11 + i, keep_last=True)
        parts.append(evaluate(engine, est, schema, held, reps))
        del schema
        if log:
            log(f"evaluated held-out schema {i + 1}/{n_eval_schemas}")
    res = summarize(parts)
    res.update({"train_plans": len(recs), "train_rows": int(sum(len(r.x) for r in recs)),
                "record_seconds": t_rec, "train_seconds": est.train_seconds,
                "train_steps": est.train_steps, "final_loss": est.losses[-1],
                "first_loss": est.losses[0], "margin": est.margin})
    est.close()
    return res

"""Benchmark of the MI355X pipelines: one JSON line on stdout.

python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one pass of the hot path over one batch of synthetic, HBM-resident
input.  The headline workload is the filtered hash join + scalar aggregate
(60,000,000 fact rows, 2,000,000-row dimension); the other workloads are
reported alongside in `workloads`.  The CPU baseline is the C oracle (a
single-threaded restatement of the reference pipeline) timed on a bounded
sample on this host.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import pathlib
import sys
import time

import numpy as np

HBM_PEAK_GBPS = 8000.0
SEED = 42
LO = -(1 << 62)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workloads",
                    default="c2,c1,c3,q6,q14,q17,q18,c4,c5,minmax,sort,copy,tpch,ceb_count")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--knobs-out", default=None,
                    help="write the autotuners' picks as a knob table (bolap_tune_save)")
    ap.add_argument("--c4-train", type=int, default=20_000, help="C4 training plans")
    ap.add_argument("--c4-eval", type=int, default=500, help="C4 held-out plans")
    ap.add_argument("--c5-queries", type=int, default=512, help="C5 batch size")
    ap.add_argument("--tpch-orders", type=int, default=15_000_000,
                    help="orders of the TPC-H database (15,000,000 = SF10)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_steps(torch, step, steps, warmup):
    """Runs `warmup` untimed steps, then times exactly `steps` steps with HIP
    events on torch's current stream (the stream libbolap issues on)."""
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        step()
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps  # ms per step


# How the CPU baseline was built and run (filled by prepare_cpu_baseline and
# cpu_time; merged into every cpu_baseline entry by cpu_bl).
CPU_INFO = {"build": "-O3 -march=x86-64-v2 (prebuilt)"}
LAST_CPU = {}


def prepare_cpu_baseline():
    """Builds the oracle's C sources with -O3 -march=native on this host (the
    reference's own flags for its timed runs) into a temporary directory and
    points the oracle at it; keeps the prebuilt copy if that fails."""
    import tempfile
    from bespokeolap_artifacts_amd import build as B
    out = pathlib.Path(tempfile.mkdtemp(prefix="bolap_cpu_")) / "liboracle_native.so"
    try:
        B.build_oracle(out, march="native")
        os.environ["BOLAP_ORACLE_LIB"] = str(out)
        CPU_INFO["build"] = "-O3 -march=native (compiled on this host)"
    except Exception as e:  # noqa: BLE001
        log("native oracle build failed, using the prebuilt one:", e)
    try:
        model = next(ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo")
                     if ln.startswith("model name"))
    except (OSError, StopIteration):
        model = "unknown"
    CPU_INFO["cpu_model"] = model
    CPU_INFO["host_cpus"] = os.cpu_count()


def cpu_time(fn, budget_s):
    """Median seconds of fn over repetitions within about budget_s, on one
    pinned core (the reference pins its query process to one CPU,
    bespoke_tpch/query_impl.cpp:60-67,143)."""
    allowed = os.sched_getaffinity(0)
    core = 3 if 3 in allowed else min(allowed)
    os.sched_setaffinity(0, {core})
    try:
        fn()  # warm-up
        times = []
        t_all = time.perf_counter()
        while True:
            t0 = time.perf_counter()
            fn()
            times.append(time.perf_counter() - t0)
            if time.perf_counter() - t_all > budget_s or len(times) >= 50:
                break
    finally:
        os.sched_setaffinity(0, allowed)
    med = float(np.median(times))
    LAST_CPU.clear()
    LAST_CPU.update({"ms": med * 1e3, "pinned_cpu": core, "runs": len(times)})
    return med, len(times)


def cpu_bl(d: dict) -> dict:
    """A cpu_baseline entry with the timing of the last cpu_time call and how
    the baseline was built."""
    return {**d, **LAST_CPU, **CPU_INFO, "cores": 1}


# ------------------------------------------------------------------- C1
def bench_c1(eng, torch, args):
    from bespokeolap_artifacts_amd import native as N
    rows = 6_000_000
    cols = eng.gen_lineitem(SEED, rows)
    a = eng.q1_args(cols, LO, 2435, 3, 2)
    out = torch.empty((6, 8), dtype=torch.int64, device="cuda")
    step = lambda: eng.q1_scan_agg_async(a, out)  # noqa: E731
    ms = time_steps(torch, step, args.steps, args.warmup)
    alg_bytes = rows * (4 + 8 + 8 + 8 + 8 + 1 + 1)
    gbps = alg_bytes / (ms * 1e-3) / 1e9
    # parity of the timed configuration
    got = eng.q1_scan_agg(cols, LO, 2435, 3, 2)
    res = {"workload": "C1 filter-scan + 4-group aggregate, 6,000,000 rows",
           "ms_per_step": ms, "GBps": gbps, "frac_hbm": gbps / HBM_PEAK_GBPS,
           "algorithmic_bytes": alg_bytes, "rows_per_s": rows / (ms * 1e-3),
           "kernel": "q1_scan_agg_kernel"}
    if not args.no_cpu:
        import oracle
        host = {k: v.cpu().numpy() for k, v in cols.items()}
        want = oracle.q1_wide(host, LO, 2435, 3, 2)
        res["parity"] = [g.__dict__ for g in got] == want
        med, reps = cpu_time(lambda: oracle.q1_wide(host, LO, 2435, 3, 2), args.cpu_seconds)
        res["cpu_baseline"] = cpu_bl({"value": alg_bytes / med / 1e9, "unit": "GB/s", "cores": 1,
                               "kind": "port",
                               "sample": f"full 6,000,000-row table, median of {reps} runs"})
    return res


# ------------------------------------------------------------------- C2
def stage_ms(trace, name):
    v = [ms for n, ms, _ in trace if n == name and ms >= 0]
    return float(np.mean(v)) if v else None


def bench_c2(eng, torch, args):
    from bespokeolap_artifacts_amd import native as N
    dim_rows, fact_rows = 2_000_000, 60_000_000
    d = eng.gen_dim(SEED, dim_rows)
    f = eng.gen_fact(SEED, fact_rows, dim_rows)
    preds = [eng.pred(d["cat"], N.OP_RANGE, 7, 7),
             eng.pred(d["cont"], N.OP_IN_SET, values=[3, 17, 29])]
    a_exact = eng.join_sum_args(d["key"], preds, f["fk"], f["price"], f["disc"])
    first = eng.join_sum_run(a_exact)

    def timed(a):
        # Timed without stage tracing (its events would add to every call);
        # the stage breakdown comes from a separate traced pass.
        for _ in range(args.warmup):
            eng.join_sum_run(a)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = eng.join_sum_run(a)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        eng.set_trace(True)
        for _ in range(args.steps):
            eng.join_sum_run(a)
        torch.cuda.synchronize()
        trace = eng.read_trace()
        eng.set_trace(False)
        return ms, trace, r

    ms_exact, tr_exact, r_exact = timed(a_exact)
    # Estimate-sized path: the table is sized from a build-size hint (here
    # the exact count, the best an estimator can give), no count pass.
    a_hint = eng.join_sum_args(d["key"], preds, f["fk"], f["price"], f["disc"],
                               capacity_hint=first["build_rows"])
    ms_hint, tr_hint, r_hint = timed(a_hint)
    # The autotuner's pick for these arguments (probe loop, filter bits,
    # workgroups per CU), timed the same way; the knobs return to the
    # defaults afterwards so that the other workloads run untuned.
    tune_ms = eng.join_autotune(a_exact, reps=3)
    tuned = {k: eng.tune_get(k) for k in
             ("join.probe_variant", "join.lds_bloom_bits", "join.blocks_per_cu")}
    ms_tuned, _, r_tuned = timed(a_exact)
    for k in tuned:
        eng.tune_set(k, 0)
    probe_ms = stage_ms(tr_exact, "join_probe")
    matches = first["matches"]
    alg_bytes = fact_rows * 4 + dim_rows * 6 + matches * 16
    probe_bytes = fact_rows * 4 + matches * 16
    res = {"workload": "C2 filtered hash join + scalar SUM, 60,000,000 fact rows, "
                       "2,000,000-row dimension",
           "ms_per_step": ms_exact, "GBps": alg_bytes / (ms_exact * 1e-3) / 1e9,
           "algorithmic_bytes": alg_bytes,
           "probes_per_s": fact_rows / (ms_exact * 1e-3),
           "probe_kernel_ms": probe_ms,
           "probe_kernel_GBps": probe_bytes / (probe_ms * 1e-3) / 1e9 if probe_ms else None,
           "probe_kernel_probes_per_s": fact_rows / (probe_ms * 1e-3) if probe_ms else None,
           "stages_ms": {k: stage_ms(tr_exact, k) for k in
                         ("join_compact", "join_hbm_build", "join_probe", "join_merge")},
           "hinted_ms_per_step": ms_hint, "exact_over_hinted": ms_exact / ms_hint,
           "build_rows": first["build_rows"], "matches": matches,
           "placement": "LDS" if first["placement"] == N.PLACE_LDS else "HBM",
           "kernel": "probe_sum_k", "kernel_bytes_per_launch": probe_bytes,
           "autotuned": {"knobs": tuned, "ms_per_step": ms_tuned, "tuner_median_ms": tune_ms},
           "consistent": (r_exact == first and r_hint["revenue"] == first["revenue"]
                          and r_tuned == first)}
    if not args.no_cpu:
        import oracle
        hd = {k: v.cpu().numpy() for k, v in d.items()}
        hf = {k: v.cpu().numpy() for k, v in f.items()}
        op = [oracle.make_pred(hd["cat"], 0, 7, 7), oracle.make_pred(hd["cont"], 1, values=[3, 17, 29])]
        run = lambda: oracle.join_sum(hd["key"], op, hf["fk"], hf["price"], hf["disc"])  # noqa: E731
        want = run()
        res["parity"] = all(first[k] == want[k] for k in ("revenue", "build_rows", "matches"))
        med, reps = cpu_time(run, args.cpu_seconds)
        res["cpu_baseline"] = cpu_bl({"value": alg_bytes / med / 1e9, "unit": "GB/s", "cores": 1,
                               "kind": "port", "ms": med * 1e3,
                               "sample": f"full 60,000,000-row fact + 2,000,000-row dimension, "
                                         f"median of {reps} runs"})
    return res


# ------------------------------------------------------------------- Q6
def bench_q6(eng, torch, args):
    """TPC-H Q6 over 60,000,000 lineitem rows in the reference layout (ship
    i16, discount u8, quantity i16 x 100, price i32: 9 B/row): the flat scan,
    and the zone-map scan over the same rows clustered by the shard key."""
    rows = 60_000_000
    c = eng.gen_q6_lineitem(SEED, rows)
    p = eng.q6_params("1994-01-01", "0.06", "24")
    a_flat = eng.q6_args(c, p)
    first = eng.q6_run(a_flat)
    ms = time_steps(torch, lambda: eng.q6_run(a_flat), args.steps, args.warmup)
    cl = eng.q6_cluster(c)
    zones = eng.q6_zones(cl)
    a_zone = eng.q6_args(cl, p, zones)
    zres = eng.q6_run(a_zone)
    ms_zone = time_steps(torch, lambda: eng.q6_run(a_zone), args.steps, args.warmup)
    eng.set_trace(True)
    for _ in range(args.steps):
        eng.q6_run(a_flat)
        eng.q6_run(a_zone)
    torch.cuda.synchronize()
    tr = eng.read_trace()
    eng.set_trace(False)
    alg = rows * 9
    k_ms = stage_ms(tr, "q6_flat")
    res = {"workload": "Q6 filter-scan, 60,000,000 lineitem rows (reference layout, 9 B/row)",
           "ms_per_step": ms, "GBps": alg / (ms * 1e-3) / 1e9, "algorithmic_bytes": alg,
           "rows_per_s": rows / (ms * 1e-3), "kernel": "q6_flat_k",
           "kernel_ms": k_ms, "kernel_GBps": alg / (k_ms * 1e-3) / 1e9 if k_ms else None,
           "zoned_ms_per_step": ms_zone, "zoned_kernel_ms": stage_ms(tr, "q6_zoned"),
           "zoned_rows_scanned": zres["rows_scanned"], "rows_emitted": first["rows_emitted"],
           "consistent": zres["revenue"] == first["revenue"]}
    if not args.no_cpu:
        import oracle
        h = {k: v.cpu().numpy() for k, v in c.items()}
        want = oracle.q6_flat(h, p)
        res["parity"] = want == first
        med, reps = cpu_time(lambda: oracle.q6_flat(h, p), args.cpu_seconds)
        res["cpu_baseline"] = cpu_bl({"value": alg / med / 1e9, "unit": "GB/s", "cores": 1,
                               "kind": "port", "ms": med * 1e3,
                               "sample": f"full table, flat loop, median of {reps} runs"})
    return res


# ------------------------------------------------------------ Q14 / Q17
def bench_q14_q17(eng, torch, args, which):
    """Q14 (one-month promo revenue) and Q17 (small-quantity revenue) over
    60,000,000 lineitem rows joined to 2,000,000 parts (reference layout)."""
    from bespokeolap_artifacts_amd import native as N
    t = eng.gen_part_lineitem(SEED, 2_000_000, 60_000_000)
    rows = 60_000_000
    if which == "q14":
        lo, hi = eng.q14_window("1995-09-01")
        a = eng.q14_args(t, lo, hi)
        step = lambda: eng.q14_run(a)  # noqa: E731
        first = step()
        ms = time_steps(torch, step, args.steps, args.warmup)
        eng.set_trace(True)
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        tr = eng.read_trace()
        eng.set_trace(False)
        # ship for every row + part key and discounted price of the rows in
        # the window (one launch of the lineitem pass)
        alg = rows * 2 + first["rows_emitted"] * 8
        k_ms = stage_ms(tr, "q14_lineitem_scan")
        res = {"workload": "Q14 promo revenue, 60,000,000 lineitem rows, 2,000,000 parts, "
                           "one-month window",
               "ms_per_step": ms, "GBps": alg / (ms * 1e-3) / 1e9, "algorithmic_bytes": alg,
               "rows_per_s": rows / (ms * 1e-3), "rows_emitted": first["rows_emitted"],
               "kernel": "q14_flat_k", "kernel_ms": k_ms,
               "kernel_GBps": alg / (k_ms * 1e-3) / 1e9 if k_ms else None,
               "stages_ms": {k: stage_ms(tr, k) for k in ("q14_part_map", "q14_lineitem_scan",
                                                          "q14_part_scan")}}
    else:
        preds = [eng.pred(t["p_brand"], N.OP_RANGE, 3, 3),
                 eng.pred(t["p_container"], N.OP_RANGE, 7, 7)]
        a = eng.q17_args(t, preds)
        step = lambda: eng.q17_run(a)  # noqa: E731
        first = step()
        ms = time_steps(torch, step, args.steps, args.warmup)
        eng.set_trace(True)
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        tr = eng.read_trace()
        eng.set_trace(False)
        # whole call: part keys once, lineitem keys once, quantity and
        # price of the matching rows; first pass: keys + matched quantity
        alg = rows * 4 + first["matches"] * 6
        k_ms = stage_ms(tr, "q17_sum_scan")
        k_alg = rows * 4 + first["matches"] * 2
        res = {"workload": "Q17 small-quantity revenue, 60,000,000 lineitem rows, "
                           "2,000,000 parts (brand and container equality)",
               "ms_per_step": ms, "GBps": alg / (ms * 1e-3) / 1e9, "algorithmic_bytes": alg,
               "probes_per_s": rows / (ms * 1e-3), "parts": first["parts"],
               "matches": first["matches"], "kernel": "q17_probe_k<1>", "kernel_ms": k_ms,
               "kernel_GBps": k_alg / (k_ms * 1e-3) / 1e9 if k_ms else None,
               "stages_ms": {k: stage_ms(tr, k) for k in ("q17_part_scan", "q17_build",
                                                          "q17_sum_scan", "q17_filter_scan")}}
    if not args.no_cpu:
        import oracle
        h = {k: v.cpu().numpy() for k, v in t.items()}
        part = {"partkey": h["p_partkey"], "promo": h["p_promo"]}
        li = {"ship": h["l_ship"], "partkey": h["l_partkey"], "disc_price": h["l_disc_price"],
              "qty": h["l_qty"], "price": h["l_price"]}
        if which == "q14":
            run = lambda: oracle.q14(li, part, lo, hi)  # noqa: E731
        else:
            op = [oracle.make_pred(h["p_brand"], 0, 3, 3),
                  oracle.make_pred(h["p_container"], 0, 7, 7)]
            run = lambda: oracle.q17(h["p_partkey"], op, li)  # noqa: E731
        res["parity"] = run() == first
        med, reps = cpu_time(run, args.cpu_seconds)
        res["cpu_baseline"] = cpu_bl({"value": alg / med / 1e9, "unit": "GB/s", "cores": 1,
                               "kind": "port", "ms": med * 1e3,
                               "sample": f"full tables, median of {reps} runs"})
    return res


# ------------------------------------------------------------------ Q18
def bench_q18(eng, torch, args):
    """TPC-H Q18: SUM(quantity) per order over 60M lineitem rows sorted by
    order key (1-7 per order), 15M orders, 1.5M customers, threshold 300."""
    t = eng.gen_q18_tables(SEED, 1_500_000, 15_000_000)
    n = int(t["l_orderkey"].numel())
    a = eng.q18_args(t, 30_000)
    first, st = eng.q18_run(a)
    ms = time_steps(torch, lambda: eng.q18_run(a), args.steps, args.warmup)
    eng.set_trace(True)
    for _ in range(args.steps):
        eng.q18_run(a)
    torch.cuda.synchronize()
    tr = eng.read_trace()
    eng.set_trace(False)
    alg = n * 6
    k_ms = stage_ms(tr, "q18_sum")
    res = {"workload": f"Q18 group-by + threshold, {n:,} lineitem rows, 15,000,000 orders, "
                       "1,500,000 customers",
           "ms_per_step": ms, "GBps": alg / (ms * 1e-3) / 1e9, "algorithmic_bytes": alg,
           "rows_per_s": n / (ms * 1e-3), "kernel": "q18_sum_sorted_k", "kernel_ms": k_ms,
           "kernel_GBps": alg / (k_ms * 1e-3) / 1e9 if k_ms else None,
           "stages_ms": {k: stage_ms(tr, k) for k in ("q18_bounds", "q18_sum", "q18_select",
                                                      "q18_sort")},
           "result_rows": int(st["rows"]), "sorted_pass": st["sorted_pass"]}
    if not args.no_cpu:
        import oracle
        h = {k: v.cpu().numpy() for k, v in t.items()}
        want, _ = oracle.q18(h, 30_000)
        res["parity"] = [tuple(r) for r in want.tolist()] == [tuple(r) for r in first.tolist()]
        med, reps = cpu_time(lambda: oracle.q18(h, 30_000), args.cpu_seconds)
        res["cpu_baseline"] = cpu_bl({"value": alg / med / 1e9, "unit": "GB/s", "cores": 1,
                               "kind": "port", "ms": med * 1e3,
                               "sample": f"full tables, median of {reps} runs"})
    return res


# ------------------------------------- TPC-H order-range and join pipelines
def _tpch_upload(torch, h):
    """The reference-layout columns of tests/tpch_data.build() on the device
    (dictionary codes as bytes, unsigned columns as their signed views)."""
    u8 = ("l_shipmode", "l_shipinstruct", "o_orderpriority", "o_orderstatus", "p_type",
          "p_brand", "p_container")
    t = {}
    for k, v in h.items():
        if not isinstance(v, np.ndarray):
            t[k] = v
            continue
        a = v.astype(np.uint8) if k in u8 else v
        if a.dtype == np.uint32:
            a = a.view(np.int32)
        elif a.dtype == np.uint16:
            a = a.view(np.int16)
        t[k] = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t


def bench_tpch(eng, torch, args):
    """TPC-H Q2, Q4, Q5, Q7-Q13, Q15, Q16, Q19-Q22 over one
    seeded SF10-sized database in the reference layout (15,000,000 orders
    sorted by date with lineitem ranges, ~60,000,000 lineitem rows in order-key
    order, 2,000,000 parts, 8,000,000 partsupp rows, 1,500,000 customers,
    100,000 suppliers).  Per query: the whole call (HIP events, results copied
    back), the traced kernel stages, the algorithmic bytes (the reference
    layout's bytes of the columns the query reads over the rows it must
    visit: the date window's orders and their lineitem rows for the window
    queries, whole columns for the scans), and the oracle on one pinned core
    over the same columns."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from tests import tpch_data as T
    import oracle
    orders = args.tpch_orders
    log(f"tpch: generating {orders:,} orders")
    r = T.raw(SEED, orders, parts=max(orders // 15 * 2, 1000), custs=max(orders // 10, 1000),
              supps=max(orders // 150, 100), strings=False)
    h = T.build(r)
    del r
    t = _tpch_upload(torch, h)
    n_li, n_o = len(h["l_orderkey"]), len(h["o_orderkey"])
    n_p, n_ps, n_c, n_s = (len(h["p_partkey"]), len(h["ps_partkey"]), len(h["c_custkey"]),
                           len(h["s_suppkey"]))
    od = h["o_orderdate"]
    li_per_order = (h["r_end"].astype(np.int64) - h["r_start"]).astype(np.int64)

    def window(lo, hi):
        a, b = np.searchsorted(od, lo), np.searchsorted(od, hi)
        return int(b - a), int(li_per_order[a:b].sum())

    d = T.date_offset
    am = T.add_months
    nls = len(h["ls_dict"])
    rc = h["rf_dict"].index("R") * nls
    q7d = (d("1995-01-01"), d("1996-12-31"), d("1996-01-01"))
    inreg2 = (np.array(h["n_regionkey"]) == 2).astype(np.uint8)
    inreg1 = (np.array(h["n_regionkey"]) == 1).astype(np.uint8)
    tc8 = h["type_dict"].index("ECONOMY ANODIZED STEEL")
    p_ok = (h["p_name_words"] == T.COLORS.index("green")).any(axis=1).astype(np.uint8)
    p_ok_dev = torch.from_numpy(p_ok).cuda()
    masks = eng.string_masks(t["o_comment_chars"], t["o_comment_offsets"])
    tb16 = np.array([x.startswith("MEDIUM POLISHED") for x in h["type_dict"]], np.uint8)
    size16 = sum(1 << s_ for s_ in [49, 14, 23, 45, 19, 3, 36, 9])
    npk, nsk = int(h["p_partkey"].max()) + 1, int(h["s_suppkey"].max()) + 1
    b45 = h["brand_dict"].index("Brand#45")
    m12 = [h["shipmode_dict"].index(x) for x in ("MAIL", "SHIP")]
    hc = [h["priority_dict"].index(x) for x in ("1-URGENT", "2-HIGH")]
    from tests.test_tpch2_oracle import q19_groups
    g19 = q19_groups(h, ("Brand#12", "Brand#23", "Brand#34"), (1, 10, 20))
    m19 = (h["shipmode_dict"].index("AIR"), h["shipmode_dict"].index("REG AIR"),
           h["instruct_dict"].index("DELIVER IN PERSON"))
    fcode = h["status_dict"].index("F")
    codes22 = [13, 31, 23, 29, 30, 18, 17]
    lo94, hi94q, hi94y = d("1994-01-01"), am("1994-01-01", 3), am("1994-01-01", 12)
    w4, w10 = window(lo94, hi94q), window(lo94, hi94q)
    w5, w8 = window(lo94, hi94y), window(q7d[0], q7d[1] + 1)
    # Q2: BRASS parts of size 15 from EUROPE's suppliers; Q15: the first
    # quarter of 1996; Q20: forest parts, CANADA, 1994, the group table sized
    # by the reference's own estimate (rows x 1/7 x colour selectivity)
    tok2 = np.array([x.endswith("BRASS") for x in h["type_dict"]], np.uint8)
    nok2 = (np.array(h["n_regionkey"]) == 3).astype(np.uint8)
    lo96, hi96q = d("1996-01-01"), am("1996-01-01", 3)
    ship = h["l_shipdate"].astype(np.int32)
    w15 = int(((ship >= lo96) & (ship < hi96q)).sum())
    w20 = int(((ship >= lo94) & (ship < hi94y)).sum())
    del ship
    p_ok20 = (h["p_name_words"][:, 0] == T.COLORS.index("forest")).astype(np.uint8)
    p_ok20_dev = torch.from_numpy(p_ok20).cuda()
    hint20 = int(max(16.0, n_li / 7.0 * p_ok20.mean())) + 1
    smap = np.full(nsk, -1, np.int16)
    smap[h["s_suppkey"]] = h["s_nationkey"]
    smap_dev = torch.from_numpy(smap).cuda()
    # (name, gpu call, oracle call, algorithmic bytes, parity check of the results)
    Q = [
        ("q4", lambda: eng.q4(t, 5, lo94, hi94q), lambda: oracle.q4(h, lo94, hi94q, 5),
         w4[0] * 11 + w4[1] * 4, lambda g, o: g == o.tolist()),
        ("q5", lambda: eng.q5(t, inreg2, lo94, hi94y), lambda: oracle.q5(h, inreg2, lo94, hi94y),
         w5[0] * 12 + w5[1] * 6, lambda g, o: g == o.tolist()),
        ("q10", lambda: eng.q10(t, rc, nls, lo94, hi94q, 25, raw=True),
         lambda: oracle.q10(h, rc, nls, lo94, hi94q, 25), w10[0] * 14 + w10[1] * 5,
         lambda g, o: [tuple(x) for x in g.tolist()] == [tuple(x) for x in o.tolist()]),
        ("q12", lambda: eng.q12(t, m12[0], m12[1], lo94, hi94y, *hc),
         lambda: oracle.q12(h, m12[0], m12[1], lo94, hi94y, *hc), n_li * 7,
         lambda g, o: g == o),
        ("q19", lambda: eng.q19(t, g19, *m19), lambda: oracle.q19(h, g19, *m19),
         n_li * 8 + n_p * 10, lambda g, o: g == o),
        ("q7", lambda: eng.q7(t, 6, 7, *q7d), lambda: oracle.q7(h, 6, 7, *q7d),
         n_o * 10 + n_li * 10, lambda g, o: g == o.reshape(-1).tolist()),
        ("q8", lambda: eng.q8(t, inreg1, 2, tc8, *q7d),
         lambda: oracle.q8(h, inreg1, 2, tc8, *q7d), w8[0] * 14 + w8[1] * 13,
         lambda g, o: g == (o[0].tolist(), o[1].tolist())),
        ("q9", lambda: eng.q9(t, p_ok_dev, 25), lambda: oracle.q9(h, p_ok, 25),
         n_li * 20 + n_ps * 12 + n_o * 2 + n_p * 5,
         lambda g, o: g[0] == o[0] and np.array_equal(g[1], o[1]) and np.array_equal(g[2], o[2])),
        ("q13", lambda: eng.q13(t, "special", "requests", masks=masks),
         lambda: oracle.q13(h, "special", "requests"), n_o * 20,
         lambda g, o: g == (o[0], o[1].tolist())),
        ("q16", lambda: eng.q16(t, len(h["brand_dict"]), b45, tb16, size16, npk, nsk),
         lambda: oracle.q16(h, b45, tb16, size16, npk, nsk), n_p * 10 + n_ps * 8 + n_s * 5,
         lambda g, o: np.array_equal(g, o)),
        ("q21", lambda: eng.q21(t, fcode, 7, nsk), lambda: oracle.q21(h, fcode, 7, nsk),
         n_o * 13 + n_li * 8, lambda g, o: g == o.tolist()),
        ("q22", lambda: eng.q22(t, codes22, [(ord("A") << 8) | ord("B")]),
         lambda: oracle.q22(h, codes22, [(ord("A") << 8) | ord("B")]), n_c * 13 + n_o * 4,
         lambda g, o: g == (o[0].tolist(), o[1].tolist())),
        ("q2", lambda: eng.q2(t, tok2, 15, nok2), lambda: oracle.q2(h, tok2, 15, nok2[h["s_nationkey"]]),
         n_p * 5, lambda g, o: np.array_equal(g, o)),
        ("q11", lambda: eng.q11(t, 7, nsk, 0.0001 / 10), lambda: oracle.q11(h, 7, 0.0001 / 10),
         n_ps * 12 + n_s * 8,
         lambda g, o: np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1]) and g[2] == o[2]),
        ("q15", lambda: eng.q15(t, lo96, hi96q, nsk), lambda: oracle.q15(h, lo96, hi96q),
         n_li * 2 + w15 * 8 + n_s * 4, lambda g, o: g == o),
        ("q20", lambda: eng.q20(t, p_ok20_dev, npk - 1, lo94, hi94y, 3, smap_dev, nsk - 1,
                                capacity_hint=hint20),
         lambda: oracle.q20(h, p_ok20, lo94, hi94y, 3), n_li * 2 + w20 * 2 + n_ps * 8 + n_p * 5,
         lambda g, o: np.array_equal(g[0], o[0]) and g[1].qualified == o[1]),
    ]
    out = {"workload": f"TPC-H SF10 reference layout: {n_o:,} orders, {n_li:,} lineitem rows, "
                       f"{n_p:,} parts, {n_ps:,} partsupp, {n_c:,} customers, {n_s:,} suppliers",
           "queries": {}}
    for name, gpu, cpu, alg, same in Q:
        log("tpch", name)
        first = gpu()
        if isinstance(first, np.ndarray):
            first = first.copy()
        ms = time_steps(torch, gpu, args.steps, args.warmup)
        eng.set_trace(True)
        for _ in range(3):
            gpu()
        torch.cuda.synchronize()
        tr = eng.read_trace()
        eng.set_trace(False)
        stages = sorted({n for n, m, _ in tr if m >= 0})
        kernel_ms = sum(stage_ms(tr, s_) for s_ in stages)
        res = {"ms_per_step": ms, "algorithmic_bytes": int(alg), "GBps": alg / (ms * 1e-3) / 1e9,
               "kernel_ms": kernel_ms, "kernel_GBps": alg / (kernel_ms * 1e-3) / 1e9
               if kernel_ms else None,
               "stages_ms": {s_: stage_ms(tr, s_) for s_ in stages}}
        if not args.no_cpu:
            want = cpu()
            res["parity"] = bool(same(first, want))
            med, reps = cpu_time(cpu, min(args.cpu_seconds, 2.0))
            res["cpu_baseline"] = cpu_bl({"value": alg / med / 1e9, "unit": "GB/s",
                                          "kind": "port", "ms": med * 1e3,
                                          "sample": f"full SF10 columns, median of {reps} runs"})
            res["speedup_vs_cpu"] = med * 1e3 / ms
        out["queries"][name] = res
        log(name, json.dumps(res))
    del t
    torch.cuda.empty_cache()
    return out


# ---------------------------------------------- CEB factorised COUNT (2a)
def bench_ceb_count(eng, torch, args):
    """CEB 2a at the IMDB sizes of the C5 configuration: title 2,500,000 rows
    (kind 3% null, production_year 10% null), movie_info 15,000,000 (2 terms:
    (info_type, info) pairs), cast_info 36,000,000 (role 3% null, persons past
    the name table), movie_keyword 4,500,000, name 4,000,000 (gender 20%
    null), in the reference builder's row orders.  The whole bolap_ceb_count call over device-resident columns and
    masks; the CPU baseline is the reference's own run_q2a (oracle/_ref,
    compiled from its sources) on one pinned core over the same tables."""
    rng = np.random.default_rng(SEED)
    T_, P, K = 2_500_000, 4_000_000, 130_000
    MI, CI, MK = 15_000_000, 36_000_000, 4_500_000
    kinds = ["movie", "tv series", "tv movie", "video movie", "tv mini series", "video game",
             "episode"]
    roles = ["actor", "actress", "producer", "writer", "cinematographer", "composer",
             "costume designer", "director", "editor", "miscellaneous crew",
             "production designer", "guest"]
    vocab = [f"value {i}" for i in range(2000)]     # movie_info.info strings
    i32 = lambda a: np.ascontiguousarray(a, np.int32)

    def nulls(n, p):
        return rng.random(n) < p

    def movies(n):
        return i32(1 + np.floor(T_ * rng.random(n) ** 2))

    title = {"id": i32(np.arange(1, T_ + 1)), "kind": i32(rng.integers(1, 8, T_)),
             "kind_null": nulls(T_, 0.03), "year": i32(rng.integers(1900, 2020, T_)),
             "year_null": nulls(T_, 0.10)}
    gender = rng.integers(0, 5, P)                      # 0 m, 1 f, 2.. (20% null at 4)
    mi = {"movie": movies(MI), "type": i32(rng.integers(1, 9, MI)),
          "info": i32(rng.integers(0, len(vocab), MI)), "info_null": nulls(MI, 0.02)}
    ci = {"movie": movies(CI), "person": i32(rng.integers(1, P + 100_000, CI)),
          "role": i32(rng.integers(1, 13, CI)), "role_null": nulls(CI, 0.03)}
    mk = {"movie": movies(MK), "kw": i32(rng.integers(1, K + 1, MK))}
    # the reference builder's row orders (bespoke_ceb/builder_impl.cpp:432-452):
    # movie_info by (movie_id, info_type_id), cast_info by (movie_id, role_id),
    # movie_keyword by keyword_id
    o = np.lexsort((mi["type"], mi["movie"]))
    mi = {k: np.ascontiguousarray(v[o]) for k, v in mi.items()}
    o = np.lexsort((np.where(ci["role_null"], 0, ci["role"]), ci["movie"]))
    ci = {k: np.ascontiguousarray(v[o]) for k, v in ci.items()}
    o = np.argsort(mk["kw"], kind="stable")
    mk = {k: np.ascontiguousarray(v[o]) for k, v in mk.items()}
    # the query: ID1 (3), ID2 (8), INFO1 (10 values), INFO2 (25 values), KIND, ROLE, GENDER
    info1, info2 = list(range(0, 20, 2)), list(range(100, 125))
    kind_ok = np.zeros(9, np.uint8); kind_ok[[1, 7]] = 1                  # movie, episode
    role_ok = np.zeros(14, np.uint8); role_ok[[1, 2]] = 1                 # actor, actress
    person_ok = np.zeros(P + 1, np.uint8); person_ok[1:][gender == 1] = 1  # f
    type_ok1 = np.zeros(114, np.uint8); type_ok1[3] = 1
    type_ok2 = np.zeros(114, np.uint8); type_ok2[8] = 1
    v1 = np.zeros(len(vocab), np.uint8); v1[info1] = 1
    v2 = np.zeros(len(vocab), np.uint8); v2[info2] = 1
    kw_ok = np.ones(K + 1, np.uint8); kw_ok[0] = 0
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    dn = lambda m: dev(m.astype(np.uint8))
    P_ = eng.ceb_pred
    g = {k: dev(v) for k, v in {"tid": title["id"], "tk": title["kind"], "ty": title["year"],
                                 "mim": mi["movie"], "mit": mi["type"], "mii": mi["info"],
                                 "cim": ci["movie"], "cip": ci["person"], "cir": ci["role"],
                                 "mkm": mk["movie"], "mkk": mk["kw"]}.items()}
    gn = {"tk": dn(title["kind_null"]), "ty": dn(title["year_null"]),
          "mii": dn(mi["info_null"]), "cir": dn(ci["role_null"])}
    gm = {k: dev(v) for k, v in {"kind": kind_ok, "role": role_ok, "person": person_ok,
                                 "t1": type_ok1, "t2": type_ok2, "v1": v1, "v2": v2,
                                 "kw": kw_ok}.items()}
    a = eng.ceb_count_args(
        {"id": g["tid"], "year": g["ty"], "year_null": gn["ty"],
         "preds": [P_(g["tk"], gm["kind"], gn["tk"])]},
        [{"movie": g["mim"], "preds": [P_(g["mit"], gm["t1"]), P_(g["mii"], gm["v1"], gn["mii"])]},
         {"movie": g["mim"], "preds": [P_(g["mit"], gm["t2"]), P_(g["mii"], gm["v2"], gn["mii"])]},
         {"movie": g["cim"], "preds": [P_(g["cir"], gm["role"], gn["cir"]),
                                       P_(g["cip"], gm["person"])]},
         {"movie": g["mkm"], "preds": [P_(g["mkk"], gm["kw"])]}],
        1980, 2015, T_ + 1)
    first = eng.ceb_count(a)
    ms = time_steps(torch, lambda: eng.ceb_count(a), args.steps, args.warmup)
    eng.set_trace(True)
    for _ in range(3):
        eng.ceb_count(a)
    torch.cuda.synchronize()
    tr = eng.read_trace()
    eng.set_trace(False)
    term_ms = [m for n, m, _ in tr if n == "ceb_term" and m >= 0]
    # columns read once each: title id/kind/year (+2 null bytes), movie_info
    # movie/type/info (+1), cast_info movie/person/role (+1), movie_keyword 2
    alg = T_ * 14 + MI * 13 + CI * 13 + MK * 8
    res = {"workload": f"CEB 2a factorised COUNT: title {T_:,}, movie_info {MI:,} (x2 terms), "
                       f"cast_info {CI:,}, movie_keyword {MK:,}, name {P:,}",
           "ms_per_step": ms, "algorithmic_bytes": alg, "GBps": alg / (ms * 1e-3) / 1e9,
           "kernel": "term_k", "count": first["count"], "joined": first["joined"],
           "stages_ms": {n: stage_ms(tr, n) for n in ("ceb_title", "ceb_term", "ceb_product")},
           "term_ms_per_call": float(np.sum(term_ms) / 3) if term_ms else None}
    if not args.no_cpu:
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        from tests import ceb_ref as CR
        if CR.available():
            def strs(ids, table):
                enc = [s.encode() for s in table]
                lens = np.array([len(e) for e in enc], np.int64)
                blob = np.frombuffer(b"".join(enc), np.uint8)
                start = np.concatenate([[0], np.cumsum(lens)[:-1]])
                ln = lens[ids]
                off = np.zeros(len(ids) + 1, np.int64)
                np.cumsum(ln, out=off[1:])
                idx = np.repeat(start[ids] - off[:-1], ln) + np.arange(off[-1])
                return {"chars": blob[idx], "offsets": off}
            gid = np.where(gender == 4, 0, gender)
            gs = strs(gid, ["m", "f", "x", "y"])
            gs["null"] = gender == 4
            mis = strs(mi["info"], vocab)
            mis["null"] = mi["info_null"]
            tabs = {
                "kind_type": {"id": (i32(np.arange(1, 8)), None), "kind": kinds},
                "role_type": {"id": (i32(np.arange(1, 13)), None), "role": roles},
                "info_type": {"id": (i32(np.arange(1, 114)), None)},
                "keyword": {"id": (i32(np.arange(1, K + 1)), None)},
                "name": {"id": (i32(np.arange(1, P + 1)), None), "gender": gs},
                "title": {"id": (title["id"], None), "kind_id": (title["kind"], title["kind_null"]),
                          "production_year": (title["year"], title["year_null"])},
                "movie_info": {"movie_id": (mi["movie"], None), "info_type_id": (mi["type"], None),
                               "info": mis},
                "cast_info": {"movie_id": (ci["movie"], None), "person_id": (ci["person"], None),
                              "role_id": (ci["role"], ci["role_null"])},
                "movie_keyword": {"movie_id": (mk["movie"], None),
                                  "keyword_id": (mk["kw"], None)}}
            ref = CR.RefCeb(tabs)
            lst = lambda v: "(" + ", ".join(f"'{x}'" for x in v) + ")"
            line = (f"2a ('3') ('8') {lst(vocab[i] for i in info1)} {lst(vocab[i] for i in info2)} "
                    f"('movie', 'episode') ('actor', 'actress') ('f') \"2015\" \"1980\"")
            n_ref = []

            def run_ref():
                n, ms_ = ref.count_timed(line)
                n_ref.append(n)
            med, reps = cpu_time(run_ref, min(args.cpu_seconds, 4.0))
            res["parity"] = all(n == first["count"] for n in n_ref)
            res["cpu_baseline"] = cpu_bl({"value": alg / med / 1e9, "unit": "GB/s",
                                          "kind": "reference", "ms": med * 1e3,
                                          "sample": f"the same tables, run_q2a, median of "
                                                    f"{reps} runs"})
            res["speedup_vs_cpu"] = med * 1e3 / ms
            ref.close()
    return res


# ------------------------------------------- CEB q10a grouped MIN/MAX (a13)
def bench_minmax(eng, torch, args):
    """Grouped MIN/MAX(year) over 4,000,000 (person, movie) join rows expanded
    through 2,500,000 movies' info ranges (0-8 values each), nullable years
    (20% null), 10% of join rows without a name group; 200,000 name groups x
    2,000 infos (400 M slots: hashed) and 2,000 x 2,000 (4 M slots: dense)."""
    import ctypes as C
    from bespokeolap_artifacts_amd import native as N
    dev = "cuda"
    g = torch.Generator(device=dev).manual_seed(SEED)
    pairs, items, keys = 4_000_000, 2_500_000, 2000
    cnt = torch.randint(0, 9, (items,), device=dev, dtype=torch.int32, generator=g)
    off = (torch.cumsum(cnt.to(torch.int64), 0) - cnt).to(torch.int32)
    values = int(cnt.sum())
    rv = torch.randint(0, keys, (values,), device=dev, dtype=torch.int32, generator=g)
    has = (torch.rand(items, device=dev, generator=g) >= 0.2).to(torch.uint8)
    year = torch.randint(1880, 2026, (items,), device=dev, dtype=torch.int32, generator=g)
    item = torch.randint(0, items, (pairs,), device=dev, dtype=torch.int32, generator=g)
    drop = torch.rand(pairs, device=dev, generator=g) < 0.1
    out = {}
    for label, groups in (("hashed", 200_000), ("dense", 2000)):
        grp = torch.randint(0, groups, (pairs,), device=dev, dtype=torch.int32, generator=g)
        grp[drop] = -1
        st, rows = eng.group_minmax(grp, item, off, cnt, has, year, rv, groups, keys)
        cap = int(st["groups"])
        buf = torch.empty((cap, 5), dtype=torch.int32, device=dev)
        a = N.MinmaxArgs(pairs, grp.data_ptr(), item.data_ptr(), items, off.data_ptr(),
                         cnt.data_ptr(), has.data_ptr(), year.data_ptr(), values,
                         rv.data_ptr(), groups, keys, 0)
        n = C.c_int64(0)

        def step():
            N.check(eng.lib.bolap_group_minmax(eng.ctx, C.byref(a), buf.data_ptr(), cap,
                                               C.byref(n), None))
        ms = time_steps(torch, step, args.steps, args.warmup)
        exp = int(st["expanded_rows"])
        # join rows (group, item) + per join row its item's offset, count, has,
        # year + the expanded range values
        alg = pairs * 8 + pairs * 13 + exp * 4
        r = {"ms_per_step": ms, "expanded_rows": exp, "groups": cap,
             "placement": "dense" if st["placement"] == N.MINMAX_DENSE else "hashed",
             "expanded_rows_per_s": exp / (ms * 1e-3), "algorithmic_bytes": alg,
             "GBps_alg": alg / (ms * 1e-3) / 1e9}
        if not args.no_cpu:
            import oracle
            h = [x.cpu().numpy() for x in (grp, item, off, cnt, has, year, rv)]
            want = oracle.group_minmax(*h, groups, keys)
            gotr = rows.cpu().numpy()
            r["parity"] = bool(len(want) == len(gotr) and all(
                np.array_equal(want[f], gotr[:, i])
                for i, f in enumerate(("group", "key", "has", "min", "max"))))
            med, reps = cpu_time(lambda: oracle.group_minmax(*h, groups, keys),
                                 args.cpu_seconds)
            r["cpu_baseline"] = cpu_bl({"value": exp / med, "unit": "expanded rows/s", "cores": 1,
                                 "kind": "port", "ms": med * 1e3,
                                 "sample": f"full input, median of {reps} runs"})
        out[label] = r
    return {"workload": "CEB q10a grouped MIN/MAX(year), 4,000,000 join rows, 2,500,000 "
                        "movies' info ranges", "ms_per_step": out["hashed"]["ms_per_step"],
            "variants": out}


# ------------------------------------------- CEB build sort order (a10)
def bench_sort(eng, torch, args):
    """cast_info's build order: 36,000,000 rows sorted by (movie_id, role_id),
    then both columns gathered into that order."""
    dev = "cuda"
    g = torch.Generator(device=dev).manual_seed(SEED)
    n = 36_000_000
    movie = torch.randint(1, 2_500_001, (n,), device=dev, dtype=torch.int32, generator=g)
    role = torch.randint(1, 13, (n,), device=dev, dtype=torch.int32, generator=g)

    def step():
        order = eng.sort_order(movie, role)
        eng.gather_rows(movie, order)
        eng.gather_rows(role, order)
    ms = time_steps(torch, step, args.steps, args.warmup)
    # The autotuner's digit width for keys of this shape (26 bits: movie id
    # offsets and role ids packed), then the leg again with it.
    keys = (torch.randint(0, 1 << 26, (n,), device=dev, dtype=torch.int64, generator=g))
    vals = torch.arange(n, device=dev, dtype=torch.int32)
    tune_ms = eng.sort_autotune(keys, vals, 26, reps=3)
    width = eng.tune_get("radix.digit_bits")
    del keys, vals
    ms_tuned = time_steps(torch, step, args.steps, args.warmup)
    eng.tune_set("radix.digit_bits", 0)
    res = {"workload": "CEB build: cast_info 36,000,000 rows sorted by (movie_id, role_id) "
                       "+ gather of both columns", "ms_per_step": ms,
           "rows_per_s": n / (ms * 1e-3),
           "autotuned": {"knobs": {"radix.digit_bits": width}, "ms_per_step": ms_tuned,
                         "tuner_median_sort_ms": tune_ms}}
    if not args.no_cpu:
        import oracle
        m = 2_000_000
        hm, hr = movie[:m].cpu().numpy(), role[:m].cpu().numpy()
        med, reps = cpu_time(lambda: oracle.sort_order_i32(hm, hr), args.cpu_seconds)
        res["cpu_baseline"] = cpu_bl({"value": m / med, "unit": "rows/s", "cores": 1, "kind": "port",
                               "ms": med * 1e3,
                               "sample": f"first 2,000,000 rows (order only), median of "
                                         f"{reps} runs"})
    return res


# ------------------------------------------------------------ C4 estimator
def bench_c4(eng, torch, args):
    """Learned-sized vs exact-sized star plans: 20,000 training plans over 16
    seeded schemas (1M-60M fact rows, Zipf s in {0, .5, 1, 1.5}), 500
    held-out plans over 4 fresh schemas of 60M fact rows."""
    from bespokeolap_artifacts_amd import estimator as E
    res = E.experiment(eng, n_train_schemas=16, train_plans=args.c4_train // 16,
                       n_eval_schemas=4, eval_plans=args.c4_eval // 4,
                       eval_fact_rows=60_000_000, seed=SEED, reps=3, log=log)
    res["workload"] = (f"C4 learned vs exact sizing, {res['train_plans']:,} training plans, "
                       f"{res['plans']} held-out plans at 60,000,000 fact rows")
    res["ms_per_step"] = res["learned_ms_per_plan"]
    return res


# ------------------------------------------------------------ C5 batch
def bench_c5(eng, torch, args):
    """512 parameterised CEB-style queries (8 templates, 1-5 joins) over six
    tables of 36M/15M/4.5M/2.5M/130k/100 rows: one at a time vs one batch
    (shared scans, hipGraph replay, autotuned stage configuration)."""
    from bespokeolap_artifacts_amd import ceb as CB
    s = CB.CebSchema(eng, seed=SEED)
    qs = CB.random_queries(s, args.c5_queries, seed=SEED)
    bound = [CB.bind(s, q) for q in qs]
    for b in bound[:8]:
        CB.run_one(eng, s, b)
    lat = []
    single = []
    t0 = time.perf_counter()
    for b in bound:
        t1 = time.perf_counter()
        single.append(CB.run_one(eng, s, b))
        lat.append(time.perf_counter() - t1)
    t_single = time.perf_counter() - t0
    tune_ms = CB.autotune(eng, s, bound)
    batch = CB.run_batch(eng, s, bound)
    reps = max(1, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(reps):
        batch = CB.run_batch(eng, s, bound)
    t_batch = (time.perf_counter() - t0) / reps
    n = len(bound)
    lat_ms = np.array(lat) * 1e3
    by_template = {}
    for q, l in zip(qs, lat_ms):
        by_template.setdefault(q["template"], []).append(l)
    return {"workload": f"C5 {n} CEB-style queries, 8 templates, 1-5 joins, "
                        "cast_info 36M / movie_info 15M / movie_keyword 4.5M / title 2.5M",
            "queries": n, "ms_per_step": t_batch * 1e3,
            "batched_queries_per_s": n / t_batch, "single_queries_per_s": n / t_single,
            "batched_over_single": t_single / t_batch,
            "batched_latency_ms": {"p50": t_batch * 1e3, "p99": t_batch * 1e3},
            "single_latency_ms": {"p50": float(np.percentile(lat_ms, 50)),
                                  "p99": float(np.percentile(lat_ms, 99))},
            "single_ms_by_template": {str(k): float(np.mean(v)) for k, v in
                                      sorted(by_template.items())},
            "autotuned": {"blocks_per_cu": eng.tune_get("ceb.blocks_per_cu"),
                          "qchunk": eng.tune_get("ceb.qchunk"), "batch_ms": tune_ms},
            "bit_identical": batch == single, "kernel": "ceb_fact_k"}


# ------------------------------------------------------------- copy ceiling
def bench_copy(eng, torch, args):
    """Plain device copy of 1 GiB: the measured bandwidth ceiling, and the
    known-size kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters."""
    n = 1 << 30
    a = torch.empty(n, dtype=torch.uint8, device="cuda")
    a.fill_(1)
    b = torch.empty_like(a)
    ms = time_steps(torch, lambda: b.copy_(a), args.steps, args.warmup)
    return {"workload": "device copy of 1 GiB (read 1 GiB + write 1 GiB)", "ms_per_step": ms,
            "GBps": 2 * n / (ms * 1e-3) / 1e9, "copy_bytes": n, "kernel": "copy"}


# ------------------------------------------------------------------- C3
def bench_c3(eng, torch, args):
    cust_rows, order_rows = 1_500_000, 15_000_000
    t = eng.gen_chain(SEED, cust_rows, order_rows)
    fact_rows = int(t["fact_order"].numel())
    # key ranges as a built database keeps them (per-column statistics, taken
    # once outside the timing); the call then skips its own min/max pass
    kr = (int(t["cust_key"].min()), int(t["cust_key"].max()), int(t["order_key"].min()),
          int(t["order_key"].max()))
    a_exact = eng.chain_args(t, 1, 1155, key_ranges=kr)
    first, st = eng.chain_topk(a_exact, 10)

    def timed(a):
        for _ in range(args.warmup):
            eng.chain_topk(a, 10)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r, s = eng.chain_topk(a, 10)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        eng.set_trace(True)  # stage breakdown from a separate traced pass
        for _ in range(args.steps):
            eng.chain_topk(a, 10)
        torch.cuda.synchronize()
        trace = eng.read_trace()
        eng.set_trace(False)
        return ms, trace, r

    ms_exact, tr, r_exact = timed(a_exact)
    a_hint = eng.chain_args(t, 1, 1155, capacity_hint=st["order_pass"], key_ranges=kr)
    ms_hint, tr_h, r_hint = timed(a_hint)
    ms_own_ranges, _, r_own = timed(eng.chain_args(t, 1, 1155))
    alg_bytes = fact_rows * (4 + 8 + 8) + order_rows * (4 + 4) + cust_rows * 1
    probe_ms = stage_ms(tr, "chain_probe")
    stages = {}
    for n, ms, _ in tr:
        if ms >= 0:
            stages.setdefault(n, []).append(ms)
    res = {"workload": "C3 chain join + ~1.44M-group GROUP BY + top-10, 1,500,000 customers, "
                       f"15,000,000 orders, {fact_rows:,} fact rows",
           "ms_per_step": ms_exact, "GBps": alg_bytes / (ms_exact * 1e-3) / 1e9,
           "algorithmic_bytes": alg_bytes, "hinted_ms_per_step": ms_hint,
           "exact_over_hinted": ms_exact / ms_hint,
           "probe_kernel_ms": probe_ms,
           "probe_kernel_GBps": fact_rows * 20 / (probe_ms * 1e-3) / 1e9 if probe_ms else None,
           "stages_ms": {k: float(np.mean(v)) for k, v in stages.items()},
           "groups": st["groups"], "fact_matches": st["fact_matches"],
           "own_key_ranges_ms_per_step": ms_own_ranges,
           "kernel": "fact_probe_k",
           "consistent": r_exact == first and r_hint == first and r_own == first}
    # The reference's own layout: each order's fact rows as a contiguous range
    # (lineitem_ranges, made on the device at build time, outside the timing).
    _, start, end = eng.order_ranges(t["order_key"], t["fact_order"], order_rows + 1)
    a_rng = eng.chain_args(t, 1, 1155, ranges=(start, end), key_ranges=kr, fact_order_sorted=True)
    ms_rng, tr_r, r_rng = timed(a_rng)
    _, st_r = eng.chain_topk(a_rng, 10)
    k_ms = stage_ms(tr_r, "chain_ranges")
    # orders' day and customer for every row; range, key and the price and
    # discount of the fact rows for the passing orders
    k_bytes = order_rows * 8 + st_r["order_pass"] * 12 + st_r["fact_matches"] * 16
    stages_r = {}
    for n, ms, _ in tr_r:
        if ms >= 0:
            stages_r.setdefault(n, []).append(ms)
    res["by_ranges"] = {"ms_per_step": ms_rng, "kernel": "range_groups_k", "kernel_ms": k_ms,
                        "kernel_bytes": k_bytes,
                        "kernel_GBps": k_bytes / (k_ms * 1e-3) / 1e9 if k_ms else None,
                        "stages_ms": {k: float(np.mean(v)) for k, v in stages_r.items()},
                        "consistent": r_rng == first and st_r["groups"] == st["groups"]}
    if not args.no_cpu:
        import oracle
        h = {k: v.cpu().numpy() for k, v in t.items()}
        # run_q3's own algorithm over the reference layout, built outside the
        # timing as the reference's build() does: orders sorted by day, each
        # order's fact rows as a contiguous range; the timed part is the
        # date-prefix walk over the ranges and the full sort of the groups.
        perm = np.argsort(h["order_day"], kind="stable")
        o_sorted = {k: np.ascontiguousarray(h[k][perm])
                    for k in ("order_key", "order_day", "order_cust")}
        r0 = np.ascontiguousarray(h["fact_offsets"][:-1][perm].astype(np.uint32))
        r1 = np.ascontiguousarray(h["fact_offsets"][1:][perm].astype(np.uint32))

        def run():
            return oracle.q3_ranges(h, 1, 1155, o_sorted, r0, r1)[:10]
        want = run()
        res["parity"] = [tuple(x) for x in want.tolist()] == [tuple(x) for x in first]
        med, reps = cpu_time(run, args.cpu_seconds)
        res["cpu_baseline"] = cpu_bl({"value": alg_bytes / med / 1e9, "unit": "GB/s", "cores": 1,
                               "kind": "port", "ms": med * 1e3,
                               "sample": f"full tables, median of {reps} runs"})
    return res


# HBM bytes per launch (read + written) of each workload's dominant kernel,
# measured by two rocprofv3 runs (FETCH_SIZE, then WRITE_SIZE; summaries under
# profiles/).  A 1 GiB device copy reads FETCH_SIZE = 524,306 KiB and
# WRITE_SIZE = 1,048,576 KiB, so FETCH_SIZE is scaled by 2.0 and WRITE_SIZE by 1.0.
MEASURED_TRAFFIC = {
    "probe_sum_k": 248.3e6 + 0.0e6,
    "q1_scan_agg_kernel": 228.3e6 + 0.1e6,
    "fact_probe_k": 622.4e6 + 59.0e6,
    "q6_flat_k": 540.6e6 + 0.0e6,
    "q18_sum_sorted_k": 382.8e6 + 112.9e6,
    "range_groups_k": 752.1e6 + 23.3e6,
    "q14_flat_k": 327.4e6 + 0.1e6,
    "q17_probe_k<1>": 248.0e6 + 4.3e6,
}


# L2 hit rate (TCC_HIT / (TCC_HIT + TCC_MISS)) of each join workload's
# dominant kernel, from a rocprofv3 counter pass (summary under profiles/).
MEASURED_L2_HIT = {
    "probe_sum_k": 0.039,
    "fact_probe_k": 0.455,
    "range_groups_k": 0.557,
    "q14_flat_k": 0.106,
    "q17_probe_k<1>": 0.043,
    "ceb_fact_k": 0.179,
}


def roofline(head):
    """Roofline of the workload's dominant kernel (algorithmic bytes per
    launch / measured kernel time)."""
    gbps = head.get("probe_kernel_GBps") or head.get("kernel_GBps") or head.get("GBps")
    if gbps is None:  # a run of workloads without a bandwidth figure (C4, C5)
        return None
    return {"bound": "hbm", "achieved": gbps, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": gbps / HBM_PEAK_GBPS, "traffic": MEASURED_TRAFFIC.get(head.get("kernel")),
            "kernel": head.get("kernel")}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); "
              "multi-GPU sharding is a design note, not built", file=sys.stderr)
        sys.exit(2)
    import torch
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from bespokeolap_artifacts_amd.engine import Engine
    if not args.no_cpu:
        prepare_cpu_baseline()
    eng = Engine(0)
    info = eng.device_info()
    log("device", info)
    results = {}
    for w in args.workloads.split(","):
        log("workload", w)
        results[w] = {"c1": bench_c1, "c2": bench_c2, "c3": bench_c3, "q6": bench_q6,
                      "q14": lambda e, t, a: bench_q14_q17(e, t, a, "q14"),
                      "q17": lambda e, t, a: bench_q14_q17(e, t, a, "q17"),
                      "q18": bench_q18, "c4": bench_c4, "c5": bench_c5,
                      "minmax": bench_minmax, "sort": bench_sort, "tpch": bench_tpch,
                      "ceb_count": bench_ceb_count,
                      "copy": bench_copy}[w](eng, torch, args)
        log(json.dumps(results[w]))
    for r in results.values():
        if r.get("kernel") in MEASURED_L2_HIT:
            r["l2_hit_rate"] = MEASURED_L2_HIT[r["kernel"]]
        br = r.get("by_ranges")
        if isinstance(br, dict) and br.get("kernel") in MEASURED_L2_HIT:
            br["l2_hit_rate"] = MEASURED_L2_HIT[br["kernel"]]
    # The autotuners' picks, as one knob table (bolap_tune_load reads it back).
    tuned_knobs = {}
    for r in results.values():
        at = r.get("autotuned")
        if isinstance(at, dict):
            if isinstance(at.get("knobs"), dict):
                tuned_knobs.update(at["knobs"])
            for k in ("blocks_per_cu", "qchunk"):
                if k in at:
                    tuned_knobs["ceb." + k] = at[k]
    if args.knobs_out and tuned_knobs:
        for k, v in tuned_knobs.items():
            eng.tune_set(k, int(v))
        eng.tune_save(args.knobs_out)
    # The headline is the first workload with a bandwidth figure (C2 by default).
    head = next((results[w] for w in results if "GBps" in results[w]),
                next(iter(results.values())))
    line = {
        "metric": "scan GB/s (% of HBM peak) + join probes/s; learned- vs exact-sized time",
        "value": head.get("GBps"), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head.get("ms_per_step"),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int64", "data": "synthetic (seeded SplitMix64 generator, seed 42)",
        "config": {"workload": head.get("workload"), "parallelism": "single GPU"},
        "roofline": roofline(head),
        "cpu_baseline": head.get("cpu_baseline"),
        "parity": head.get("parity"),
        "workloads": results,
        "tuned_knobs": tuned_knobs,
        "device": info,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

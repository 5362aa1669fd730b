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
    ap.add_argument("--workloads", default="c2,c1,c3,q6,q14,q17,q18,c4,c5,minmax,sort,copy")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--c4-train", type=int, default=20_000, help="C4 training plans")
    ap.add_argument("--c4-eval", type=int, default=500, help="C4 held-out plans")
    ap.add_argument("--c5-queries", type=int, default=512, help="C5 batch size")
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
           "consistent": r_exact == first and r_hint["revenue"] == first["revenue"]}
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
    res = {"workload": "CEB build: cast_info 36,000,000 rows sorted by (movie_id, role_id) "
                       "+ gather of both columns", "ms_per_step": ms,
           "rows_per_s": n / (ms * 1e-3)}
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
    "fact_probe_k": 612.6e6 + 91.5e6,
    "q6_flat_k": 540.6e6 + 0.0e6,
    "q18_sum_sorted_k": 378.9e6 + 107.7e6,
    "range_groups_k": 749.9e6 + 23.3e6,
    "q14_flat_k": 327.5e6 + 0.1e6,
    "q17_probe_k<1>": 248.1e6 + 4.3e6,
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
                      "minmax": bench_minmax, "sort": bench_sort,
                      "copy": bench_copy}[w](eng, torch, args)
        log(json.dumps(results[w]))
    for r in results.values():
        if r.get("kernel") in MEASURED_L2_HIT:
            r["l2_hit_rate"] = MEASURED_L2_HIT[r["kernel"]]
        br = r.get("by_ranges")
        if isinstance(br, dict) and br.get("kernel") in MEASURED_L2_HIT:
            br["l2_hit_rate"] = MEASURED_L2_HIT[br["kernel"]]
    # The headline is the first workload with a bandwidth figure (C2 by default).
    head = next((results[w] for w in results if "GBps" in results[w]),
                next(iter(results.values())))
    line = {
        "metric": "scan GB/s (% of HBM peak) + join probes/s; learned- vs exact-sized time",
        "value": head.get("GBps"), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int64", "data": "synthetic (seeded SplitMix64 generator, seed 42)",
        "config": {"workload": head["workload"], "parallelism": "single GPU"},
        "roofline": roofline(head),
        "cpu_baseline": head.get("cpu_baseline"),
        "parity": head.get("parity"),
        "workloads": results,
        "device": info,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

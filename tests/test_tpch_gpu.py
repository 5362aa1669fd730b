"""Q6, Q18 and the build derivations on the GPU vs the oracle (bit-exact)."""
import numpy as np
import pytest

import oracle
from bespokeolap_artifacts_amd import native as N

pytestmark = pytest.mark.gpu

SEED = 42
Q6_CASES = [("1994-01-01", "0.06", "24"), ("1996-02-29", "0.00", "24.5"),
            ("1992-01-01", "0.10", "50"), ("1998-06-15", "0.05", "1")]


def host(d: dict) -> dict:
    return {k: v.cpu().numpy() for k, v in d.items()}


# --------------------------------------------------------------------- Q6
@pytest.mark.parametrize("rows", [0, 1, 7, 8, 9, 2047, 2048, 2049, 100_003, 6_000_000])
def test_q6_flat_and_zoned_match_oracle(engine, rows):
    c = engine.gen_q6_lineitem(SEED, rows)
    h = host(c)
    zones = engine.q6_zones(c)
    for case in Q6_CASES:
        p = engine.q6_params(*case)
        want = oracle.q6_flat(h, p)
        got = engine.q6(c, p)
        assert got == want, (rows, case, got, want)
        z = engine.q6(c, p, zones)
        assert (z["revenue"], z["rows_emitted"]) == (want["revenue"], want["rows_emitted"])
        assert z["rows_scanned"] <= rows


def test_q6_cluster_is_a_stable_sort_on_the_shard_key(engine):
    rows = 300_001
    c = engine.gen_q6_lineitem(SEED, rows)
    h = host(c)
    cl = engine.q6_cluster(c, with_perm=True)
    days = h["ship"].astype(np.int64) + 8035
    dt = days.astype("datetime64[D]")
    year = dt.astype("datetime64[Y]").astype(np.int64) + 1970
    month = dt.astype("datetime64[M]").astype(np.int64) % 12
    key = ((year - 1900) * 12 + month) * (1 << 18) + h["disc"].astype(np.int64) * (1 << 10) + \
        (h["qty"].astype(np.int64) // 100 + 512)
    want = np.argsort(key, kind="stable")
    perm = cl["perm"].cpu().numpy()
    assert np.array_equal(perm, want)
    hc = host({k: cl[k] for k in ("ship", "disc", "qty", "price")})
    for k in hc:
        assert np.array_equal(hc[k], h[k][want]), k


@pytest.mark.parametrize("rows", [2048 * 3 + 5, 6_000_000, 60_000_000])
def test_q6_zoned_on_clustered_columns(engine, rows):
    c = engine.gen_q6_lineitem(SEED + 1, rows)
    h = host(c)
    cl = engine.q6_cluster(c)
    zones = engine.q6_zones(cl)
    for case in Q6_CASES:
        p = engine.q6_params(*case)
        want = oracle.q6_flat(h, p)
        got = engine.q6(cl, p, zones)
        assert (got["revenue"], got["rows_emitted"]) == (want["revenue"], want["rows_emitted"])
        flat = engine.q6(cl, p)
        assert flat == want
    # one year, three discounts, 23 quantities: the zones skip most blocks
    got = engine.q6(cl, engine.q6_params("1994-01-01", "0.06", "24"), zones)
    assert got["rows_scanned"] < 0.05 * rows + 8 * 2048


def test_q6_rejects_misaligned_columns(engine):
    c = engine.gen_q6_lineitem(SEED, 1000)
    bad = {k: v[1:] for k, v in c.items()}
    with pytest.raises(N.BolapError):
        engine.q6(bad, engine.q6_params(*Q6_CASES[0]))


# -------------------------------------------------------------------- Q18
def q18_frame(rows: np.ndarray) -> list:
    return [tuple(int(r[k]) for k in rows.dtype.names) for r in rows]


def run_q18(engine, t, threshold, sorted_hint=True, max_rows=1 << 16):
    got, st = engine.q18(t, threshold, orderkey_sorted=sorted_hint, max_rows=max_rows)
    want, dense = oracle.q18(host(t), threshold)
    assert st["dense_path"] == int(dense)
    assert q18_frame(got) == q18_frame(want), (threshold, st, len(got), len(want))
    return got, st


@pytest.mark.parametrize("cust,orders", [(100, 1000), (20_000, 300_000)])
@pytest.mark.parametrize("threshold", [-1, 15_000, 25_000, 32_000, 10 ** 6])
def test_q18_dense_keys_sorted_lineitem(engine, cust, orders, threshold):
    t = engine.gen_q18_tables(SEED, cust, orders)
    got, st = run_q18(engine, t, threshold, max_rows=orders + 1)
    assert st["sorted_pass"] == 1 and st["fallbacks"] == 0


def test_q18_long_runs_cross_the_lookahead(engine):
    # Eight orders' rows merged under one key: runs of about 32 rows, past
    # the 16-row window a lane sums in registers.
    import torch
    t = engine.gen_q18_tables(SEED, 5000, 100_000)
    keys, inv = torch.unique_consecutive(t["l_orderkey"], return_inverse=True)
    first = torch.arange(keys.numel(), device=keys.device) // 8 * 8
    t["l_orderkey"] = keys[first][inv].contiguous()
    for thr in (50_000, 100_000, 140_000):
        got, st = run_q18(engine, t, thr)
        assert st["sorted_pass"] == 1 and len(got) > 0


def test_q18_date_sorted_orders_take_the_max_key_path(engine):
    import torch
    t = engine.gen_q18_tables(SEED, 20_000, 300_000)
    order = engine.date_order(t["o_orderdate"]).long()
    for k in ("o_orderkey", "o_custkey", "o_orderdate", "o_totalprice"):
        t[k] = t[k][order].contiguous()
    got, st = run_q18(engine, t, 25_000)
    assert st["dense_path"] == 0 and len(got) > 0
    torch.cuda.synchronize()


def test_q18_unsorted_lineitem_falls_back_to_atomics(engine):
    import torch
    t = engine.gen_q18_tables(SEED, 5000, 100_000)
    g = torch.Generator(device="cpu").manual_seed(7)
    perm = torch.randperm(t["l_orderkey"].numel(), generator=g).cuda()
    t["l_orderkey"] = t["l_orderkey"][perm].contiguous()
    t["l_qty"] = t["l_qty"][perm].contiguous()
    _, st = run_q18(engine, t, 20_000, sorted_hint=True)
    assert st["fallbacks"] == 1 and st["sorted_pass"] == 0
    _, st = run_q18(engine, t, 20_000, sorted_hint=False)
    assert st["fallbacks"] == 0 and st["sorted_pass"] == 0


def test_q18_duplicate_customer_keys_and_missing_customers(engine):
    import torch
    t = engine.gen_q18_tables(SEED, 3000, 50_000)
    ck = t["c_custkey"].clone()
    ck[::7] = ck[1::7][: ck[::7].numel()] if ck.numel() > 7 else ck[::7]
    ck[5] = 10 ** 7  # a key no order refers to
    perm = torch.randperm(ck.numel(), generator=torch.Generator().manual_seed(3)).cuda()
    t["c_custkey"] = ck[perm].contiguous()
    run_q18(engine, t, 15_000)


def test_q18_full_size(engine):
    t = engine.gen_q18_tables(SEED, 1_500_000, 15_000_000)
    got, st = run_q18(engine, t, 30_000)
    assert 100 < len(got) < 10_000 and st["sorted_pass"] == 1


def test_q18_more_rows_than_max_rows_is_an_error(engine):
    t = engine.gen_q18_tables(SEED, 100, 1000)
    with pytest.raises(N.BolapError):
        engine.q18(t, -1, max_rows=10)


def test_q18_empty_tables(engine):
    import torch
    t = engine.gen_q18_tables(SEED, 100, 1000)
    e = {k: v[:0] for k, v in t.items()}
    got, st = engine.q18(e, 0)
    assert len(got) == 0
    t2 = dict(t)
    t2["c_custkey"] = t["c_custkey"][:0]
    got, _ = engine.q18(t2, -1)
    assert len(got) == 0
    torch.cuda.synchronize()


# ------------------------------------------------------------------ build
def test_derive_lineitem_matches_oracle(engine):
    import torch
    t = engine.gen_q18_tables(SEED, 10_000, 250_000)
    n = t["l_orderkey"].numel()
    w = engine.gen_lineitem(SEED, n)
    price = w["price"].to(torch.int32)
    disc = w["disc"].to(torch.uint8)
    rf, ls = w["flag_a"], w["flag_b"]
    dp, code, s = engine.derive_lineitem(price, disc, rf, ls, 2, t["l_orderkey"])
    hdp, hcode, hs = oracle.derive_lineitem(price.cpu().numpy(), disc.cpu().numpy(),
                                            rf.cpu().numpy(), ls.cpu().numpy(), 2,
                                            t["l_orderkey"].cpu().numpy())
    assert np.array_equal(dp.cpu().numpy(), hdp) and np.array_equal(code.cpu().numpy(), hcode)
    assert s and hs
    shuffled = t["l_orderkey"].flip(0).contiguous()
    assert not engine.derive_lineitem(price, disc, rf, ls, 2, shuffled)[2]


@pytest.mark.parametrize("rows", [1, 2, 1000, 1_500_001])
def test_date_order_matches_oracle(engine, rows):
    t = engine.gen_q18_tables(SEED, 100, rows)
    d = t["o_orderdate"]
    got = engine.date_order(d).cpu().numpy().astype(np.uint32)
    assert np.array_equal(got, oracle.date_order(d.cpu().numpy()))


def test_order_ranges_match_oracle(engine):
    import torch
    t = engine.gen_q18_tables(SEED, 1000, 200_000)
    order = engine.date_order(t["o_orderdate"]).long()
    ok = t["o_orderkey"][order].contiguous()
    lk = t["l_orderkey"]
    m, start, end = engine.order_ranges(ok, lk, int(ok.max().item()) + 1)
    hm, hs, he = oracle.order_ranges(ok.cpu().numpy(), lk.cpu().numpy(), True)
    assert np.array_equal(m.cpu().numpy(), hm)
    assert np.array_equal(start.cpu().numpy().astype(np.uint32), hs)
    assert np.array_equal(end.cpu().numpy().astype(np.uint32), he)
    perm = torch.randperm(lk.numel(), generator=torch.Generator().manual_seed(1)).cuda()
    lk2 = lk[perm].contiguous()
    m, start, end = engine.order_ranges(ok, lk2, int(ok.max().item()) + 1)
    hm, hs, he = oracle.order_ranges(ok.cpu().numpy(), lk2.cpu().numpy(), False)
    assert np.array_equal(start.cpu().numpy().astype(np.uint32), hs)
    assert np.array_equal(end.cpu().numpy().astype(np.uint32), he)


# ------------------------------------------------------------- Q14 / Q17
def host_part_li(t: dict):
    h = host(t)
    part = {"partkey": h["p_partkey"], "promo": h["p_promo"], "brand": h["p_brand"],
            "container": h["p_container"]}
    li = {"ship": h["l_ship"], "partkey": h["l_partkey"], "disc_price": h["l_disc_price"],
          "qty": h["l_qty"], "price": h["l_price"]}
    return part, li


@pytest.mark.parametrize("parts,rows", [(100, 1000), (20_000, 600_003), (2_000_000, 60_000_000)])
def test_q14_matches_oracle(engine, parts, rows):
    t = engine.gen_part_lineitem(SEED, parts, rows)
    part, li = host_part_li(t)
    for date in ("1995-09-01", "1992-01-01", "1998-11-30"):
        lo, hi = engine.q14_window(date)
        got = engine.q14(t, lo, hi)
        assert got == oracle.q14(li, part, lo, hi), date


def test_q14_over_zones_of_clustered_rows(engine):
    import torch
    t = engine.gen_part_lineitem(SEED, 200_000, 6_000_000)
    cl = engine.q6_cluster({"ship": t["l_ship"], "disc": t["l_disc"], "qty": t["l_qty"],
                            "price": t["l_price"]}, with_perm=True)
    perm = cl["perm"].long()
    t2 = dict(t)
    for k in ("l_ship", "l_partkey", "l_disc_price", "l_qty", "l_price", "l_disc"):
        t2[k] = t[k][perm].contiguous()
    zones = engine.q6_zones({"ship": t2["l_ship"], "disc": t2["l_disc"], "qty": t2["l_qty"]})
    part, li = host_part_li(t)
    lo, hi = engine.q14_window("1995-09-01")
    got = engine.q14(t2, lo, hi, zones)
    assert got == oracle.q14(li, part, lo, hi)
    torch.cuda.synchronize()


def test_q14_duplicate_and_negative_part_keys(engine):
    import torch
    t = engine.gen_part_lineitem(SEED, 50_000, 1_000_000)
    pk = t["p_partkey"].clone()
    pk[::9] = pk[1::9][: pk[::9].numel()]
    pk[3] = -7
    t["p_partkey"] = pk
    part, li = host_part_li(t)
    lo, hi = engine.q14_window("1995-09-01")
    assert engine.q14(t, lo, hi) == oracle.q14(li, part, lo, hi)
    t["p_partkey"] = torch.full_like(pk, -1)
    assert engine.q14(t, lo, hi)["total_sum"] == 0


def test_q14_sparse_part_keys_rerun_at_the_exact_size(engine):
    # Keys far above the dense-key guess (1.5 x parts + 1024) take the exact
    # rerun; lineitem keys are mapped the same way so rows still match.
    t = dict(engine.gen_part_lineitem(SEED, 20_000, 600_003))
    t["p_partkey"] = t["p_partkey"] * 7 + 100_000
    t["l_partkey"] = t["l_partkey"] * 7 + 100_000
    part, li = host_part_li(t)
    lo, hi = engine.q14_window("1995-09-01")
    got = engine.q14(t, lo, hi)
    assert got == oracle.q14(li, part, lo, hi) and got["rows_emitted"] > 0


@pytest.mark.parametrize("size", [1, 1000, 20_001, 1 << 20])
def test_q14_result_independent_of_array_size(engine, size):
    t = engine.gen_part_lineitem(SEED, 20_000, 600_003)
    part, li = host_part_li(t)
    lo, hi = engine.q14_window("1995-09-01")
    engine.tune_set("q14.array_size", size)
    try:
        got = engine.q14(t, lo, hi)
    finally:
        engine.tune_set("q14.array_size", 0)
    assert got == oracle.q14(li, part, lo, hi)


@pytest.mark.parametrize("parts,rows", [(100, 1000), (20_000, 600_003), (2_000_000, 60_000_000)])
def test_q17_matches_oracle(engine, parts, rows):
    t = engine.gen_part_lineitem(SEED, parts, rows)
    part, li = host_part_li(t)
    for brand, container in ((3, 7), (24, 39), (0, 0)):
        preds = [engine.pred(t["p_brand"], N.OP_RANGE, brand, brand),
                 engine.pred(t["p_container"], N.OP_RANGE, container, container)]
        got = engine.q17(t, preds)
        opreds = [oracle.make_pred(part["brand"], 0, brand, brand),
                  oracle.make_pred(part["container"], 0, container, container)]
        assert got == oracle.q17(part["partkey"], opreds, li), (brand, container)


@pytest.mark.parametrize("cap", [1, 3, 64])
def test_q17_match_list_overflow_rescans(engine, cap):
    # Workgroups whose match list overflows re-scan their own rows in pass 2;
    # with cap 1 nearly all do, with 64 only some.
    t = engine.gen_part_lineitem(SEED, 20_000, 600_003)
    part, li = host_part_li(t)
    preds = [engine.pred(t["p_brand"], N.OP_RANGE, 0, 2)]
    engine.tune_set("q17.list_cap", cap)
    try:
        got = engine.q17(t, preds)
    finally:
        engine.tune_set("q17.list_cap", 0)
    want = oracle.q17(part["partkey"], [oracle.make_pred(part["brand"], 0, 0, 2)], li)
    assert got == want and got["matches"] > 0


@pytest.mark.parametrize("brand_hi", [0, 2, 12])  # LDS table / LDS filter + HBM table / HBM
def test_q17_duplicate_negative_and_min_keys(engine, brand_hi):
    import torch
    t = dict(engine.gen_part_lineitem(SEED, 20_000, 600_003))
    pk = t["p_partkey"].clone()
    idx = torch.arange(0, pk.numel() - 1, 7, device=pk.device)
    pk[idx] = pk[idx + 1]          # duplicated part rows
    pk[5] = -3
    t["p_partkey"] = pk
    lk = t["l_partkey"].clone()
    lk[::1000] = -2 ** 31          # the table's empty marker as a fact key
    lk[1::1000] = -3
    t["l_partkey"] = lk
    part, li = host_part_li(t)
    preds = [engine.pred(t["p_brand"], N.OP_RANGE, 0, brand_hi)]
    got = engine.q17(t, preds)
    want = oracle.q17(part["partkey"], [oracle.make_pred(part["brand"], 0, 0, brand_hi)], li)
    assert got == want and got["matches"] > 0


def test_q17_many_parts_take_the_hbm_table(engine):
    t = engine.gen_part_lineitem(SEED, 200_000, 3_000_000)
    part, li = host_part_li(t)
    preds = [engine.pred(t["p_brand"], N.OP_RANGE, 0, 12)]  # about half of the parts
    got = engine.q17(t, preds)
    want = oracle.q17(part["partkey"], [oracle.make_pred(part["brand"], 0, 0, 12)], li)
    assert got == want and got["parts"] > 4096


def test_q17_duplicate_and_negative_part_keys(engine):
    t = engine.gen_part_lineitem(SEED, 50_000, 1_000_000)
    pk = t["p_partkey"].clone()
    pk[::9] = pk[1::9][: pk[::9].numel()]
    pk[3] = -7
    t["p_partkey"] = pk
    part, li = host_part_li(t)
    preds = [engine.pred(t["p_brand"], N.OP_RANGE, 0, 3)]
    got = engine.q17(t, preds)
    assert got == oracle.q17(part["partkey"], [oracle.make_pred(part["brand"], 0, 0, 3)], li)

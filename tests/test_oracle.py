"""Pins the C oracle (oracle/oracle.c) without the reference binary.

The reference cannot be compiled here (Arrow + absent harness headers), and
it holds no golden vectors, so each oracle function is checked against an
independent formulation (pandas/numpy group-bys, joins and sorts written
differently from the C loops), closed forms and invariants.  The mutation
tests plant a plausible mistake in the oracle and require these checks to
catch it.
"""
import numpy as np
import pandas as pd
import pytest

import oracle

SEED = 42
LO = -(1 << 62)


@pytest.fixture(autouse=True)
def _no_mutation():
    oracle.set_mutation(0)
    yield
    oracle.set_mutation(0)


# ----------------------------------------------------------------- generators
def test_generator_value_ranges():
    c = oracle.gen_lineitem(SEED, 200_000)
    assert c["ship"].min() == 0 and c["ship"].max() == 2525
    assert c["qty"].min() == 1 and c["qty"].max() == 50
    assert c["disc"].min() == 0 and c["disc"].max() == 10
    assert c["tax"].min() == 0 and c["tax"].max() == 8
    unit_lo = c["qty"] * 90000 // 100
    unit_hi = c["qty"] * 200000 // 100
    assert np.all(c["price"] >= unit_lo) and np.all(c["price"] <= unit_hi)
    combos = set(zip(c["flag_a"].tolist(), c["flag_b"].tolist()))
    # TPC-H: N-O, N-F, A-F, R-F only.
    assert combos == {(1, 1), (1, 0), (0, 0), (2, 0)}


def test_generator_is_counter_based():
    full = oracle.gen_lineitem(SEED, 5000)
    part = oracle.gen_lineitem(SEED, 1000, row0=3000)
    for k in full:
        assert np.array_equal(full[k][3000:4000], part[k])
    other = oracle.gen_lineitem(SEED + 1, 5000)
    assert not np.array_equal(full["price"], other["price"])


def test_dim_keys_are_a_permutation():
    for n in (1, 7, 1000, 2_000_000 // 64):
        d = oracle.gen_dim(SEED, n)
        assert np.array_equal(np.sort(d["key"]), np.arange(1, n + 1, dtype=np.int32))
    d = oracle.gen_dim(SEED, 100_000)
    assert d["cat"].max() == 149 and d["cont"].max() == 39


def test_chain_generator_shapes():
    t = oracle.gen_chain(SEED, 1000, 5000)
    per = np.diff(t["fact_offsets"])
    assert per.min() == 1 and per.max() == 7
    assert len(t["fact_order"]) == t["fact_offsets"][-1]
    assert np.array_equal(np.repeat(t["order_key"], per), t["fact_order"])
    assert t["order_cust"].min() >= 1 and t["order_cust"].max() <= 1000
    assert t["order_day"].max() <= 2405


def test_uniform_closed_form():
    # og_uniform maps the top 32 hash bits linearly onto [lo, hi].
    L = oracle.lib()
    assert L.og_uniform(0, 5, 9) == 5
    assert L.og_uniform((1 << 64) - 1, 5, 9) == 9
    assert L.og_uniform(1 << 63, 0, 1) == 1
    # SplitMix64 output function, first value of the published sequence
    # from state 0 (state += golden ratio, then mix).
    assert L.og_mix64(0x9E3779B97F4A7C15) == 0xE220A8397B1DCDAF


# ------------------------------------------------------------------------ Q1
def q1_pandas(c, lo, hi, n_a, n_b, dp=None):
    df = pd.DataFrame({k: c[k] for k in ("ship", "qty", "price", "disc", "tax")})
    df["a"] = c["flag_a"].astype(np.int64)
    df["b"] = c["flag_b"].astype(np.int64)
    df = df[(df.ship >= lo) & (df.ship <= hi) & (df.a < n_a) & (df.b < n_b)]
    df = df.assign(dp=df.price * (100 - df.disc) if dp is None else dp[df.index])
    df = df.assign(ch=df.dp * (100 + df.tax))
    g = df.groupby(["a", "b"]).agg(sum_qty=("qty", "sum"), sum_base_price=("price", "sum"),
                                   sum_disc_price=("dp", "sum"), sum_charge=("ch", "sum"),
                                   sum_disc=("disc", "sum"), count=("qty", "size"))
    out = []
    for a in range(n_a):
        for b in range(n_b):
            if (a, b) in g.index:
                r = g.loc[(a, b)]
                out.append({k: int(r[k]) for k in ("sum_qty", "sum_base_price",
                                                   "sum_disc_price", "sum_charge", "sum_disc",
                                                   "count")})
            else:
                out.append(dict(sum_qty=0, sum_base_price=0, sum_disc_price=0,
                                sum_charge=0, sum_disc=0, count=0))
    return out


@pytest.mark.parametrize("rows", [0, 1, 17, 100_003])
def test_q1_wide_matches_pandas(rows):
    c = oracle.gen_lineitem(SEED, rows)
    assert oracle.q1_wide(c, LO, 2435, 3, 2) == q1_pandas(c, LO, 2435, 3, 2)


def test_q1_filter_pass_rate_closed_form():
    rows = 400_000
    c = oracle.gen_lineitem(SEED, rows)
    got = sum(g["count"] for g in oracle.q1_wide(c, LO, 2435, 3, 2))
    p = 2436 / 2526
    assert abs(got - rows * p) < 5 * np.sqrt(rows * p * (1 - p))


def test_q1_ranges_and_dictionary_bounds():
    c = oracle.gen_lineitem(SEED, 50_000)
    assert oracle.q1_wide(c, 100, 900, 3, 2) == q1_pandas(c, 100, 900, 3, 2)
    # A dictionary smaller than the codes drops the rows outside it.
    assert oracle.q1_wide(c, LO, 2435, 2, 1) == q1_pandas(c, LO, 2435, 2, 1)
    none = oracle.q1_wide(c, 3000, 4000, 3, 2)
    assert all(g["count"] == 0 for g in none)


def test_q1_invariant_groups_partition_the_filter():
    c = oracle.gen_lineitem(SEED, 60_000)
    gs = oracle.q1_wide(c, LO, 2435, 3, 2)
    m = c["ship"] <= 2435
    assert sum(g["count"] for g in gs) == int(m.sum())
    assert sum(g["sum_base_price"] for g in gs) == int(c["price"][m].sum())
    assert sum(g["sum_qty"] for g in gs) == int(c["qty"][m].sum())


def ref_layout(c):
    """Reference Database columns (builder_impl.hpp:39-66) from wide columns."""
    dp = (c["price"] * (100 - c["disc"])) // 100  # truncated, builder_impl.cpp:527-531
    return {"ship": c["ship"].astype(np.int16), "qty": (c["qty"] * 100).astype(np.int16),
            "price": c["price"].astype(np.int32), "disc_price": dp.astype(np.int32),
            "disc": c["disc"].astype(np.uint8), "tax": c["tax"].astype(np.uint8),
            "code": (c["flag_a"] * 2 + c["flag_b"]).astype(np.uint8)}


def test_q1_ref_layout_matches_pandas():
    c = oracle.gen_lineitem(SEED, 30_011)
    r = ref_layout(c)
    got = oracle.q1_ref(r, LO, 2435, 6)
    w = {k: r[k].astype(np.int64) for k in ("ship", "qty", "price", "disc", "tax")}
    w["flag_a"] = (r["code"] // 2).astype(np.uint8)
    w["flag_b"] = (r["code"] % 2).astype(np.uint8)
    want = q1_pandas(w, LO, 2435, 3, 2, dp=r["disc_price"].astype(np.int64))
    assert got == want


def test_q1_charge_is_128_bit():
    # Rows whose charge sum exceeds int64: 5 rows of price 2^58.
    n = 5
    c = {"ship": np.zeros(n, np.int32), "qty": np.ones(n, np.int64),
         "price": np.full(n, 1 << 58, np.int64), "disc": np.zeros(n, np.int64),
         "tax": np.zeros(n, np.int64), "flag_a": np.zeros(n, np.uint8),
         "flag_b": np.zeros(n, np.uint8)}
    c["price"][:] = (1 << 56)
    g = oracle.q1_wide(c, LO, 10, 1, 1)[0]
    # dp = 2^56*100 wraps in int64 exactly as the reference's int64 product;
    # the charge accumulates the int64 row products in 128 bits.
    def wrap64(x):
        x &= (1 << 64) - 1
        return x - (1 << 64) if x >> 63 else x
    dp = wrap64((1 << 56) * 100)
    ch = wrap64(dp * 100)
    assert g["sum_charge"] == n * ch
    assert g["sum_disc_price"] == wrap64(n * dp)


# ---------------------------------------------------------------- join + sum
def join_numpy(d, cat, conts, f):
    keep = (d["cat"] == cat) & np.isin(d["cont"], conts)
    keys = d["key"][keep]
    hit = np.isin(f["fk"], keys)
    rev = int((f["price"][hit] * (100 - f["disc"][hit])).sum())
    return {"revenue": rev, "build_rows": int(keep.sum()), "matches": int(hit.sum())}


def c2_preds(d, cat, conts):
    return [oracle.make_pred(d["cat"], 0, cat, cat),
            oracle.make_pred(d["cont"], 1, values=conts)]


@pytest.mark.parametrize("dim_rows,fact_rows", [(1, 10), (1000, 20_000), (200_000, 300_000)])
def test_join_sum_matches_numpy(dim_rows, fact_rows):
    d = oracle.gen_dim(SEED, dim_rows)
    f = oracle.gen_fact(SEED, fact_rows, dim_rows)
    for cat, conts in ((7, [3, 17, 29]), (0, list(range(40))), (149, [39])):
        got = oracle.join_sum(d["key"], c2_preds(d, cat, conts), f["fk"], f["price"],
                              f["disc"])
        assert got == join_numpy(d, cat, conts, f)


def test_join_sum_expected_matches_closed_form():
    # E[matches] = fact_rows * P(cat = c) * P(cont in 3 of 40) = fact/2000.
    d = oracle.gen_dim(SEED, 200_000)
    f = oracle.gen_fact(SEED, 2_000_000, 200_000)
    got = oracle.join_sum(d["key"], c2_preds(d, 7, [3, 17, 29]), f["fk"], f["price"],
                          f["disc"])
    exp_build = 200_000 / 2000
    assert abs(got["build_rows"] - exp_build) < 5 * np.sqrt(exp_build)
    # Given the build set, each fact row hits with probability build/dim.
    p = got["build_rows"] / 200_000
    assert abs(got["matches"] - 2e6 * p) < 5 * np.sqrt(2e6 * p)


def test_join_sum_bitmap_predicate_and_narrow_columns():
    d = oracle.gen_dim(SEED, 5000)
    f = oracle.gen_fact(SEED, 40_000, 5000)
    bitmap = np.zeros(8, np.uint32)  # 256 bits
    for v in (3, 17, 29, 200):
        bitmap[v >> 5] |= np.uint32(1 << (v & 31))
    preds = [oracle.make_pred(d["cat"], 0, 0, 74),
             oracle.make_pred(d["cont"], 1, set_bitmap=bitmap)]
    price32 = f["price"].astype(np.int32)
    disc8 = f["disc"].astype(np.uint8)
    got = oracle.join_sum(d["key"], preds, f["fk"], price32, disc8)
    keep = (d["cat"] <= 74) & np.isin(d["cont"], [3, 17, 29])
    hit = np.isin(f["fk"], d["key"][keep])
    assert got["revenue"] == int((f["price"][hit] * (100 - f["disc"][hit])).sum())


# ------------------------------------------------------------ chain + group-by
def chain_pandas(t, seg, day_lt):
    cust = pd.DataFrame({"c": t["cust_key"], "seg": t["cust_segment"]})
    orders = pd.DataFrame({"key": t["order_key"], "day": t["order_day"],
                           "c": t["order_cust"], "row": np.arange(len(t["order_key"]))})
    fact = pd.DataFrame({"key": t["fact_order"],
                         "v": t["fact_price"] * (100 - t["fact_disc"])})
    o = orders[orders.day < day_lt].merge(cust[cust.seg == seg][["c"]], on="c")
    j = fact.merge(o[["key", "day", "row"]], on="key")
    g = j.groupby(["key", "day", "row"], as_index=False)["v"].sum().sort_values("row")
    return g


@pytest.mark.parametrize("cust,orders", [(10, 50), (1500, 15_000), (20_000, 150_000)])
def test_chain_groups_match_pandas(cust, orders):
    t = oracle.gen_chain(SEED, cust, orders)
    got, stats = oracle.chain_groups(t, 1, 1155)
    want = chain_pandas(t, 1, 1155)
    assert np.array_equal(got["key"], want["key"].to_numpy())
    assert np.array_equal(got["day"], want["day"].to_numpy())
    assert np.array_equal(got["revenue"], want["v"].to_numpy())
    assert stats["groups"] == len(want)
    assert stats["fact_matches"] == int(
        np.isin(t["fact_order"], want["key"].to_numpy()).sum())


def date_sorted_with_ranges(t):
    """The reference build's layout of the chain tables: orders sorted by day
    (stable) and each order's contiguous fact range (the generated fact rows
    are clustered by order key, order rows by key)."""
    perm = np.argsort(t["order_day"], kind="stable")
    o = {k: np.ascontiguousarray(t[k][perm]) for k in ("order_key", "order_day", "order_cust")}
    off = t["fact_offsets"]
    return o, off[:-1][perm], off[1:][perm]


@pytest.mark.parametrize("cust,orders", [(10, 50), (1500, 15_000), (20_000, 150_000)])
def test_q3_range_walk_matches_pandas(cust, orders):
    """run_q3's own algorithm (date-prefix of the sorted orders, range sums,
    full sort) against the pandas merge/group-by, ordered by the full key."""
    t = oracle.gen_chain(SEED, cust, orders)
    o, start, end = date_sorted_with_ranges(t)
    got = oracle.q3_ranges(t, 1, 1155, o, start, end)
    want = chain_pandas(t, 1, 1155).sort_values(["v", "day", "key"],
                                                ascending=[False, True, True])
    assert np.array_equal(got["key"], want["key"].to_numpy())
    assert np.array_equal(got["day"], want["day"].to_numpy())
    assert np.array_equal(got["revenue"], want["v"].to_numpy())
    # an empty prefix and an unknown segment give no rows
    assert len(oracle.q3_ranges(t, 1, 0, o, start, end)) == 0
    assert len(oracle.q3_ranges(t, 9, 1155, o, start, end)) == 0


def test_chain_groups_shuffled_fact_and_disc_price():
    t = oracle.gen_chain(SEED, 2000, 20_000)
    rng = np.random.default_rng(0)
    perm = rng.permutation(len(t["fact_order"]))
    for k in ("fact_order", "fact_price", "fact_disc"):
        t[k] = np.ascontiguousarray(t[k][perm])
    got, _ = oracle.chain_groups(t, 2, 800)
    want = chain_pandas(t, 2, 800)
    assert np.array_equal(got["revenue"], want["v"].to_numpy())
    # Reference layout: sum of the truncated discounted price column.
    dp = ((t["fact_price"] * (100 - t["fact_disc"])) // 100).astype(np.int32)
    got2, _ = oracle.chain_groups(t, 2, 800, disc_price=dp)
    fact = pd.DataFrame({"key": t["fact_order"], "v": dp.astype(np.int64)})
    w2 = fact[fact.key.isin(want["key"])].groupby("key")["v"].sum()
    assert np.array_equal(got2["revenue"], w2.loc[got2["key"]].to_numpy())


def test_sort_groups_matches_python_sorted():
    rng = np.random.default_rng(1)
    rows = np.zeros(5000, oracle.GROUP_DTYPE)
    rows["key"] = rng.permutation(5000) + 1
    rows["day"] = rng.integers(0, 30, 5000)
    rows["revenue"] = rng.integers(0, 50, 5000)  # many ties
    got = oracle.sort_groups(rows)
    want = sorted(rows.tolist(), key=lambda r: (-r[2], r[1], r[0]))
    assert got.tolist() == want


# ----------------------------------------------------------------- radix sort
@pytest.mark.parametrize("n,bits", [(0, 64), (1, 64), (1000, 20), (100_000, 64),
                                    (65_537, 33)])
def test_radix_sort_is_stable_sort(n, bits):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << 63, n, dtype=np.uint64) >> np.uint64(rng.integers(0, 50))
    vals = np.arange(n, dtype=np.uint32)
    k, v = oracle.radix_sort_u64(keys, vals, bits)
    mask = np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    order = np.argsort(keys & mask, kind="stable")
    assert np.array_equal(v, vals[order])
    assert np.array_equal(k, keys[order])


def test_topk_matches_sorted_with_ties():
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 100, 10_000).astype(np.uint64)
    vals = np.arange(10_000, dtype=np.uint32)
    k, v = oracle.topk_u64(keys, vals, 25)
    order = np.argsort(keys, kind="stable")[:25]
    assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])


# ------------------------------------------------- planted mistakes are caught
def test_mutation_q1_tax_sign_is_caught():
    c = oracle.gen_lineitem(SEED, 5000)
    oracle.set_mutation(1)
    assert oracle.q1_wide(c, LO, 2435, 3, 2) != q1_pandas(c, LO, 2435, 3, 2)


def test_mutation_join_discount_factor_is_caught():
    d = oracle.gen_dim(SEED, 1000)
    f = oracle.gen_fact(SEED, 50_000, 1000)
    oracle.set_mutation(2)
    got = oracle.join_sum(d["key"], c2_preds(d, 0, list(range(40))), f["fk"], f["price"],
                          f["disc"])
    assert got != join_numpy(d, 0, list(range(40)), f)


def test_chain_groups_day_boundary():
    t = oracle.gen_chain(SEED, 1500, 15_000)
    t["order_day"][::7] = 1155  # exactly on the exclusive bound
    got, _ = oracle.chain_groups(t, 1, 1155)
    want = chain_pandas(t, 1, 1155)
    assert np.array_equal(got["key"], want["key"].to_numpy())


def test_mutation_chain_day_bound_is_caught():
    t = oracle.gen_chain(SEED, 1500, 15_000)
    t["order_day"][::7] = 1155
    oracle.set_mutation(3)
    got, _ = oracle.chain_groups(t, 1, 1155)
    want = chain_pandas(t, 1, 1155)
    assert len(got) != len(want) or not np.array_equal(got["key"], want["key"].to_numpy())


def test_mutation_sort_direction_is_caught():
    rows = np.zeros(100, oracle.GROUP_DTYPE)
    rows["key"] = np.arange(100) + 1
    rows["revenue"] = np.arange(100) * 7 % 13
    oracle.set_mutation(4)
    got = oracle.sort_groups(rows)
    want = sorted(rows.tolist(), key=lambda r: (-r[2], r[1], r[0]))
    assert got.tolist() != want


def test_mutation_radix_digit_shift_is_caught():
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 1 << 20, 5000, dtype=np.uint64)
    vals = np.arange(5000, dtype=np.uint32)
    oracle.set_mutation(5)
    k, v = oracle.radix_sort_u64(keys, vals, 64)
    assert not np.array_equal(v, vals[np.argsort(keys, kind="stable")])


@pytest.mark.parametrize("ship_gt", [0, 1300, 2525])
def test_chain_groups_ship_filter_matches_pandas(ship_gt):
    """run_q3's l_shipdate > DATE on the fact side: only rows shipped after
    the date are summed, and an order without such a row forms no group."""
    t = oracle.gen_chain(SEED, 3000, 30_000)
    ship = np.random.default_rng(ship_gt).integers(0, 2526, len(t["fact_order"])).astype(np.int16)
    got, stats = oracle.chain_groups(t, 1, 1155, fact_ship=ship, ship_gt=ship_gt)
    keep = ship > ship_gt
    tf = dict(t)
    for k in ("fact_order", "fact_price", "fact_disc"):
        tf[k] = t[k][keep]
    want = chain_pandas(tf, 1, 1155)
    assert np.array_equal(got["key"], want["key"].to_numpy())
    assert np.array_equal(got["revenue"], want["v"].to_numpy())
    assert stats["groups"] == len(want)
    assert stats["fact_matches"] == int(np.isin(tf["fact_order"], want["key"].to_numpy()).sum())
    if ship_gt == 2525:
        assert len(got) == 0

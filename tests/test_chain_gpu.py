"""Chain join + high-cardinality group-by + ordering on the GPU vs the oracle."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SEED = 42


def host(t):
    return {k: v.cpu().numpy() for k, v in t.items()}


def want_sorted(t, seg, day_lt, disc_price=None):
    rows, stats = oracle.chain_groups(host(t), seg, day_lt,
                                      None if disc_price is None else disc_price.cpu().numpy())
    return oracle.sort_groups(rows), stats


def as_array(rows):
    a = np.zeros(len(rows), oracle.GROUP_DTYPE)
    for i, (k, d, r) in enumerate(rows):
        a[i] = (k, d, r)
    return a


def check_stats(got, st):
    for k in ("cust_pass", "order_pass", "fact_matches", "groups"):
        assert got[k] == st[k], k


@pytest.mark.parametrize("cust,orders", [(10, 40), (1500, 15_000), (150_000, 1_500_000)])
def test_chain_topk_and_all_groups(engine, cust, orders):
    t = engine.gen_chain(SEED, cust, orders)
    want, st = want_sorted(t, 1, 1155)
    a = engine.chain_args(t, 1, 1155)
    top, gs = engine.chain_topk(a, 10)
    assert np.array_equal(as_array(top), want[:10])
    check_stats(gs, st)
    allg, gs2 = engine.chain_groups_sorted(a, len(want) + 5)
    assert np.array_equal(allg, want)
    check_stats(gs2, st)


def test_chain_full_c3_size(engine):
    t = engine.gen_chain(SEED, 1_500_000, 15_000_000)
    want, st = want_sorted(t, 1, 1155)
    top, gs = engine.chain_topk(engine.chain_args(t, 1, 1155), 10)
    assert np.array_equal(as_array(top), want[:10])
    check_stats(gs, st)
    assert 1_300_000 < st["groups"] < 1_600_000


@pytest.mark.parametrize("hint", [1, 1000, 100_000, 10_000_000])
def test_chain_capacity_hints_never_change_results(engine, hint):
    # The hashed group table (forced; dense keys would take the directly
    # addressed one, which ignores hints).
    t = engine.gen_chain(SEED, 20_000, 200_000)
    want, st = want_sorted(t, 3, 1800)
    engine.tune_set("chain.direct", 2)
    try:
        top, gs = engine.chain_topk(engine.chain_args(t, 3, 1800, capacity_hint=hint), 25)
    finally:
        engine.tune_set("chain.direct", 0)
    assert np.array_equal(as_array(top), want[:25])
    assert gs["fallbacks"] == (1 if hint * 2 < st["order_pass"] else gs["fallbacks"])
    if hint >= st["order_pass"]:
        assert gs["fallbacks"] == 0
    top, gs = engine.chain_topk(engine.chain_args(t, 3, 1800, capacity_hint=hint), 25)
    assert np.array_equal(as_array(top), want[:25]) and gs["fallbacks"] == 0


@pytest.mark.parametrize("mode", [1, 2])  # directly addressed, hashed
def test_chain_group_table_modes_agree(engine, mode):
    import torch
    t = engine.gen_chain(SEED, 5000, 60_000)
    perm = torch.randperm(t["fact_order"].numel(), generator=torch.Generator().manual_seed(9))
    perm = perm.cuda()
    for k in ("fact_order", "fact_price", "fact_disc"):
        t[k] = t[k][perm].contiguous()
    want, st = want_sorted(t, 4, 1300)
    engine.tune_set("chain.direct", mode)
    try:
        allg, gs = engine.chain_groups_sorted(engine.chain_args(t, 4, 1300), 100_000)
    finally:
        engine.tune_set("chain.direct", 0)
    assert np.array_equal(allg, want)
    check_stats(gs, st)


def test_chain_shuffled_fact_narrow_types_and_disc_price(engine):
    import torch
    t = engine.gen_chain(SEED, 5000, 60_000)
    perm = torch.randperm(t["fact_order"].numel(), generator=torch.Generator().manual_seed(3))
    perm = perm.cuda()
    for k in ("fact_order", "fact_price", "fact_disc"):
        t[k] = t[k][perm].contiguous()
    want, _ = want_sorted(t, 2, 900)
    top, _ = engine.chain_topk(engine.chain_args(t, 2, 900), 50)
    assert np.array_equal(as_array(top), want[:50])
    t32 = dict(t)
    t32["fact_price"] = t["fact_price"].to(torch.int32)
    t32["fact_disc"] = t["fact_disc"].to(torch.uint8)
    allg, _ = engine.chain_groups_sorted(engine.chain_args(t32, 2, 900), 100_000)
    assert np.array_equal(allg, want)
    dp = ((t["fact_price"] * (100 - t["fact_disc"])) // 100).to(torch.int32)
    want_dp, _ = want_sorted(t, 2, 900, disc_price=dp)
    allg, _ = engine.chain_groups_sorted(engine.chain_args(t, 2, 900, disc_price=dp), 100_000)
    assert np.array_equal(allg, want_dp)


def test_chain_negative_keys_and_wide_ranges_fall_back(engine):
    import torch
    t = engine.gen_chain(SEED, 3000, 30_000)
    # Negative, widely spread order keys (unique) and huge revenues: the
    # packed ordering does not fit, the three-pass sort takes over.
    okey = t["order_key"].to(torch.int64) * 7_001 - 100_000_000
    okey = okey.to(torch.int32)
    remap = {int(a): int(b) for a, b in zip(t["order_key"].cpu().tolist(), okey.cpu().tolist())}
    fo = torch.tensor([remap[int(x)] for x in t["fact_order"].cpu().tolist()],
                      dtype=torch.int32).cuda()
    t2 = dict(t)
    t2["order_key"], t2["fact_order"] = okey.contiguous(), fo
    t2["fact_price"] = t["fact_price"] * (1 << 30)
    want, st = want_sorted(t2, 0, 2000)
    allg, gs = engine.chain_groups_sorted(engine.chain_args(t2, 0, 2000), 100_000)
    assert np.array_equal(allg, want)
    assert gs["sort_passes"] > 8


def test_chain_empty_results(engine):
    t = engine.gen_chain(SEED, 1000, 5000)
    top, gs = engine.chain_topk(engine.chain_args(t, 9, 1155), 10)  # no such segment
    assert top == [] and gs["groups"] == 0
    top, gs = engine.chain_topk(engine.chain_args(t, 1, 0), 10)  # no order before day 0
    assert top == [] and gs["order_pass"] == 0


# ---------------------------------------------------------------- by ranges
def offsets_ranges(t):
    import torch
    off = t["fact_offsets"]
    return off[:-1].to(torch.int32).contiguous(), off[1:].to(torch.int32).contiguous()


@pytest.mark.parametrize("cust,orders", [(10, 40), (1500, 15_000), (150_000, 1_500_000),
                                         (1_500_000, 15_000_000)])
def test_chain_by_ranges_matches_oracle(engine, cust, orders):
    t = engine.gen_chain(SEED, cust, orders)
    want, st = want_sorted(t, 1, 1155)
    _, start, end = engine.order_ranges(t["order_key"], t["fact_order"],
                                        int(t["order_key"].max().item()) + 1)
    # the generated fact rows are clustered by order key: the ranges are
    # exact, and checking each row's key changes nothing
    for sorted_ in (True, False):
        a = engine.chain_args(t, 1, 1155, ranges=(start, end), fact_order_sorted=sorted_)
        top, gs = engine.chain_topk(a, 10)
        assert np.array_equal(as_array(top), want[:10])
        check_stats(gs, st)
        if orders <= 1_500_000:
            allg, gs2 = engine.chain_groups_sorted(a, len(want) + 5)
            assert np.array_equal(allg, want)
            check_stats(gs2, st)


@pytest.mark.parametrize("orders", [40, 60_000, 1_500_000])
def test_chain_by_ranges_unsorted_fact_rows(engine, orders):
    """Fact rows out of order (shuffled within blocks of 256 rows, or wholly
    for the smallest case): bolap_order_ranges gives min/max spans that cover
    other orders' rows (the reference builder's unsorted branch), and the
    ranged sum must skip rows of other orders, as run_q3's unsorted branch
    does."""
    import torch
    t = engine.gen_chain(SEED, max(10, orders // 10), orders)
    n = t["fact_order"].numel()
    g = torch.Generator().manual_seed(orders)
    block = (torch.arange(n, dtype=torch.float64) // 256) if orders > 40 else torch.zeros(n, dtype=torch.float64)
    perm = torch.argsort(block + torch.rand(n, generator=g, dtype=torch.float64)).cuda()
    t2 = dict(t)
    for k in ("fact_order", "fact_price", "fact_disc"):
        t2[k] = t[k][perm].contiguous()
    want, st = want_sorted(t2, 1, 1155)
    _, start, end = engine.order_ranges(t2["order_key"], t2["fact_order"],
                                        int(t2["order_key"].max().item()) + 1)
    assert int((end - start).max().item()) > 7  # spans wider than any order's rows
    a = engine.chain_args(t2, 1, 1155, ranges=(start, end))
    allg, gs = engine.chain_groups_sorted(a, len(want) + 5)
    assert np.array_equal(allg, want)
    check_stats(gs, st)
    top, _ = engine.chain_topk(a, 10)
    assert np.array_equal(as_array(top), want[:10])


def test_chain_by_ranges_shuffled_orders_and_orders_without_facts(engine):
    import torch
    t = engine.gen_chain(SEED, 5000, 60_000)
    # Drop the fact rows of every tenth order, then shuffle the order rows.
    keep = (t["fact_order"] % 10) != 3
    t2 = dict(t)
    for k in ("fact_order", "fact_price", "fact_disc"):
        t2[k] = t[k][keep].contiguous()
    perm = torch.randperm(t["order_key"].numel(), generator=torch.Generator().manual_seed(5))
    perm = perm.cuda()
    for k in ("order_key", "order_day", "order_cust"):
        t2[k] = t[k][perm].contiguous()
    want, st = want_sorted(t2, 2, 1500)
    _, start, end = engine.order_ranges(t2["order_key"], t2["fact_order"],
                                        int(t2["order_key"].max().item()) + 1)
    assert int((end == start).sum().item()) > 0
    allg, gs = engine.chain_groups_sorted(engine.chain_args(t2, 2, 1500, ranges=(start, end)),
                                          100_000)
    assert np.array_equal(allg, want)
    check_stats(gs, st)
    assert gs["order_pass"] > gs["groups"]


def test_chain_by_ranges_narrow_types_and_disc_price(engine):
    import torch
    t = engine.gen_chain(SEED, 5000, 60_000)
    rng = offsets_ranges(t)
    t32 = dict(t)
    t32["fact_price"] = t["fact_price"].to(torch.int32)
    t32["fact_disc"] = t["fact_disc"].to(torch.uint8)
    want, _ = want_sorted(t, 2, 900)
    allg, _ = engine.chain_groups_sorted(engine.chain_args(t32, 2, 900, ranges=rng), 100_000)
    assert np.array_equal(allg, want)
    dp = ((t["fact_price"] * (100 - t["fact_disc"])) // 100).to(torch.int32)
    want_dp, _ = want_sorted(t, 2, 900, disc_price=dp)
    allg, _ = engine.chain_groups_sorted(engine.chain_args(t, 2, 900, disc_price=dp, ranges=rng),
                                         100_000)
    assert np.array_equal(allg, want_dp)


def test_chain_by_ranges_negative_keys_and_wide_values(engine):
    import torch
    t = engine.gen_chain(SEED, 3000, 30_000)
    okey = (t["order_key"].to(torch.int64) * 7_001 - 100_000_000).to(torch.int32)
    remap = {int(a): int(b) for a, b in zip(t["order_key"].cpu().tolist(), okey.cpu().tolist())}
    fo = torch.tensor([remap[int(x)] for x in t["fact_order"].cpu().tolist()],
                      dtype=torch.int32).cuda()
    t2 = dict(t)
    t2["order_key"], t2["fact_order"] = okey.contiguous(), fo
    t2["fact_price"] = t["fact_price"] * (1 << 30)
    want, _ = want_sorted(t2, 0, 2000)
    allg, gs = engine.chain_groups_sorted(engine.chain_args(t2, 0, 2000,
                                                            ranges=offsets_ranges(t2)), 100_000)
    assert np.array_equal(allg, want)
    assert gs["sort_passes"] > 8


def test_chain_by_ranges_empty_results(engine):
    t = engine.gen_chain(SEED, 1000, 5000)
    rng = offsets_ranges(t)
    top, gs = engine.chain_topk(engine.chain_args(t, 9, 1155, ranges=rng), 10)
    assert top == [] and gs["groups"] == 0
    top, gs = engine.chain_topk(engine.chain_args(t, 1, 0, ranges=rng), 10)
    assert top == [] and gs["order_pass"] == 0


@pytest.mark.parametrize("mode", [1, 2])  # directly addressed, hashed group table
@pytest.mark.parametrize("widen", [0, 1000])
def test_chain_known_key_ranges_match_computed(engine, mode, widen):
    """Key ranges from the caller (the tables' statistics) instead of the
    call's own min/max pass: same rows on both group-table paths and by
    ranges, also for ranges wider than the keys, with order rows shuffled."""
    import torch
    t = engine.gen_chain(SEED, 20_000, 200_000)
    perm = torch.randperm(t["order_key"].numel(), generator=torch.Generator().manual_seed(9))
    perm = perm.cuda()
    t2 = dict(t)
    for k in ("order_key", "order_day", "order_cust"):
        t2[k] = t[k][perm].contiguous()
    want, st = want_sorted(t2, 3, 1200)
    kr = (int(t2["cust_key"].min()) - widen, int(t2["cust_key"].max()) + widen,
          int(t2["order_key"].min()) - widen, int(t2["order_key"].max()) + widen)
    engine.tune_set("chain.direct", mode)
    try:
        allg, gs = engine.chain_groups_sorted(engine.chain_args(t2, 3, 1200, key_ranges=kr),
                                              len(want) + 5)
    finally:
        engine.tune_set("chain.direct", 0)
    assert np.array_equal(allg, want)
    check_stats(gs, st)
    _, start, end = engine.order_ranges(t2["order_key"], t2["fact_order"],
                                        int(t2["order_key"].max().item()) + 1)
    allr, gr = engine.chain_groups_sorted(
        engine.chain_args(t2, 3, 1200, ranges=(start, end), key_ranges=kr), len(want) + 5)
    assert np.array_equal(allr, want)
    check_stats(gr, st)


@pytest.mark.parametrize("which", ["cust", "order"])
def test_chain_key_outside_given_range_is_an_error(engine, which):
    from bespokeolap_artifacts_amd import native as N
    t = engine.gen_chain(SEED, 2000, 20_000)
    kr = [int(t["cust_key"].min()), int(t["cust_key"].max()),
          int(t["order_key"].min()), int(t["order_key"].max())]
    if which == "cust":
        kr[1] -= 1
    else:
        kr[3] -= 1
    with pytest.raises(N.BolapError):
        engine.chain_groups_sorted(engine.chain_args(t, 1, 1155, key_ranges=kr), 50_000)
    if which == "cust":
        _, start, end = engine.order_ranges(t["order_key"], t["fact_order"],
                                            int(t["order_key"].max().item()) + 1)
        with pytest.raises(N.BolapError):
            engine.chain_groups_sorted(
                engine.chain_args(t, 1, 1155, ranges=(start, end), key_ranges=kr), 50_000)
    # the context stays usable
    want, _ = want_sorted(t, 1, 1155)
    allg, _ = engine.chain_groups_sorted(engine.chain_args(t, 1, 1155), len(want) + 5)
    assert np.array_equal(allg, want)


@pytest.mark.parametrize("path", ["direct", "hashed", "ranges"])
@pytest.mark.parametrize("ship_gt", [1300, 2525])
def test_chain_ship_filter_matches_oracle(engine, path, ship_gt):
    """The fact-side l_shipdate > DATE filter of run_q3 on every path."""
    import torch
    t = engine.gen_chain(SEED, 20_000, 200_000)
    n = int(t["fact_order"].numel())
    ship_h = np.random.default_rng(ship_gt).integers(0, 2526, n).astype(np.int16)
    ship = torch.from_numpy(ship_h).cuda()
    rows, st = oracle.chain_groups(host(t), 2, 1500, fact_ship=ship_h, ship_gt=ship_gt)
    want = oracle.sort_groups(rows)
    kw = {"fact_ship": ship, "ship_gt": ship_gt}
    if path == "ranges":
        _, start, end = engine.order_ranges(t["order_key"], t["fact_order"],
                                            int(t["order_key"].max().item()) + 1)
        kw["ranges"] = (start, end)
    engine.tune_set("chain.direct", 2 if path == "hashed" else 1)
    try:
        allg, gs = engine.chain_groups_sorted(engine.chain_args(t, 2, 1500, **kw), len(want) + 5)
        top, _ = engine.chain_topk(engine.chain_args(t, 2, 1500, **kw), 10)
    finally:
        engine.tune_set("chain.direct", 0)
    assert np.array_equal(allg, want)
    assert np.array_equal(as_array(top), want[:10])
    check_stats(gs, st)


@pytest.mark.parametrize("k", [1, 10, 64])
def test_chain_topk_when_revenues_tie(engine, k):
    """Every group has the same revenue, so one revenue code holds every row:
    the top-k cut cannot narrow the candidates and the call orders the rows
    by the radix path instead (day, then key, break the ties).  A second
    table with a few distinct revenues on top checks the cut itself."""
    import torch
    rng = np.random.default_rng(5)
    orders, cust = 30_000, 3_000
    t = {"cust_key": torch.arange(1, cust + 1, dtype=torch.int32),
         "cust_segment": torch.from_numpy(rng.integers(0, 2, cust).astype(np.uint8)),
         "order_key": torch.arange(1, orders + 1, dtype=torch.int32),
         "order_day": torch.from_numpy(rng.integers(0, 50, orders).astype(np.int32)),
         "order_cust": torch.from_numpy(rng.integers(1, cust + 1, orders).astype(np.int32)),
         "fact_order": torch.arange(1, orders + 1, dtype=torch.int32),
         "fact_price": torch.full((orders,), 1000, dtype=torch.int64),
         "fact_disc": torch.full((orders,), 5, dtype=torch.int64)}
    for bump in (False, True):
        if bump:  # a few orders stand out: the cut is narrow again
            p = t["fact_price"].clone()
            p[rng.integers(0, orders, 100)] += torch.from_numpy(rng.integers(1, 10**6, 100))
            t["fact_price"] = p
        d = {key: v.cuda() for key, v in t.items()}
        want, _ = want_sorted(d, 1, 25)
        assert len(want) > 2048
        for ranges in (None, "ranges"):
            r = None
            if ranges:
                ar = torch.arange(orders, dtype=torch.int32)
                r = (ar.cuda(), (ar + 1).cuda())
            top, _ = engine.chain_topk(engine.chain_args(d, 1, 25, ranges=r,
                                                         fact_order_sorted=True), k)
            assert np.array_equal(as_array(top), want[:k])

"""Star join + group-by, table statistics and the star generators on the GPU,
bit-exact against the oracle."""
import numpy as np
import pytest

import oracle as O
from bespokeolap_artifacts_amd import native as N
from bespokeolap_artifacts_amd import workload as W

pytestmark = pytest.mark.gpu

SMALL = dict(seed=5, fact_rows=1_000_003, dim_rows=[200_000, 30_000, 700],
             fk_s=[0.0, 1.0, 1.5], n_groups=20_000, group_s=0.5)


@pytest.fixture(scope="module")
def small(engine):
    s = W.StarSchema(engine, **SMALL)
    return s, O.gen_star(**SMALL)


def host(t):
    return t.cpu().numpy()


def plans(n, seed=0):
    rng = np.random.default_rng(seed)
    return [W.random_plan(rng) for _ in range(n)]


def test_generators_match_oracle(small):
    s, t = small
    for j in range(3):
        for c in ("key", "cat", "val"):
            np.testing.assert_array_equal(host(s.dims[j][c]), t["dims"][j][c])
    for c in ("fk0", "fk1", "fk2", "grp", "price", "disc"):
        np.testing.assert_array_equal(host(s.fact[c]), t["fact"][c])


def test_hll_matches_oracle(engine, small):
    s, t = small
    for c in ("fk0", "fk1", "fk2", "grp"):
        np.testing.assert_array_equal(engine.hll(s.fact[c]), O.hll_i32(t["fact"][c]))
    np.testing.assert_array_equal(engine.hll(s.dims[2]["key"][:0]), np.zeros(4096, np.uint8))


def test_sample_counts_match_oracle(engine, small):
    s, t = small
    ps = plans(40, seed=3)
    got = engine.sample_counts([W.bind(s, p) for p in ps], 17)
    for i, p in enumerate(ps):
        np.testing.assert_array_equal(got[i], O.star_sample_counts(t, p, 17))


def check_plan(engine, s, t, plan, pass_hints=None, group_hint=0):
    a = W.bind(s, plan, pass_hints, group_hint)
    st, keys, rev = engine.star_groupby(a, want_rows=True)
    ost, okeys, orev = O.star_groupby(t, plan)
    np.testing.assert_array_equal(host(keys), okeys)
    np.testing.assert_array_equal(host(rev), orev)
    assert st["dim_pass"] == ost["dim_pass"]
    assert st["matches"] == ost["matches"]
    assert st["groups"] == ost["groups"]
    assert st["revenue"] == ost["revenue"]
    return st, ost


def test_exact_plans_match_oracle(engine, small):
    s, t = small
    seen = set()
    for plan in plans(30):
        st, _ = check_plan(engine, s, t, plan)
        assert st["fallbacks"] == 0 and st["count_passes"] == 2
        seen.update((p, ) for p in st["dim_placement"])
        seen.add(("g", st["group_placement"]))
    # both dimension placements occurred; groups pre-aggregate in LDS by default
    assert {(N.PLACE_LDS,), (N.PLACE_HBM,), ("g", N.PLACE_LDS)} <= seen


def test_hinted_plans_skip_counting_and_match(engine, small):
    s, t = small
    for plan in plans(20, seed=1):
        ost, _, _ = O.star_groupby(t, plan)
        hints = [max(1, p) for p in ost["dim_pass"]]
        st, _ = check_plan(engine, s, t, plan, hints, max(1, ost["groups"]))
        assert st["fallbacks"] == 0 and st["count_passes"] == 0


def test_undersized_hints_fall_back_to_exact(engine, small):
    s, t = small
    n_fb = 0
    for plan in plans(20, seed=2):
        ost, _, _ = O.star_groupby(t, plan)
        big = ost["groups"] > 1024 or any(p > 1024 for p in ost["dim_pass"])
        st, _ = check_plan(engine, s, t, plan, [1] * len(plan["dims"]), 1)
        if big:
            assert st["fallbacks"] == 1
        n_fb += st["fallbacks"]
    assert n_fb > 0


def test_group_placement_knob_gives_same_result(engine, small):
    s, t = small
    plan = {"dims": [{"dim": 2, "preds": [{"col": "cat", "op": 0, "lo": 0, "hi": 200}]}],
            "group": "fk2"}
    st1, _ = check_plan(engine, s, t, plan)
    assert st1["group_placement"] == N.PLACE_LDS
    engine.tune_set("star.group_lds", 2)
    try:
        st2, _ = check_plan(engine, s, t, plan)
        assert st2["group_placement"] == N.PLACE_HBM
    finally:
        engine.tune_set("star.group_lds", 0)


def test_wide_group_range_orders_through_radix_sort(engine, small):
    """A group key range above 2^26 keys takes the radix-sort ordering path;
    the groups must equal the bitmap-ordered ones."""
    s, t = small
    import torch
    for plan in plans(8, seed=6):
        a = W.bind(s, plan)
        st1, k1, r1 = engine.star_groupby(a, want_rows=True)
        b = W.bind(s, plan)
        b.group_min, b.group_max = -(1 << 30), 1 << 30
        st2, k2, r2 = engine.star_groupby(b, want_rows=True, max_rows=max(1, st1["groups"]))
        assert torch.equal(k1, k2) and torch.equal(r1, r2)
        assert st1["groups"] == st2["groups"] and st1["revenue"] == st2["revenue"]


def test_edge_cases(engine):
    import torch
    dev = "cuda:0"
    # negative keys, one-row dimension, group range of one key, keys outside the dimension
    dkey = torch.tensor([-7, -6, -5, 3], dtype=torch.int32, device=dev)
    dval = torch.tensor([1, 2, 3, 4], dtype=torch.int32, device=dev)
    fk = torch.tensor([-7, -5, 3, 99, -1000, 3, -6], dtype=torch.int32, device=dev)
    grp = torch.tensor([5, 5, 5, 5, 5, 5, 5], dtype=torch.int32, device=dev)
    price = torch.arange(1, 8, dtype=torch.int64, device=dev) * 1000
    disc = torch.tensor([0, 1, 2, 3, 4, 5, 6], dtype=torch.int64, device=dev)
    from bespokeolap_artifacts_amd.engine import Engine
    for hint in (0, 1, 1000):
        a = N.StarArgs()
        a.n_dims = 1
        d = a.dims[0]
        d.rows, d.key, d.key_min, d.key_max = 4, dkey.data_ptr(), -7, 3
        d.n_preds = 1
        d.preds[0] = Engine.pred(dval, N.OP_RANGE, 1, 3)      # keys -7, -6, -5
        d.fact_fk = fk.data_ptr()
        d.pass_hint = hint
        a.fact_rows = 7
        a.group_key, a.group_min, a.group_max = grp.data_ptr(), 5, 5
        a.fact_price, a.price_dtype = price.data_ptr(), N.I64
        a.fact_disc, a.disc_dtype = disc.data_ptr(), N.I64
        a.group_hint = hint
        st, keys, rev = engine.star_groupby(a, want_rows=True)
        # rows 0 (-7), 1 (-5), 6 (-6) pass
        want = 1000 * 100 + 2000 * 99 + 7000 * 94
        assert st["matches"] == 3 and st["groups"] == 1 and st["dim_pass"] == [3]
        assert host(keys).tolist() == [5] and host(rev).tolist() == [want]
    # empty fact table
    a.fact_rows = 0
    st, keys, rev = engine.star_groupby(a, want_rows=True)
    assert st["matches"] == 0 and st["groups"] == 0 and keys.numel() == 0
    # a key outside the declared range is an argument error, not a silent result
    a.fact_rows = 7
    a.group_min = a.group_max = 6
    with pytest.raises(N.BolapError):
        engine.star_groupby(a)


def test_full_size_hinted_equals_exact(engine):
    """60M fact rows: learned-size and exact-size runs give identical groups."""
    s = W.StarSchema(engine, seed=42, fact_rows=60_000_000,
                     dim_rows=[4_000_000, 1_000_000, 100_000], fk_s=[1.0, 0.5, 1.5],
                     n_groups=1_000_000, group_s=1.0)
    import torch
    for plan in plans(6, seed=9):
        a = W.bind(s, plan)
        st, k1, r1 = engine.star_groupby(a, want_rows=True)
        hints = [max(1, 2 * p) for p in st["dim_pass"]]
        b = W.bind(s, plan, hints, max(1, 2 * st["groups"]))
        st2, k2, r2 = engine.star_groupby(b, want_rows=True)
        assert st2["fallbacks"] == 0 and st2["count_passes"] == 0
        assert torch.equal(k1, k2) and torch.equal(r1, r2)
        assert int(r1.sum().item()) == st["revenue"] == st2["revenue"]
        assert st["matches"] == st2["matches"]
        if st["groups"]:
            assert bool((k1[1:] > k1[:-1]).all().item())


def test_full_size_matches_oracle(engine):
    """60M fact rows against the oracle itself (not only hinted against
    exact): the groups, their revenues and the plan counts of three plans,
    exact-sized and estimate-sized (the first plan's group hint is 1, far
    too small for its groups)."""
    cfg = dict(seed=43, fact_rows=60_000_000, dim_rows=[4_000_000, 1_000_000, 100_000],
               fk_s=[1.0, 0.5, 1.5], n_groups=1_000_000, group_s=1.0)
    s = W.StarSchema(engine, **cfg)
    t = O.gen_star(**cfg)
    for i, plan in enumerate(plans(3, seed=11)):
        a = W.bind(s, plan)
        st, keys, rev = engine.star_groupby(a, want_rows=True)
        ost, okeys, orev = O.star_groupby(t, plan)
        np.testing.assert_array_equal(host(keys), okeys)
        np.testing.assert_array_equal(host(rev), orev)
        small_hint = 1 if i == 0 else max(1, 2 * st["groups"])
        b = W.bind(s, plan, [max(1, 2 * p) for p in st["dim_pass"]], small_hint)
        st2, k2, r2 = engine.star_groupby(b, want_rows=True)
        np.testing.assert_array_equal(host(k2), okeys)
        np.testing.assert_array_equal(host(r2), orev)

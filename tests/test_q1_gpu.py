"""Q1 fused filter-scan + group-by on the GPU vs the oracle (bit-exact)."""
import numpy as np
import pytest

import oracle
from bespokeolap_artifacts_amd import native as N

pytestmark = pytest.mark.gpu

SEED = 42
LO = -(1 << 62)


def host(cols):
    return {k: v.cpu().numpy() for k, v in cols.items()}


def as_dicts(groups):
    return [g.__dict__ for g in groups]


def test_generators_match_oracle(engine):
    rows = 100_003
    g = host(engine.gen_lineitem(SEED, rows))
    o = oracle.gen_lineitem(SEED, rows)
    for k in o:
        assert np.array_equal(g[k], o[k]), k
    gd = host(engine.gen_dim(SEED, 50_000))
    od = oracle.gen_dim(SEED, 50_000)
    for k in od:
        assert np.array_equal(gd[k], od[k]), k
    gf = host(engine.gen_fact(SEED, 70_001, 50_000))
    of = oracle.gen_fact(SEED, 70_001, 50_000)
    for k in of:
        assert np.array_equal(gf[k], of[k]), k
    gc = host(engine.gen_chain(SEED, 3000, 20_000))
    oc = oracle.gen_chain(SEED, 3000, 20_000)
    for k in oc:
        assert np.array_equal(gc[k], oc[k]), k


@pytest.mark.parametrize("rows", [0, 1, 2, 3, 255, 1000, 65_537, 1_000_003])
def test_q1_wide_parity(engine, rows):
    cols = engine.gen_lineitem(SEED, rows)
    got = engine.q1_scan_agg(cols, LO, 2435, 3, 2)
    assert as_dicts(got) == oracle.q1_wide(host(cols), LO, 2435, 3, 2)


@pytest.mark.parametrize("lo,hi", [(LO, -1), (LO, 1 << 40), (100, 900), (2435, 2435)])
def test_q1_filter_edges(engine, lo, hi):
    cols = engine.gen_lineitem(SEED + 1, 300_001)
    got = engine.q1_scan_agg(cols, lo, hi, 3, 2)
    assert as_dicts(got) == oracle.q1_wide(host(cols), lo, hi, 3, 2)


@pytest.mark.parametrize("n_a,n_b", [(1, 1), (2, 1), (3, 2), (4, 4), (2, 8)])
def test_q1_group_counts(engine, n_a, n_b):
    import torch
    rows = 200_000
    cols = engine.gen_lineitem(SEED, rows)
    g = torch.Generator().manual_seed(7)
    # Random codes, some outside the dictionaries (dropped by both sides).
    cols["flag_a"] = torch.randint(0, n_a + 1, (rows,), generator=g,
                                   dtype=torch.uint8).cuda()
    cols["flag_b"] = torch.randint(0, n_b + 1, (rows,), generator=g,
                                   dtype=torch.uint8).cuda()
    got = engine.q1_scan_agg(cols, LO, 2435, n_a, n_b)
    assert as_dicts(got) == oracle.q1_wide(host(cols), LO, 2435, n_a, n_b)


def test_q1_unaligned_columns_take_scalar_path(engine):
    rows = 100_001
    cols = engine.gen_lineitem(SEED, rows + 1)
    view = {k: v[1:] for k, v in cols.items()}  # 1 element past 256-B alignment
    got = engine.q1_scan_agg(view, LO, 2435, 3, 2)
    assert as_dicts(got) == oracle.q1_wide(host(view), LO, 2435, 3, 2)


def ref_layout(c):
    import torch
    dp = (c["price"] * (100 - c["disc"])) // 100  # builder_impl.cpp:527-531
    return {"ship": c["ship"].to(torch.int16), "qty": (c["qty"] * 100).to(torch.int16),
            "price": c["price"].to(torch.int32), "disc_price": dp.to(torch.int32),
            "disc": c["disc"].to(torch.uint8), "tax": c["tax"].to(torch.uint8),
            "flag_a": (c["flag_a"] * 2 + c["flag_b"]).to(torch.uint8)}


@pytest.mark.parametrize("rows", [5, 4099, 1_000_001])
def test_q1_reference_layout_parity(engine, rows):
    cols = ref_layout(engine.gen_lineitem(SEED, rows))
    got = engine.q1_scan_agg(cols, LO, 2435, 3, 2, layout=N.Q1_LAYOUT_REF)
    h = host(cols)
    h["code"] = h.pop("flag_a")
    assert as_dicts(got) == oracle.q1_ref(h, LO, 2435, 6)


def test_q1_charge_exceeds_int64(engine):
    import torch
    rows = 4096
    cols = engine.gen_lineitem(SEED, rows)
    cols["price"] = torch.full((rows,), 1 << 45, dtype=torch.int64, device="cuda")
    got = engine.q1_scan_agg(cols, LO, 1 << 40, 3, 2)
    want = oracle.q1_wide(host(cols), LO, 1 << 40, 3, 2)
    assert as_dicts(got) == want
    assert max(abs(w["sum_charge"]) for w in want) > (1 << 63)


def test_q1_full_c1_size(engine):
    rows = 6_000_000
    cols = engine.gen_lineitem(SEED, rows)
    got = engine.q1_scan_agg(cols, LO, 2435, 3, 2)
    want = oracle.q1_wide(host(cols), LO, 2435, 3, 2)
    assert as_dicts(got) == want
    assert sum(1 for w in want if w["count"]) == 4


def test_q1_repeated_calls_reset_state(engine):
    cols = engine.gen_lineitem(SEED, 123_457)
    first = as_dicts(engine.q1_scan_agg(cols, LO, 2435, 3, 2))
    for _ in range(5):
        assert as_dicts(engine.q1_scan_agg(cols, LO, 2435, 3, 2)) == first


@pytest.mark.parametrize("finish", [1, 2])
@pytest.mark.parametrize("rows,per_cu", [(6_000_000, 1), (1_234_567, 3), (70_001, 1)])
def test_finish_variants_match_oracle(engine, finish, rows, per_cu):
    """Both finishes of the scan (the last workgroup merges every partial
    row; a separate merge kernel), at several grid sizes, give the oracle's
    groups on two calls in a row (the finish counter returns to zero)."""
    cols = engine.gen_lineitem(SEED + rows, rows)
    want = oracle.q1_wide(host(cols), LO, 2435, 3, 2)
    engine.tune_set("q1.finish", finish)
    engine.tune_set("q1.blocks_per_cu", per_cu)
    try:
        for _ in range(2):
            assert as_dicts(engine.q1_scan_agg(cols, LO, 2435, 3, 2)) == want
    finally:
        engine.tune_set("q1.finish", 0)
        engine.tune_set("q1.blocks_per_cu", 0)

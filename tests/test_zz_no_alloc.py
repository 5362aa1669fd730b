"""The allocation-free join entry point (bolap_join_sum_noalloc) against
bolap_join_sum on the same inputs: two calls in a row, the second smaller,
so that the scratch array's leftover contents from the first call must not
leak into the second; and a dimension too large for the scratch array is
rejected before any launch."""
import ctypes as C

import pytest

from bespokeolap_artifacts_amd import native as N

pytestmark = pytest.mark.gpu


def case(eng, seed, dim_rows, fact_rows, cats, conts):
    d = eng.gen_dim(seed, dim_rows)
    f = eng.gen_fact(seed, fact_rows, dim_rows)
    preds = [eng.pred(d["cat"], N.OP_RANGE, cats[0], cats[1]),
             eng.pred(d["cont"], N.OP_IN_SET, values=conts)]
    return (d, f), eng.join_sum_args(d["key"], preds, f["fk"], f["price"], f["disc"])


FIELDS = ("revenue", "build_rows", "matches", "capacity", "placement", "fallbacks")


def test_noalloc_matches_join_sum_twice_second_smaller(engine):
    keep1, big = case(engine, 7, 2_000_000, 6_000_001, (3, 40), [3, 17, 29])
    want1 = engine.join_sum_run(big)
    got1 = engine.join_sum_run(big, noalloc=True)
    print("call 1", got1, flush=True)
    assert {k: got1[k] for k in FIELDS} == {k: want1[k] for k in FIELDS}
    assert want1["placement"] == N.PLACE_HBM  # the first call fills the HBM table
    keep2, small = case(engine, 8, 20_000, 300_001, (7, 9), [3, 17, 29])
    want2 = engine.join_sum_run(small)
    got2 = engine.join_sum_run(small, noalloc=True)
    print("call 2", got2, flush=True)
    assert {k: got2[k] for k in FIELDS} == {k: want2[k] for k in FIELDS}
    assert want2["placement"] == N.PLACE_LDS
    # the exact answer again after the smaller call (nothing left over)
    again = engine.join_sum_run(big, noalloc=True)
    assert {k: again[k] for k in FIELDS} == {k: want1[k] for k in FIELDS}


def test_noalloc_forced_placements_match(engine):
    keep, a = case(engine, 9, 200_000, 1_000_003, (0, 149), [0, 1, 2, 3])
    for placement in (N.PLACE_HBM, N.PLACE_AUTO):
        a.placement = placement
        want = engine.join_sum_run(a)
        got = engine.join_sum_run(a, noalloc=True)
        assert {k: got[k] for k in FIELDS} == {k: want[k] for k in FIELDS}, placement


def test_noalloc_rejects_a_dimension_beyond_the_scratch(engine):
    import torch
    rows = 8_000_000
    key = torch.arange(1, rows + 1, dtype=torch.int32, device="cuda")
    cat = torch.zeros(rows, dtype=torch.uint8, device="cuda")
    fk = torch.ones(1024, dtype=torch.int32, device="cuda")
    price = torch.ones(1024, dtype=torch.int64, device="cuda")
    a = engine.join_sum_args(key, [engine.pred(cat, N.OP_RANGE, 0, 0)], fk, price, price)
    r = N.JoinSumResult()
    s = engine.lib.bolap_join_sum_noalloc(engine.ctx, C.byref(a), C.byref(r))
    assert s == N.BOLAP_ERR_CAPACITY
    assert b"scratch" in engine.lib.bolap_last_error()
    # the allocating entry point still serves it
    assert engine.join_sum_run(a)["build_rows"] == rows

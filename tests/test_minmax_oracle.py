"""The grouped MIN/MAX oracle (oracle/ceb.c, bespoke_ceb/query_q10a.cpp:838-1107)
against an independent pandas formulation, in both of the reference's
placements (dense matrix, hashed), with nulls, dropped join rows, empty
ranges and a planted mistake."""
import numpy as np
import pytest

import oracle as O
from tests.minmax_data import gen, pandas_minmax


def run(d, dense_max=0):
    return O.group_minmax(d["pair_group"], d["pair_item"], d["item_offset"], d["item_count"],
                          d["item_has"], d["item_value"], d["range_value"], d["groups"],
                          d["keys"], dense_max)


@pytest.mark.parametrize("seed,pairs,items,groups,keys", [
    (1, 2000, 300, 40, 25), (2, 50_000, 5000, 300, 200), (3, 1, 1, 1, 1), (4, 3000, 50, 7, 3)])
@pytest.mark.parametrize("dense_max", [0, 1])
def test_oracle_matches_pandas(seed, pairs, items, groups, keys, dense_max):
    d = gen(seed, pairs, items, groups, keys)
    got = run(d, dense_max)
    want = pandas_minmax(d)
    np.testing.assert_array_equal(got, want)


def test_placements_agree_and_nulls_stay_groups():
    d = gen(5, 20_000, 2000, 100, 60, null=0.6)
    a, b = run(d, 0), run(d, 1)
    np.testing.assert_array_equal(a, b)
    assert (a["has"] == 0).any() and (a["has"] == 1).any()
    assert ((a["min"] == 0) & (a["max"] == 0))[a["has"] == 0].all()
    assert (a["min"] <= a["max"]).all()


def test_all_dropped_all_null_and_empty():
    d = gen(6, 500, 40, 10, 10, drop=1.0)
    assert len(run(d)) == 0
    d = gen(7, 500, 40, 10, 10, null=1.0)
    r = run(d)
    assert len(r) > 0 and (r["has"] == 0).all()
    d = gen(8, 0, 10, 3, 3)
    assert len(run(d)) == 0


def test_out_of_bounds_is_an_error():
    d = gen(9, 100, 20, 5, 5)
    d["range_value"][0] = 5
    d["item_count"][:] = np.maximum(d["item_count"], 1)
    d["item_offset"][:] = 0
    d["pair_group"][:] = 0
    with pytest.raises(ValueError):
        run(d)


def test_planted_null_as_zero_is_caught():
    d = gen(10, 5000, 500, 30, 20, null=0.5)
    O.set_mutation(14)
    try:
        got = run(d)
    finally:
        O.set_mutation(0)
    assert not np.array_equal(got, pandas_minmax(d))

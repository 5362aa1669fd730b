"""Grouped MIN/MAX on the GPU (bolap_group_minmax) bit-exact against the oracle:
both placements, nulls, dropped rows, empty ranges, a large hashed case, and
bounds errors."""
import numpy as np
import pytest
import torch

import oracle as O
from bespokeolap_artifacts_amd import native as N
from tests.minmax_data import gen

pytestmark = pytest.mark.gpu


def dev(d, with_has=True):
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda()
         for k, v in d.items() if isinstance(v, np.ndarray)}
    if not with_has:
        t["item_has"] = None
    return t


def gpu(engine, d, dense_max=0, with_has=True):
    t = dev(d, with_has)
    st, rows = engine.group_minmax(t["pair_group"], t["pair_item"], t["item_offset"],
                                   t["item_count"], t["item_has"], t["item_value"],
                                   t["range_value"], d["groups"], d["keys"], dense_max)
    out = np.zeros(rows.shape[0], O.MINMAX_DTYPE)
    r = rows.cpu().numpy()
    for i, f in enumerate(("group", "key", "has", "min", "max")):
        out[f] = r[:, i]
    return st, out


def want(d, dense_max=0, with_has=True):
    return O.group_minmax(d["pair_group"], d["pair_item"], d["item_offset"], d["item_count"],
                          d["item_has"] if with_has else None, d["item_value"],
                          d["range_value"], d["groups"], d["keys"], dense_max)


@pytest.mark.parametrize("seed,pairs,items,groups,keys", [
    (1, 2000, 300, 40, 25), (2, 200_000, 50_000, 3000, 1000), (3, 1, 1, 1, 1),
    (4, 70_001, 9, 3, 2)])
@pytest.mark.parametrize("dense_max", [0, 1])
def test_matches_oracle(engine, seed, pairs, items, groups, keys, dense_max):
    d = gen(seed, pairs, items, groups, keys)
    st, got = gpu(engine, d, dense_max)
    np.testing.assert_array_equal(got, want(d, dense_max))
    slots = groups * keys
    dense = slots <= (dense_max or 5_000_000)
    assert st["placement"] == (N.MINMAX_DENSE if dense else N.MINMAX_HASH)
    assert st["groups"] == len(got)


def test_q10a_sized_hashed_and_dense(engine):
    # 1,000,000 join rows over 500,000 movies with up to 8 info values each:
    # 50,000 name groups x 2,000 infos = 100 M slots (hashed) and a 2,000 x
    # 2,000 variant (4 M slots, dense)
    d = gen(11, 1_000_000, 500_000, 50_000, 2000, max_range=8)
    st, got = gpu(engine, d)
    assert st["placement"] == N.MINMAX_HASH and st["expanded_rows"] > 3_000_000
    np.testing.assert_array_equal(got, want(d))
    d2 = dict(d, pair_group=(d["pair_group"] % 2000).astype(np.int32), groups=2000)
    d2["pair_group"][d["pair_group"] < 0] = -1
    st2, got2 = gpu(engine, d2)
    assert st2["placement"] == N.MINMAX_DENSE
    np.testing.assert_array_equal(got2, want(d2))


def test_nulls_drops_and_missing_has_column(engine):
    d = gen(12, 30_000, 3000, 200, 50, null=0.7, drop=0.5)
    for dm in (0, 1):
        np.testing.assert_array_equal(gpu(engine, d, dm)[1], want(d, dm))
        np.testing.assert_array_equal(gpu(engine, d, dm, with_has=False)[1],
                                      want(d, dm, with_has=False))
    d = gen(13, 5000, 100, 10, 10, drop=1.0)
    st, got = gpu(engine, d)
    assert len(got) == 0 and st["expanded_rows"] == 0
    d = gen(14, 0, 10, 3, 3)
    assert len(gpu(engine, d, 1)[1]) == 0


def test_extreme_values_and_one_hot_group(engine):
    d = gen(15, 100_000, 1000, 1, 1, null=0.1)  # every row into one group
    d["item_value"][0] = np.iinfo(np.int32).min
    d["item_value"][1] = np.iinfo(np.int32).max
    d["item_has"][:2] = 1
    d["item_count"][:2] = np.maximum(d["item_count"][:2], 1)
    d["item_offset"] = np.concatenate([[0], np.cumsum(d["item_count"][:-1])]).astype(np.int32)
    d["range_value"] = np.zeros(int(d["item_count"].sum()), np.int32)
    d["pair_item"][:2] = [0, 1]
    d["pair_group"][:2] = 0
    for dm in (0, 1):
        np.testing.assert_array_equal(gpu(engine, d, dm)[1], want(d, dm))
    # the same rows beyond the dense limit (2 slots > 1): one run of every
    # expanded row through the sort, spanning every wave's chunk
    d2 = dict(d, keys=2)
    st, got = gpu(engine, d2, 1)
    assert st["placement"] == N.MINMAX_HASH and len(got) == 1
    np.testing.assert_array_equal(got, want(d2, 1))


@pytest.mark.parametrize("field", ["range_value", "pair_group", "pair_item"])
def test_out_of_bounds_raise(engine, field):
    d = gen(16, 1000, 100, 10, 10)
    d["item_count"][:] = np.maximum(d["item_count"], 1)
    d["item_offset"] = np.concatenate([[0], np.cumsum(d["item_count"][:-1])]).astype(np.int32)
    d["range_value"] = np.zeros(int(d["item_count"].sum()), np.int32)
    d["pair_group"][:] = 0
    d["pair_item"][0] = 0  # item 0's range holds range_value[0]
    d[field][0] = {"range_value": 10, "pair_group": 10, "pair_item": 100}[field]
    with pytest.raises(N.BolapError):
        gpu(engine, d)


@pytest.mark.parametrize("dense_max", [0, 1])
def test_long_ranges_next_to_empty_and_dropped_rows(engine, dense_max):
    """A few items with very long value ranges (hundreds to thousands of
    values, spanning many 64-value batches of a wave) among empty ranges and
    dropped join rows (group -1) in the same waves, as real q10a movie_info
    ranges are."""
    d = gen(17, 20_000, 2000, 300, 4000, max_range=3, drop=0.3)
    rng = np.random.default_rng(17)
    count = d["item_count"].copy()
    long_items = rng.choice(2000, 12, replace=False)
    count[long_items] = rng.integers(200, 5000, 12)
    count[rng.choice(2000, 400, replace=False)] = 0  # empty ranges
    count[long_items] = np.maximum(count[long_items], 200)
    d["item_count"] = count.astype(np.int32)
    d["item_offset"] = np.concatenate([[0], np.cumsum(count[:-1])]).astype(np.int32)
    d["range_value"] = rng.integers(0, 4000, int(count.sum())).astype(np.int32)
    # every tenth join row points at a long item, next to dropped rows
    d["pair_item"][::10] = rng.choice(long_items, len(d["pair_item"][::10]))
    d["pair_group"][1::10] = -1
    st, got = gpu(engine, d, dense_max)
    assert st["expanded_rows"] > 1_000_000
    np.testing.assert_array_equal(got, want(d, dense_max))

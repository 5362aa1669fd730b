"""Table sort orders and row gathers on the GPU (bolap_sort_order_i32,
bolap_gather_rows) bit-exact against the oracle, up to cast_info's 36 M rows."""
import numpy as np
import pytest
import torch

import oracle as O
from bespokeolap_artifacts_amd import native as N

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,two,spread", [(1, True, 5), (1000, True, 7), (1000, False, 7),
                                          (300_001, True, 100_000), (300_001, False, 1 << 30)])
def test_sort_order_matches_oracle(engine, n, two, spread):
    rng = np.random.default_rng(n + spread)
    p = rng.integers(-spread, spread, n).astype(np.int32)
    if n > 3:
        p[:3] = [np.iinfo(np.int32).min, np.iinfo(np.int32).max, 0]
    s = rng.integers(-4, 4, n).astype(np.int32) if two else None
    order = engine.sort_order(torch.from_numpy(p).cuda(),
                              torch.from_numpy(s).cuda() if two else None)
    got = order.cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(got, O.sort_order_i32(p, s))


def test_cast_info_sized_sort_and_gather(engine):
    # 36,000,000 rows sorted by (movie_id, role_id) as the reference builder
    # sorts cast_info; properties at full size, oracle order on a prefix run
    n = 36_000_000
    g = torch.Generator(device="cuda").manual_seed(3)
    movie = torch.randint(1, 2_500_001, (n,), device="cuda", dtype=torch.int32, generator=g)
    role = torch.randint(1, 13, (n,), device="cuda", dtype=torch.int32, generator=g)
    order = engine.sort_order(movie, role)
    sm = engine.gather_rows(movie, order)
    sr = engine.gather_rows(role, order)
    key = sm.to(torch.int64) * 16 + sr.to(torch.int64)
    assert bool((key[1:] >= key[:-1]).all())
    # a permutation: every row exactly once
    seen = torch.zeros(n, dtype=torch.int32, device="cuda")
    seen.index_add_(0, order.to(torch.int64), torch.ones(n, dtype=torch.int32, device="cuda"))
    assert bool((seen == 1).all())
    # ties in input order: within equal keys the row numbers ascend
    o64 = order.to(torch.int64)
    same = key[1:] == key[:-1]
    assert bool((o64[1:][same] > o64[:-1][same]).all())
    del seen
    m = 200_000
    np.testing.assert_array_equal(
        engine.sort_order(movie[:m], role[:m]).cpu().numpy().view(np.uint32),
        O.sort_order_i32(movie[:m].cpu().numpy(), role[:m].cpu().numpy()))


@pytest.mark.parametrize("dtype", [torch.uint8, torch.int16, torch.int32, torch.int64])
def test_gather_rows_all_widths(engine, dtype):
    n = 100_003
    src = torch.arange(n, device="cuda").to(dtype)
    order = torch.randperm(n, device="cuda").to(torch.int32)
    got = engine.gather_rows(src, order)
    assert torch.equal(got, src[order.to(torch.int64)])


def test_gather_rejects_out_of_range_order(engine):
    src = torch.arange(10, device="cuda", dtype=torch.int32)
    order = torch.tensor([0, 3, 10], device="cuda", dtype=torch.int32)
    with pytest.raises(N.BolapError):
        engine.gather_rows(src, order)


def test_sort_order_unaligned_views(engine):
    """Columns starting one row in (not 16-byte aligned) and a ragged length:
    the min/max pass reads them row by row."""
    rng = np.random.default_rng(21)
    p = torch.from_numpy(rng.integers(-9, 9, 100_003).astype(np.int32)).cuda()
    s = torch.from_numpy(rng.integers(0, 3, 100_003).astype(np.int32)).cuda()
    got = engine.sort_order(p[1:], s[1:]).cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(got, O.sort_order_i32(p[1:].cpu().numpy(), s[1:].cpu().numpy()))
    got1 = engine.sort_order(p[3:]).cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(got1, O.sort_order_i32(p[3:].cpu().numpy()))

"""LSD radix sort and radix-select top-k on the GPU vs the oracle."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def dev(a, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(dtype)).cuda()


@pytest.mark.parametrize("n,bits,spread", [(0, 64, 64), (1, 64, 64), (2, 64, 1), (1000, 20, 20),
                                           (65_537, 64, 64), (1_000_003, 33, 33),
                                           (3_000_000, 64, 12),
                                           # 5- and 6-bit digits (10 bits in two passes,
                                           # 17 bits in three)
                                           (200_000, 10, 10), (300_001, 17, 17)])
def test_radix_sort_matches_stable_sort(engine, n, bits, spread):
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, 1 << 62, n, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    keys = keys >> np.uint64(64 - spread)  # many duplicates when spread is small
    vals = np.arange(n, dtype=np.uint32)
    import torch
    k = dev(keys, np.int64)
    v = dev(vals, np.int32)
    engine.radix_sort(k, v, bits)
    ok, ov = oracle.radix_sort_u64(keys, vals, bits)
    assert np.array_equal(k.cpu().numpy().view(np.uint64), ok)
    assert np.array_equal(v.cpu().numpy().view(np.uint32), ov)


@pytest.mark.parametrize("n,k,spread", [(5, 10, 64), (100, 10, 64), (100_000, 10, 64),
                                        (1_000_000, 100, 40), (1_440_000, 10, 63),
                                        (50_000, 50, 3), (20_000, 7, 0), (300_000, 4096, 20)])
def test_topk_matches_oracle(engine, n, k, spread):
    rng = np.random.default_rng(n + k)
    keys = rng.integers(0, 1 << 63, n, dtype=np.uint64)
    keys = keys >> np.uint64(64 - spread) if spread < 64 else keys
    if spread == 0:
        keys[:] = 0  # every key ties: order falls back to the input index
    vals = rng.integers(0, 1 << 31, n).astype(np.uint32)
    ko, vo = engine.topk(dev(keys, np.int64), dev(vals, np.int32), k)
    wk, wv = oracle.topk_u64(keys, vals, k)
    assert np.array_equal(ko.cpu().numpy().view(np.uint64), wk)
    assert np.array_equal(vo.cpu().numpy().view(np.uint32), wv)

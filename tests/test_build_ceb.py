"""CEB build-time data on the host: the string pool (bolap_string_pool_*,
host code in libbolap, no GPU) against a Python dict in first-appearance
order, and the oracle's table sort order against numpy's stable lexsort."""
import numpy as np
import pytest

import oracle as O
from bespokeolap_artifacts_amd import native as N
from bespokeolap_artifacts_amd.ceb import StringPool


def test_string_pool_ids_follow_first_appearance_across_columns():
    lib = N.load()
    pool = StringPool(lib)
    rng = np.random.default_rng(1)
    words = ["Alpha", "beta", "GAMMA", "", "delta ", "Beta", "é-accent", "x" * 300]
    cols = [[words[i] for i in rng.integers(0, len(words), 500)] for _ in range(3)]
    valid = (rng.random(500) > 0.2).astype(np.uint8)
    seen = {}
    for c, col in enumerate(cols):
        ids = pool.intern(col, valid if c == 1 else None)
        for i, s in enumerate(col):
            if c == 1 and not valid[i]:
                assert ids[i] == -1
                continue
            seen.setdefault(s.encode(), len(seen))
            assert ids[i] == seen[s.encode()]
    chars, offs = pool.export()
    order = sorted(seen, key=seen.get)
    assert [chars[offs[i]:offs[i + 1]] for i in range(len(order))] == order
    lchars, loffs = pool.export(lower=True)
    assert np.array_equal(loffs, offs)
    assert [lchars[loffs[i]:loffs[i + 1]] for i in range(len(order))] == \
        [s.lower() for s in order]  # bytes.lower(): ASCII only, as std::tolower in "C"
    pool.close()


def test_string_pool_empty_and_bad_offsets():
    lib = N.load()
    pool = StringPool(lib)
    assert len(pool.intern([])) == 0
    chars, offs = pool.export()
    assert chars == b"" and list(offs) == [0]
    ids = np.zeros(1, np.int32)
    bad = np.array([0, 5, 2], np.int64)
    buf = np.zeros(8, np.uint8)
    assert lib.bolap_string_pool_intern(pool.h, buf.ctypes.data, bad.ctypes.data, 2, None,
                                        ids.ctypes.data) != 0
    pool.close()


@pytest.mark.parametrize("n,two", [(0, False), (1, True), (5000, True), (5000, False),
                                   (100_000, True)])
def test_oracle_sort_order_is_numpy_stable_lexsort(n, two):
    rng = np.random.default_rng(n)
    p = rng.integers(-50, 50, n).astype(np.int32)
    if n > 2:
        p[:2] = [np.iinfo(np.int32).min, np.iinfo(np.int32).max]
    s = rng.integers(-3, 3, n).astype(np.int32) if two else None
    got = O.sort_order_i32(p, s)
    want = np.lexsort((s, p)) if two else np.argsort(p, kind="stable")
    np.testing.assert_array_equal(got, want.astype(np.uint32))

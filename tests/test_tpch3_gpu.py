"""Q2, Q11, Q15 and Q20 on the GPU (tpch_supp.hip) bit-exact against the
oracle (oracle/tpch3.c) over the same reference-layout columns: small and
medium databases, partsupp with and without the fixed stride, the edge cases
the reference handles (empty windows, nations without suppliers, ties at the
maximum revenue, zero-sum groups under a negative fraction) and Q20's table
sized from a short estimate (the exact rerun), from the exact count and from
a generous estimate.  The SF10 sizes run in tests/test_tpch2_gpu.py."""
import numpy as np
import pytest
import torch

import oracle
from tests import tpch_data as T
from tests.test_tpch2_gpu import upload

pytestmark = pytest.mark.gpu

CASES = {"small": (15, 3000, 700, 300, 60), "medium": (16, 250_000, 20_000, 30_000, 2000)}


@pytest.fixture(scope="module", params=["small", "medium"])
def db(request):
    seed, orders, parts, custs, supps = CASES[request.param]
    r = T.raw(seed, orders, parts=parts, custs=custs, supps=supps, strings=False)
    h = T.build(r)
    return r, h, upload(h)


def first_word_ok(h, prefix: str) -> np.ndarray:
    """p_name starts with `prefix` (names are five colour words)."""
    ok = np.array([c.startswith(prefix) for c in T.COLORS])
    return ok[h["p_name_words"][:, 0]].astype(np.uint8)


# ------------------------------------------------------------------------ Q2
@pytest.mark.parametrize("suffix,size,region", [("BRASS", 15, 3), ("TIN", 3, 2), ("", 7, 0),
                                                ("NOSUCH", 15, 3), ("STEEL", 49, 7)])
def test_q2(engine, db, suffix, size, region):
    r, h, t = db
    nation_ok = (np.asarray(r["n_regionkey"]) == region).astype(np.uint8)
    t_ok = np.array([x.endswith(suffix) for x in h["type_dict"]], np.uint8)
    got = engine.q2(t, t_ok, size, nation_ok)
    want = oracle.q2(h, t_ok, size, nation_ok[h["s_nationkey"]])
    np.testing.assert_array_equal(got.reshape(-1, 2), want.reshape(-1, 2))


def test_q2_without_the_stride_and_with_ties(engine):
    """partsupp rows not a multiple of the part rows (the binary-search path),
    parts without partsupp rows, suppliers outside the table, cost ties."""
    rng = np.random.default_rng(3)
    parts, supps = 5000, 300
    per = rng.integers(0, 9, parts)
    pk = np.repeat(np.arange(1, parts + 1, dtype=np.int32), per)
    h = {"p_size": rng.integers(1, 4, parts).astype(np.int32),
         "p_type": rng.integers(0, 3, parts).astype(np.uint16),
         "ps_partkey": pk, "ps_suppkey": rng.integers(0, supps + 20, len(pk)).astype(np.int32),
         "ps_supplycost": rng.integers(1, 6, len(pk)).astype(np.int32),
         "s_nationkey": rng.integers(0, 30, supps).astype(np.int32)}
    assert len(pk) % parts != 0
    nation_ok = (np.arange(25) % 3 == 0).astype(np.uint8)   # nations 25..29 are outside
    s_ok = np.where(h["s_nationkey"] < 25, nation_ok[np.minimum(h["s_nationkey"], 24)], 0)
    t_ok = np.array([1, 0, 1], np.uint8)
    d = {k: torch.from_numpy(v.astype(np.uint8) if k == "p_type" else v).cuda()
         for k, v in h.items()}
    got = engine.q2(d, t_ok, 2, nation_ok)
    want = oracle.q2(h, t_ok, 2, s_ok.astype(np.uint8))
    assert len(want) > 100
    np.testing.assert_array_equal(got, want)
    with pytest.raises(RuntimeError, match="room"):
        engine.q2(d, t_ok, 2, nation_ok, max_rows=len(want) - 1)


# ----------------------------------------------------------------------- Q11
@pytest.mark.parametrize("fraction", [0.0001, 0.01, 0.0, -0.5, 3.0])
def test_q11(engine, db, fraction):
    _, h, t = db
    t["ps_value"] = torch.from_numpy(h["ps_value"]).cuda()
    nk = int(h["s_suppkey"].max()) + 1
    for nation in (int(h["s_nationkey"][0]), 24, 99):
        pk, val, total = engine.q11(t, nation, nk, fraction)
        wpk, wval, wtotal, _ = oracle.q11(h, nation, fraction)
        assert total == wtotal
        np.testing.assert_array_equal(pk, wpk)
        np.testing.assert_array_equal(val, wval)


def test_q11_long_runs_and_keys_outside_the_mark(engine):
    """Runs of hundreds of rows, suppkeys beyond the marked range, negative
    values that cancel to a zero sum."""
    rng = np.random.default_rng(11)
    per = np.where(rng.random(3000) < 0.02, rng.integers(100, 900, 3000), rng.integers(1, 5, 3000))
    pk = np.repeat(np.arange(3000, dtype=np.int32) * 3 - 7, per)
    h = {"s_suppkey": np.arange(1, 51, dtype=np.int32),
         "s_nationkey": (np.arange(50) % 4).astype(np.int32),
         "ps_partkey": pk, "ps_suppkey": rng.integers(-3, 60, len(pk)).astype(np.int32),
         "ps_value": rng.integers(-1000, 1000, len(pk)).astype(np.int32)}
    d = {k: torch.from_numpy(v).cuda() for k, v in h.items()}
    for fraction in (0.0, 0.001, -0.25):
        pk_g, val_g, tot_g = engine.q11(d, 2, 51, fraction)
        wpk, wval, wtot, _ = oracle.q11(h, 2, fraction)
        assert tot_g == wtot
        np.testing.assert_array_equal(pk_g, wpk)
        np.testing.assert_array_equal(val_g, wval)


# ----------------------------------------------------------------------- Q15
@pytest.mark.parametrize("date", ["1996-01-01", "1993-02-27", "1992-01-01", "2010-01-01"])
def test_q15(engine, db, date):
    _, h, t = db
    nk = int(h["s_suppkey"].max()) + 1
    lo, hi = T.date_offset(date), T.add_months(date, 3)
    assert engine.q15(t, lo, hi, nk) == oracle.q15(h, lo, hi)


def test_q15_ties_keys_without_rows_and_ragged_tail(engine):
    rng = np.random.default_rng(15)
    n = 100_003
    h = {"l_shipdate": rng.integers(0, 40, n).astype(np.int16),
         "l_suppkey": rng.integers(-2, 40, n).astype(np.int32),
         "l_disc_price": np.ones(n, np.int32),
         "s_suppkey": np.array([5, 9, 30, 31, 5, 200], np.int32)}
    # suppliers 5 and 9 get 2000 rows each in the window and no others: a tie
    # at the maximum (every other key has about 600); key 5 has two rows
    sk = h["l_suppkey"]
    sk[sk == 5], sk[sk == 9] = 6, 8
    sk[:4000] = np.where(np.arange(4000) % 2 == 0, 5, 9)
    h["l_shipdate"][:4000] = 3
    d = {k: torch.from_numpy(v).cuda() for k, v in h.items()}
    got = engine.q15(d, 0, 10, 201)
    want = oracle.q15(h, 0, 10)
    assert got == want and len(want[0]) == 2


# ----------------------------------------------------------------------- Q20
@pytest.mark.parametrize("prefix,date,hint", [("g", "1994-01-01", 0), ("f", "1995-01-01", 1),
                                              ("", "1993-01-01", 1 << 20), ("zz", "1994-01-01", 0),
                                              ("b", "1996-02-29", 3)])
def test_q20(engine, db, prefix, date, hint):
    _, h, t = db
    p_ok = first_word_ok(h, prefix)
    smax = int(h["s_suppkey"].max())
    m = np.full(smax + 1, -1, np.int16)
    m[h["s_suppkey"]] = h["s_nationkey"]
    lo, hi = T.date_offset(date), T.add_months(date, 12)
    t["ps_availqty"] = torch.from_numpy(h["ps_availqty"]).cuda()
    for nation in (int(h["s_nationkey"][0]), int(h["s_nationkey"][1]), 99):
        q, st = engine.q20(t, torch.from_numpy(p_ok).cuda(), int(h["p_partkey"].max()), lo, hi,
                           nation, torch.from_numpy(m).cuda(), smax, capacity_hint=hint)
        want, n, pairs = oracle.q20(h, p_ok, lo, hi, nation)
        assert np.array_equal(q, want)
        assert (st.qualified, st.pairs) == (n, pairs)
        if hint == 1 and pairs > 5:
            assert st.fallbacks == 1
        if hint in (0, 1 << 20):
            assert st.fallbacks == 0

"""Pins the star join + group-by oracle and the estimator statistics on the CPU
by independent formulations (pandas / numpy), closed forms and planted
mistakes."""
import ctypes as C

import numpy as np
import pandas as pd
import pytest

import oracle as O
from bespokeolap_artifacts_amd import workload as W

SCHEMA = dict(seed=5, fact_rows=120_000, dim_rows=[30_000, 8_000, 500],
              fk_s=[0.0, 1.0, 1.5], n_groups=3_000, group_s=0.5)


@pytest.fixture(scope="module")
def tables():
    return O.gen_star(**SCHEMA)


def pandas_star(t, plan):
    """The same query written as pandas merges and a group-by."""
    f = pd.DataFrame({k: v for k, v in t["fact"].items()})
    dim_pass = []
    for d in plan["dims"]:
        dt = t["dims"][d["dim"]]
        ok = np.ones(len(dt["key"]), bool)
        for p in d["preds"]:
            col = dt[p["col"]].astype(np.int64)
            if p["op"] == 0:
                ok &= (col >= p["lo"]) & (col <= p["hi"])
            else:
                ok &= np.isin(col, p["values"])
        dim_pass.append(int(ok.sum()))
        f = f[f[f"fk{d['dim']}"].isin(dt["key"][ok])]
    rev = f["price"].to_numpy() * (100 - f["disc"].to_numpy())
    g = pd.Series(rev).groupby(f[plan["group"]].to_numpy()).sum().sort_index()
    return dim_pass, len(f), g.index.to_numpy(np.int32), g.to_numpy(np.int64)


def plans(n, seed=0):
    rng = np.random.default_rng(seed)
    return [W.random_plan(rng) for _ in range(n)]


def test_star_groupby_matches_pandas(tables):
    for plan in plans(25):
        st, keys, rev = O.star_groupby(tables, plan)
        dim_pass, matches, pk, pr = pandas_star(tables, plan)
        assert st["dim_pass"] == dim_pass
        assert st["matches"] == matches
        assert st["groups"] == len(pk)
        np.testing.assert_array_equal(keys, pk)
        np.testing.assert_array_equal(rev, pr)
        assert st["revenue"] == int(pr.sum())


@pytest.mark.parametrize("mutation", [6, 7])
def test_planted_mistakes_are_caught(tables, mutation):
    caught = False
    O.set_mutation(mutation)
    try:
        for plan in plans(25):
            if len(plan["dims"]) < 2 and mutation == 7:
                continue
            st, keys, rev = O.star_groupby(tables, plan)
            _, matches, pk, pr = pandas_star(tables, plan)
            if st["matches"] != matches or len(keys) != len(pk) or not np.array_equal(rev, pr):
                caught = True
                break
    finally:
        O.set_mutation(0)
    assert caught


def test_zipf_generator_closed_forms():
    n = 1000
    # s = 0: the inverse CDF is ceil(i n / 65536), so every key is equally likely.
    q0 = W.zipf_quantiles(n, 0.0)
    np.testing.assert_array_equal(q0[1:], np.ceil(np.arange(1, 65537) * n / 65536).astype(int))
    x = O.zipf_i32(1, 7, 400_000, q0, n, W.coprime_mul(n))
    assert x.min() >= 1 and x.max() <= n
    counts = np.bincount(x, minlength=n + 1)[1:]
    assert abs(counts.mean() - 400.0) < 1e-9 and counts.std() < 40  # binomial sd 20
    # s = 1.5: the most frequent key carries 1 / H(n, 1.5) of the mass.
    q = W.zipf_quantiles(n, 1.5)
    y = O.zipf_i32(2, 9, 400_000, q, n, W.coprime_mul(n))
    h = (np.arange(1, n + 1, dtype=np.float64) ** -1.5).sum()
    top = np.bincount(y).max() / len(y)
    assert abs(top - 1 / h) < 0.01, (top, 1 / h)
    # ranks map to keys one to one: rank 1 is key 1 (offset-free scatter)
    assert np.bincount(y).argmax() == 1


def test_zipf_quantiles_large_domain_tail():
    n = 60_000_000
    for s in (0.5, 1.0, 1.5):
        q = W.zipf_quantiles(n, s)
        assert q[0] == 1 and q[-1] == n and np.all(np.diff(q) >= 0)
    # Zipf(1): the median rank m satisfies H(m) = H(n) / 2, H(x) ~ ln x + 0.5772
    q1 = W.zipf_quantiles(n, 1.0)
    m = q1[32768]
    hn = np.log(n) + 0.5772156649
    hm = np.log(m) + 0.5772156649
    assert abs(hm / hn - 0.5) < 0.01


def hll_estimate(regs: np.ndarray) -> float:
    """Flajolet et al. (2007) HyperLogLog estimator with the small-range
    (linear counting) correction, m = 4096."""
    m = 4096
    alpha = 0.7213 / (1 + 1.079 / m)
    est = alpha * m * m / np.sum(2.0 ** -regs.astype(np.float64))
    zeros = int((regs == 0).sum())
    if est <= 2.5 * m and zeros:
        est = m * np.log(m / zeros)
    return est


@pytest.mark.parametrize("n", [100, 5_000, 200_000, 2_000_000])
def test_hll_error_within_bound(n):
    rng = np.random.default_rng(n)
    vals = rng.permutation(np.arange(-n // 2, n - n // 2, dtype=np.int32))
    col = np.concatenate([vals, vals[: n // 3]])  # duplicates do not change the sketch
    regs = O.hll_i32(col)
    np.testing.assert_array_equal(regs, O.hll_i32(vals))
    est = hll_estimate(regs)
    # 1.04 / sqrt(4096) = 1.6% standard error; 4 sigma
    assert abs(est - n) / n < 0.065, (est, n)


def test_hll_registers_merge_by_max():
    rng = np.random.default_rng(3)
    a = rng.integers(-10**9, 10**9, 50_000).astype(np.int32)
    b = rng.integers(-10**9, 10**9, 70_000).astype(np.int32)
    ra, rb = O.hll_i32(a), O.hll_i32(b)
    np.testing.assert_array_equal(O.hll_i32(np.concatenate([a, b])), np.maximum(ra, rb))
    one = O.hll_i32(np.full(1000, 42, np.int32))
    assert (one > 0).sum() == 1


def test_sample_counts_match_numpy(tables):
    L = O.lib()
    seed = 11

    def rows_of(stream, n):
        return np.array([(L.og_hash(seed, stream, i) >> 32) * n >> 32 for i in range(4096)],
                        dtype=np.int64)

    f_rows = rows_of(0x53460000, SCHEMA["fact_rows"])
    for plan in plans(12, seed=4):
        got = O.star_sample_counts(tables, plan, seed)
        want = np.zeros(16, np.int64)
        all_ok = np.ones(4096, bool)
        for i, d in enumerate(plan["dims"]):
            dt = tables["dims"][d["dim"]]
            ok = np.ones(len(dt["key"]), bool)
            for p in d["preds"]:
                col = dt[p["col"]].astype(np.int64)
                ok &= ((col >= p["lo"]) & (col <= p["hi"])) if p["op"] == 0 else \
                    np.isin(col, p["values"])
            want[i] = ok[rows_of(0x53440000, len(dt["key"]))].sum()
            fk = tables["fact"][f"fk{d['dim']}"][f_rows].astype(np.int64) - 1
            j_ok = ok[fk]
            want[3 + i] = j_ok.sum()
            all_ok &= j_ok
        want[6] = all_ok.sum()
        _, c = np.unique(tables["fact"][plan["group"]][f_rows][all_ok], return_counts=True)
        want[7], want[8], want[9] = len(c), (c == 1).sum(), (c == 2).sum()
        np.testing.assert_array_equal(got, want)

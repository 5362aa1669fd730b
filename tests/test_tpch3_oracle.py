"""The oracle's Q2, Q11, Q15 and Q20 (oracle/tpch3.c, over the reference
Database layout) against independent pandas formulations on seeded
TPC-H-shaped tables, plus planted mistakes that those checks must catch."""
from __future__ import annotations

import numpy as np
import pandas as pd
import pytest

import oracle
from tests import tpch_data as T


@pytest.fixture(scope="module", params=[(3, 40, 30, 12), (4_001, 700, 600, 60)])
def data(request):
    orders, parts, custs, supps = request.param
    r = T.raw(71 + orders, orders, parts=parts, custs=custs, supps=supps)
    return r, T.build(r)


def type_ok(db, suffix):
    return np.array([t.endswith(suffix) for t in db["type_dict"]], np.uint8)


def region_ok(db, r, region):
    reg = r["r_name"].index(region) if region in r["r_name"] else -1
    nat_ok = np.asarray(r["n_regionkey"]) == reg
    return nat_ok[db["s_nationkey"]].astype(np.uint8)


# ------------------------------------------------------------------------ Q2
def q2_pandas(db, suffix, size, s_ok):
    ok = (db["p_size"] == size) & type_ok(db, suffix)[db["p_type"]].astype(bool)
    parts = pd.DataFrame({"prow": np.nonzero(ok)[0]})
    parts["partkey"] = parts["prow"] + 1
    ps = pd.DataFrame({"psrow": np.arange(len(db["ps_partkey"])), "partkey": db["ps_partkey"],
                       "srow": db["ps_suppkey"].astype(np.int64) - 1,
                       "cost": db["ps_supplycost"]})
    ps = ps[(ps["srow"] >= 0) & (ps["srow"] < len(s_ok))]
    ps = ps[s_ok[ps["srow"].to_numpy()].astype(bool)]
    m = parts.merge(ps, on="partkey")
    m = m[m["cost"] == m.groupby("prow")["cost"].transform("min")]
    m = m.sort_values(["prow", "psrow"])
    return m[["prow", "srow"]].to_numpy(np.int32)


@pytest.mark.parametrize("suffix,size,region", [("BRASS", 15, "EUROPE"), ("TIN", 3, "ASIA"),
                                                ("STEEL", 49, "AMERICA"), ("", 7, "AFRICA"),
                                                ("NOSUCH", 15, "EUROPE"), ("BRASS", 15, "NONE")])
def test_q2_matches_pandas(data, suffix, size, region):
    r, db = data
    s_ok = region_ok(db, r, region)
    got = oracle.q2(db, type_ok(db, suffix), size, s_ok)
    want = q2_pandas(db, suffix, size, s_ok)
    np.testing.assert_array_equal(got.reshape(-1, 2), want.reshape(-1, 2))


def test_q2_keeps_every_supplier_at_the_minimum():
    """Ties at the minimum cost are all emitted, in partsupp row order; rows
    whose supplier is outside the table or outside the region drop out."""
    db = {"p_size": np.array([5, 5, 6], np.int32), "p_type": np.array([0, 0, 0], np.uint16),
          "ps_partkey": np.array([1, 1, 1, 1, 2, 2, 3], np.int32),
          "ps_suppkey": np.array([3, 1, 2, 9, 0, 2, 1], np.int32),
          "ps_supplycost": np.array([70, 70, 60, 10, 5, 80, 1], np.int32)}
    s_ok = np.array([1, 0, 1], np.uint8)
    got = oracle.q2(db, np.array([1], np.uint8), 5, s_ok)
    # part 1: suppkey 9 (supplier row 8) is outside the table and suppkey 2
    # (row 1, the cheapest) outside the region; suppliers 3 and 1 (rows 2 and
    # 0) tie at 70.  Part 2: suppkey 0 (row -1) is outside, supplier row 1 out
    # of the region.  Part 3 has another size.
    np.testing.assert_array_equal(got, [[0, 2], [0, 0]])


# ----------------------------------------------------------------------- Q11
def q11_pandas(db, nation, fraction):
    mark = db["s_suppkey"][db["s_nationkey"] == nation]
    ps = pd.DataFrame({"pk": db["ps_partkey"], "sk": db["ps_suppkey"],
                       "v": db["ps_value"].astype(np.int64)})
    ps["v"] = np.where(ps["sk"].isin(mark), ps["v"], 0)
    total = int(ps["v"].sum())
    g = ps.groupby("pk", sort=True)["v"].sum().reset_index()
    g = g[g["v"] != 0]
    g = g[g["v"].astype(float) > float(total) * fraction]
    g = g.sort_values(["v", "pk"], ascending=[False, True])
    return g["pk"].to_numpy(np.int32), g["v"].to_numpy(np.int64), total


@pytest.mark.parametrize("nation,fraction", [(7, 0.0001), (3, 0.01), (12, 0.0), (0, -0.5),
                                             (24, 2.0), (99, 0.001)])
def test_q11_matches_pandas(data, nation, fraction):
    _, db = data
    pk, val, total, _ = oracle.q11(db, nation, fraction)
    wpk, wval, wtotal = q11_pandas(db, nation, fraction)
    assert total == wtotal
    np.testing.assert_array_equal(pk, wpk)
    np.testing.assert_array_equal(val, wval)


def test_q11_groups_are_runs_of_the_sorted_table():
    """The value column is supplycost * availqty as int32 (the builder's
    product); the threshold is compared in double."""
    db = {"s_suppkey": np.array([1, 2, 3], np.int32), "s_nationkey": np.array([4, 4, 1], np.int32),
          "ps_partkey": np.array([5, 5, 6, 7, 7], np.int32),
          "ps_suppkey": np.array([1, 3, 3, 2, 1], np.int32),
          "ps_value": np.array([10, 1000, 50, 20, 30], np.int32)}
    pk, val, total, groups = oracle.q11(db, 4, 0.5)
    assert (total, groups) == (60, 2)          # part 6 sums to 0: no group
    np.testing.assert_array_equal(pk, [7])     # 50 > 30, 10 is not
    np.testing.assert_array_equal(val, [50])


# ----------------------------------------------------------------------- Q15
def q15_window(db, lo, hi):
    ship = db["l_shipdate"].astype(np.int32)
    return int(((ship >= lo) & (ship < hi)).sum())


def q15_pandas(db, lo, hi):
    ship = db["l_shipdate"].astype(np.int32)
    inw = (ship >= lo) & (ship < hi)
    if not inw.any():
        return []
    df = pd.DataFrame({"sk": db["l_suppkey"][inw], "p": db["l_disc_price"][inw].astype(np.int64)})
    rev = df.groupby("sk")["p"].sum()
    rev = rev[rev > 0]
    if rev.empty:
        return []
    mx = rev.max()
    last_row = {int(k): i for i, k in enumerate(db["s_suppkey"].tolist())}
    return [(int(k), last_row[int(k)], int(v)) for k, v in rev[rev == mx].items()
            if int(k) in last_row]


@pytest.mark.parametrize("date", ["1996-01-01", "1993-02-27", "1992-01-01", "1999-01-01",
                                  "1998-07-15"])
def test_q15_matches_pandas(data, date):
    _, db = data
    lo, hi = T.date_offset(date), T.add_months(date, 3)
    got, emitted = oracle.q15(db, lo, hi)
    assert got == q15_pandas(db, lo, hi)
    assert emitted == q15_window(db, lo, hi)


def test_q15_ties_and_keys_without_a_supplier_row():
    db = {"l_shipdate": np.array([10, 11, 12, 13, 40], np.int16),
          "l_suppkey": np.array([2, 4, 4, 5, 2], np.int32),
          "l_disc_price": np.array([300, 100, 200, 300, 999], np.int32),
          "s_suppkey": np.array([4, 2, 3, 2], np.int32)}
    got, emitted = oracle.q15(db, 10, 40)
    # suppliers 2, 4 and 5 tie at 300; 5 has no supplier row; supplier 2's
    # last row is row 3; the row on day 40 is outside [10, 40)
    assert emitted == 4
    assert got == [(2, 3, 300), (4, 0, 300)]


# ----------------------------------------------------------------------- Q20
def q20_pandas(db, r, color, lo, hi, nation):
    p_ok = np.array([n.startswith(color) for n in r["p_name"]])
    colored = set(db["p_partkey"][p_ok].tolist())
    ship = db["l_shipdate"].astype(np.int32)
    li = pd.DataFrame({"pk": db["l_partkey"], "sk": db["l_suppkey"],
                       "q": db["l_quantity"].astype(np.int64)})
    li = li[(ship >= lo) & (ship < hi) & (db["l_supp_nation"] == nation) &
            li["pk"].isin(colored)]
    sums = li.groupby(["pk", "sk"])["q"].sum().reset_index()
    snat = dict(zip(db["s_suppkey"].tolist(), db["s_nationkey"].tolist()))
    ps = pd.DataFrame({"pk": db["ps_partkey"], "sk": db["ps_suppkey"],
                       "avail": db["ps_availqty"].astype(np.int64)})
    ps = ps[ps["pk"].isin(colored) & (ps["sk"].map(snat) == nation)]
    m = ps.merge(sums, on=["pk", "sk"])
    return sorted(set(m.loc[m["avail"] * 200 > m["q"], "sk"].tolist())), len(sums)


@pytest.mark.parametrize("color,date,nation", [("forest", "1994-01-01", 3), ("g", "1995-01-01", 7),
                                               ("", "1993-01-01", 12), ("plum", "1996-02-29", 0),
                                               ("zzz", "1994-01-01", 3)])
def test_q20_matches_pandas(data, color, date, nation):
    r, db = data
    p_ok = np.array([n.startswith(color) for n in r["p_name"]], np.uint8)
    lo, hi = T.date_offset(date), T.add_months(date, 12)
    q, n, pairs = oracle.q20(db, p_ok, lo, hi, nation)
    want, wpairs = q20_pandas(db, r, color, lo, hi, nation)
    assert np.nonzero(q)[0].tolist() == want
    assert (n, pairs) == (len(want), wpairs)


def test_q20_availability_boundary():
    """availqty * 200 must exceed the quantity sum (hundredths): equal is out."""
    db = {"p_partkey": np.array([1, 2], np.int32), "l_shipdate": np.array([5, 5, 5], np.int16),
          "l_partkey": np.array([1, 1, 2], np.int32), "l_suppkey": np.array([1, 1, 2], np.int32),
          "l_supp_nation": np.array([3, 3, 3], np.int16),
          "l_quantity": np.array([1000, 1000, 1000], np.int16),
          "ps_partkey": np.array([1, 2], np.int32), "ps_suppkey": np.array([1, 2], np.int32),
          "ps_availqty": np.array([10, 6], np.int32),
          "s_suppkey": np.array([1, 2], np.int32), "s_nationkey": np.array([3, 3], np.int32)}
    q, n, pairs = oracle.q20(db, np.array([1, 1], np.uint8), 0, 10, 3)
    assert (q.tolist(), n, pairs) == ([0, 0, 1], 1, 2)  # 2000 == 2000 out; 1200 > 1000 in


# ------------------------------------------------------- planted mistakes
@pytest.mark.parametrize("mutation", [27, 28, 29, 30])
def test_planted_mistakes_are_caught(mutation):
    r = T.raw(4_072, 4_001, parts=700, custs=600, supps=60)
    db = T.build(r)
    caught = False
    oracle.set_mutation(mutation)
    try:
        if mutation == 27:
            s_ok = region_ok(db, r, "EUROPE")
            caught = not np.array_equal(oracle.q2(db, type_ok(db, "BRASS"), 15, s_ok).reshape(-1, 2),
                                        q2_pandas(db, "BRASS", 15, s_ok).reshape(-1, 2))
        elif mutation == 28:
            nation = int(db["s_nationkey"][0])  # a nation with suppliers
            pk, _, _, _ = oracle.q11(db, nation, -0.5)
            caught = not np.array_equal(pk, q11_pandas(db, nation, -0.5)[0])
        elif mutation == 29:
            for date in ("1996-01-01", "1993-02-27", "1995-04-01"):
                lo, hi = T.date_offset(date), T.add_months(date, 3)
                got, emitted = oracle.q15(db, lo, hi)
                caught |= got != q15_pandas(db, lo, hi) or emitted != q15_window(db, lo, hi)
        else:
            # many suppliers with few pairs each, so that one pair decides
            r = T.raw(4_072, 4_001, parts=2000, custs=600, supps=600)
            db = T.build(r)
            for nation in db["s_nationkey"][:4].tolist():
                color = ""
                p_ok = np.ones(len(r["p_name"]), np.uint8)
                lo, hi = T.date_offset("1994-01-01"), T.add_months("1994-01-01", 12)
                q, _, pairs = oracle.q20(db, p_ok, lo, hi, nation)
                want, wpairs = q20_pandas(db, r, color, lo, hi, nation)
                caught |= np.nonzero(q)[0].tolist() != want or pairs != wpairs
    finally:
        oracle.set_mutation(0)
    assert caught

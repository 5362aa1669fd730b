"""Host-side result rows and CSV output (libbolap, no GPU) against a Python
restatement of run_q1's finalisation and the reference's CSV writers.

Python's int -> float conversion and '%.2f' formatting are correctly rounded,
as are the C++ conversions and iostream fixed output, so the two must agree
byte for byte.
"""
import datetime
import random

import pytest

from bespokeolap_artifacts_amd import native as N
from bespokeolap_artifacts_amd.engine import Engine, Q1Group

SCALES = {N.Q1_LAYOUT_REF: (100.0, 100.0, 100.0, 100.0 * 100.0, 100.0),
          N.Q1_LAYOUT_WIDE: (1.0, 100.0, 100.0 * 100.0, 100.0 * 100.0 * 100.0, 100.0)}


def py_finalize(groups, rf, ls, layout):
    sq, sp, sdp, sc, sd = SCALES[layout]
    rows = []
    for idx, g in enumerate(groups):
        if g.count == 0:
            continue
        cnt = float(g.count)
        r = {"returnflag": rf[idx // len(ls)], "linestatus": ls[idx % len(ls)],
             "sum_qty": float(g.sum_qty) / sq, "sum_base_price": float(g.sum_base_price) / sp,
             "sum_disc_price": float(g.sum_disc_price) / sdp,
             "sum_charge": float(g.sum_charge) / sc, "count_order": g.count}
        r["avg_qty"] = r["sum_qty"] / cnt
        r["avg_price"] = r["sum_base_price"] / cnt
        r["avg_disc"] = float(g.sum_disc) / sd / cnt
        rows.append(r)
    rows.sort(key=lambda r: (r["returnflag"].encode(), r["linestatus"].encode()))
    return rows


def py_q1_csv(rows):
    out = ["l_returnflag,l_linestatus,sum_qty,sum_base_price,sum_disc_price,sum_charge,avg_qty,"
           "avg_price,avg_disc,count_order\n"]
    for r in rows:
        out.append(f"{r['returnflag']},{r['linestatus']},{r['sum_qty']:.2f},"
                   f"{r['sum_base_price']:.2f},{r['sum_disc_price']:.2f},{r['sum_charge']:.2f},"
                   f"{r['avg_qty']:.2f},{r['avg_price']:.2f},{r['avg_disc']:.4f},"
                   f"{r['count_order']}\n")
    return "".join(out)


def random_groups(rng, n, big=False):
    gs = []
    for _ in range(n):
        if rng.random() < 0.2:
            gs.append(Q1Group(0, 0, 0, 0, 0, 0))
            continue
        m = 1 << 62 if big else 10 ** 12
        charge = rng.randrange(-(1 << 100), 1 << 100) if big else rng.randrange(0, 10 ** 16)
        gs.append(Q1Group(rng.randrange(-m, m), rng.randrange(-m, m), rng.randrange(-m, m),
                          charge, rng.randrange(0, 10 ** 9), rng.randrange(1, 10 ** 9)))
    return gs


@pytest.mark.parametrize("layout", [N.Q1_LAYOUT_REF, N.Q1_LAYOUT_WIDE])
@pytest.mark.parametrize("big", [False, True])
def test_q1_finalize_and_csv_match_python(tmp_path, layout, big):
    rng = random.Random(7 + big)
    rf, ls = ["N", "R", "A"], ["O", "F"]  # first-appearance dictionary order
    for trial in range(20):
        groups = random_groups(rng, 6, big)
        got = Engine.q1_finalize(groups, rf, ls, layout)
        want = py_finalize(groups, rf, ls, layout)
        assert got == want, trial
        path = tmp_path / f"q1_{trial}.csv"
        Engine.q1_write_csv(str(path), got)
        assert path.read_text() == py_q1_csv(want)


def test_q1_rows_sort_by_strings_not_codes():
    groups = [Q1Group(1, 1, 1, 1, 1, 1) for _ in range(4)]
    rows = Engine.q1_finalize(groups, ["R", "A"], ["O", "F"], N.Q1_LAYOUT_REF)
    assert [(r["returnflag"], r["linestatus"]) for r in rows] == \
        [("A", "F"), ("A", "O"), ("R", "F"), ("R", "O")]


def test_q1_finalize_rejects_too_many_groups():
    with pytest.raises(N.BolapError):
        Engine.q1_finalize([Q1Group(0, 0, 0, 0, 0, 0)] * 20, ["a"] * 5, ["b"] * 4)


def test_q6_csv(tmp_path):
    for rev in (0, 123456789, -5, 10 ** 15 + 12345):
        p = tmp_path / "q6.csv"
        Engine.q6_write_csv(str(p), {"revenue": rev, "rows_emitted": 0, "rows_scanned": 0})
        assert p.read_text() == f"revenue\n{rev / 10000.0:.2f}\n"


def test_q18_csv(tmp_path):
    import numpy as np
    dt = np.dtype([("custkey", "<i4"), ("orderkey", "<i4"), ("orderdate", "<i4"),
                   ("totalprice", "<i4"), ("sum_qty", "<i8"), ("cust_row", "<i4"),
                   ("order_row", "<i4")])
    rows = np.array([(7, 100, 0, 12345678, 31200, 1, 99), (3, 5, 2405, 99, 30100, 0, 4),
                     (9, 6, 789, 500000000, 40000, 2, 5)], dtype=dt)
    names = ['Customer#000000003', 'He said "hi"', "plain,comma"]
    p = tmp_path / "q18.csv"
    Engine.q18_write_csv(str(p), rows, names)
    base = datetime.date(1992, 1, 1)
    lines = ["c_name,c_custkey,o_orderkey,o_orderdate,o_totalprice,sum(l_quantity)\n"]
    for r in rows:
        name = names[r["cust_row"]].replace('"', '\\"')
        d = base + datetime.timedelta(days=int(r["orderdate"]))
        lines.append(f'"{name}",{r["custkey"]},{r["orderkey"]},{d.isoformat()},'
                     f'{r["totalprice"] / 100.0:.2f},{r["sum_qty"] / 100.0:.2f}\n')
    assert p.read_text() == "".join(lines)


def test_q14_q17_csv(tmp_path):
    from bespokeolap_artifacts_amd import native as N
    lib = N.load()
    for promo, total in ((0, 0), (123, 1000), (987654321, 1234567890)):
        r = N.Q14Result(promo, total, 1, 100.0 * promo / total if total > 0 else 0.0)
        p = tmp_path / "q14.csv"
        N.check(lib.bolap_q14_write_csv(str(p).encode(), r))
        assert p.read_text() == f"promo_revenue\n{r.promo_revenue:.2f}\n"
    for s in (0, 7, 123456789012, -35):
        r = N.Q17Result(s, 0, 0, 0)
        p = tmp_path / "q17.csv"
        N.check(lib.bolap_q17_write_csv(str(p).encode(), r))
        assert p.read_text() == f"avg_yearly\n{s / 700.0:.2f}\n"


def test_q3_csv(tmp_path):
    """write_q3_csv (bespoke_tpch/query_q3.cpp:367-380): key, revenue / 100 with
    two decimals, ISO order date, shippriority."""
    rows = [(172484, 699, 45567367), (5, 0, 1), (2**31 - 1, 2405, 123456789012),
            (7, 30, 0)]
    p = tmp_path / "q3.csv"
    Engine.q3_write_csv(str(p), rows)
    base = datetime.date(1992, 1, 1)
    want = ["l_orderkey,revenue,o_orderdate,o_shippriority\n"]
    for k, d, r in rows:
        want.append(f"{k},{r / 100.0:.2f},{(base + datetime.timedelta(days=d)).isoformat()},0\n")
    assert p.read_text() == "".join(want)
    Engine.q3_write_csv(str(p), rows, shippriority=[0, 1, 2, 3])
    assert [ln.rsplit(",", 1)[1] for ln in p.read_text().splitlines()[1:]] == ["0", "1", "2", "3"]

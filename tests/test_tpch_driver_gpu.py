"""The TPC-H driver (bespokeolap_artifacts_amd/bolap_tpch): the reference's
load() / build() / query() sequence over a seeded lineitem Parquet file, with
the reference's stdin requests and resultN.csv files, compared byte for byte
with a numpy computation of Q1 and Q6 and the Python restatements of the CSV
writers."""
import datetime
import decimal
import pathlib
import subprocess

import numpy as np
import pytest

from bespokeolap_artifacts_amd import build as B
from bespokeolap_artifacts_amd.engine import Engine, Q1Group
from tests.test_results import py_finalize, py_q1_csv
from bespokeolap_artifacts_amd import native as N

pytestmark = pytest.mark.gpu

BASE = datetime.date(1992, 1, 1)


SEGMENTS = ["BUILDING", "AUTOMOBILE", "MACHINERY", "HOUSEHOLD", "FURNITURE"]
BRANDS = [f"Brand#{i}{j}" for i in range(1, 6) for j in range(1, 6)]
CONTAINERS = [f"{a} {b}" for a in ("SM", "MED", "LG", "JUMBO", "WRAP")
              for b in ("CASE", "BOX", "BAG", "JAR", "PKG", "PACK", "CAN", "DRUM")]


def tables(seed: int, orders: int, parts: int = 2000, custs: int = 300):
    """A small TPC-H-shaped database (orders 1..n with 1-7 lineitems each, in
    order-key order) as Parquet tables, and its columns in the reference's
    integer units."""
    pa = pytest.importorskip("pyarrow")
    rng = np.random.default_rng(seed)
    per = rng.integers(1, 8, orders) if orders > 1 else np.array([1])
    okey = np.repeat(np.arange(1, orders + 1), per)
    n = len(okey)
    qty = rng.integers(1, 51, n)                                  # units
    price = qty * rng.integers(90_000, 200_001, n) // 100         # cents
    disc = rng.integers(0, 11, n)                                 # hundredths
    tax = rng.integers(0, 9, n)
    ship = rng.integers(1, 2526, n)                               # days from 1992-01-01
    # returnflag/linestatus correlated as in TPC-H: R/A ship early (F), N late (O or F)
    late = ship > 1260
    rf = np.where(late, "N", np.where(rng.random(n) < 0.5, "R", "A"))
    ls = np.where(late & (rng.random(n) < 0.97), "O", "F")
    pkey = rng.integers(1, parts + 1, n)
    D = decimal.Decimal
    dec = lambda cents: pa.array([D(int(c)).scaleb(-2) for c in cents], pa.decimal128(15, 2))
    day = lambda offs: pa.array([BASE + datetime.timedelta(days=int(s)) for s in offs],
                                pa.date32())
    li = pa.table({
        "l_orderkey": pa.array(okey.astype(np.int32)), "l_partkey": pa.array(pkey.astype(np.int32)),
        "l_quantity": pa.array([D(int(q)) for q in qty], pa.decimal128(15, 2)),
        "l_extendedprice": dec(price), "l_discount": dec(disc), "l_tax": dec(tax),
        "l_returnflag": pa.array(rf.tolist()), "l_linestatus": pa.array(ls.tolist()),
        "l_shipdate": day(ship)})
    brand = rng.integers(0, len(BRANDS), parts)
    cont = rng.integers(0, len(CONTAINERS), parts)
    promo = rng.random(parts) < 0.25
    ptype = [("PROMO " if p else "STANDARD ") + "ANODIZED TIN" for p in promo]
    part = pa.table({"p_partkey": pa.array(np.arange(1, parts + 1, dtype=np.int32)),
                     "p_brand": pa.array([BRANDS[i] for i in brand]),
                     "p_container": pa.array([CONTAINERS[i] for i in cont]),
                     "p_type": pa.array(ptype)})
    ocust = rng.integers(1, custs + 1, orders)
    oday = rng.integers(0, 2406, orders)
    ototal = rng.integers(100_000, 50_000_000, orders)            # cents
    oprio = rng.integers(0, 2, orders)
    ords = pa.table({"o_orderkey": pa.array(np.arange(1, orders + 1, dtype=np.int32)),
                     "o_custkey": pa.array(ocust.astype(np.int32)),
                     "o_orderdate": day(oday), "o_totalprice": dec(ototal),
                     "o_shippriority": pa.array(oprio.astype(np.int32))})
    names = [f"Customer#{i:09d}" for i in range(1, custs + 1)]
    names[0], names[1] = 'He said "hi"', "plain,comma"
    cseg = rng.integers(0, len(SEGMENTS), custs)
    cust = pa.table({"c_custkey": pa.array(np.arange(1, custs + 1, dtype=np.int32)),
                     "c_name": pa.array(names),
                     "c_mktsegment": pa.array([SEGMENTS[i] for i in cseg])})
    cols = {"qty": qty * 100, "price": price, "disc": disc, "tax": tax, "ship": ship,
            "rf": rf, "ls": ls, "okey": okey, "pkey": pkey, "brand": brand, "cont": cont,
            "promo": promo, "ocust": ocust, "oday": oday, "ototal": ototal, "names": names,
            "oprio": oprio, "cseg": cseg}
    return {"lineitem": li, "part": part, "orders": ords, "customer": cust}, cols


def write(tabs, d: pathlib.Path):
    pq = pytest.importorskip("pyarrow.parquet")
    d.mkdir()
    for name, t in tabs.items():
        pq.write_table(t, d / f"{name}.parquet")


def first_appearance(values):
    order = list(dict.fromkeys(values.tolist()))
    return order, np.array([order.index(v) for v in values.tolist()])


def want_q1(c, delta: int) -> str:
    cutoff = (datetime.date(1998, 12, 1) - datetime.timedelta(days=delta) - BASE).days
    rf, rfc = first_appearance(c["rf"])
    ls, lsc = first_appearance(c["ls"])
    keep = c["ship"] <= cutoff
    dp = c["price"] * (100 - c["disc"]) // 100                    # truncated, all >= 0
    groups = []
    for code in range(len(rf) * len(ls)):
        m = keep & (rfc * len(ls) + lsc == code)
        groups.append(Q1Group(int(c["qty"][m].sum()), int(c["price"][m].sum()),
                              int(dp[m].sum()), int((dp[m] * (100 + c["tax"][m])).sum()),
                              int(c["disc"][m].sum()), int(m.sum())))
    return py_q1_csv(py_finalize(groups, rf, ls, N.Q1_LAYOUT_REF))


def want_q6(c, date: str, discount: str, quantity: str) -> str:
    p = Engine.q6_params(date, discount, quantity)
    m = ((c["ship"] >= p["ship_lo"]) & (c["ship"] < p["ship_hi"]) & (c["disc"] >= p["disc_lo"]) &
         (c["disc"] <= p["disc_hi"]) & (c["qty"] < p["qty_lt"]))
    rev = int((c["price"][m] * c["disc"][m]).sum())
    return f"revenue\n{rev / 10000.0:.2f}\n"


def want_q14(c, date: str) -> str:
    lo, hi = Engine.q14_window(date)
    m = (c["ship"] >= lo) & (c["ship"] < hi)
    dp = c["price"] * (100 - c["disc"]) // 100
    total = int(dp[m].sum())
    promo = int(dp[m & c["promo"][c["pkey"] - 1]].sum())
    return f"promo_revenue\n{(100.0 * promo / total if total > 0 else 0.0):.2f}\n"


def want_q17(c, brand: str, container: str) -> str:
    s = 0
    if brand in BRANDS and container in CONTAINERS:
        ok = (c["brand"] == BRANDS.index(brand)) & (c["cont"] == CONTAINERS.index(container))
        for p in np.nonzero(ok)[0] + 1:
            rows = c["pkey"] == p
            q = c["qty"][rows]
            below = q * 5 * len(q) < q.sum()
            s += int(c["price"][rows][below].sum())
    return f"avg_yearly\n{s / 700.0:.2f}\n"


def want_q18(c, quantity: str) -> str:
    out = ["c_name,c_custkey,o_orderkey,o_orderdate,o_totalprice,sum(l_quantity)\n"]
    if quantity == "<<NULL>>":
        return "".join(out)
    thr = round(float(quantity) * 100)
    sums = np.bincount(c["okey"], weights=c["qty"]).astype(np.int64)
    rows = []
    for k in np.nonzero(sums > thr)[0]:
        o = k - 1
        rows.append((-int(c["ototal"][o]), int(c["oday"][o]), int(k)))
    for neg_total, day, k in sorted(rows):
        cust = int(c["ocust"][k - 1])
        name = c["names"][cust - 1].replace('"', '\\"')
        d = (BASE + datetime.timedelta(days=day)).isoformat()
        out.append(f'"{name}",{cust},{k},{d},{-neg_total / 100.0:.2f},'
                   f'{int(sums[k]) / 100.0:.2f}\n')
    return "".join(out)


def want_q3(c, segment: str, date: str) -> str:
    out = ["l_orderkey,revenue,o_orderdate,o_shippriority\n"]
    if segment not in SEGMENTS:
        return "".join(out)
    d = (datetime.date.fromisoformat(date) - BASE).days
    cust_ok = c["cseg"] == SEGMENTS.index(segment)
    order_ok = (c["oday"] < d) & cust_ok[c["ocust"] - 1]
    row_ok = (c["ship"] > d) & order_ok[c["okey"] - 1]
    dp = c["price"] * (100 - c["disc"]) // 100
    rev = np.bincount(c["okey"][row_ok], weights=dp[row_ok], minlength=len(c["oday"]) + 1)
    has = np.bincount(c["okey"][row_ok], minlength=len(c["oday"]) + 1) > 0
    rows = sorted((-int(rev[k]), int(c["oday"][k - 1]), int(k)) for k in np.nonzero(has)[0])
    for neg, day, k in rows:
        out.append(f"{k},{-neg / 100.0:.2f},{(BASE + datetime.timedelta(days=day)).isoformat()},"
                   f"{int(c['oprio'][k - 1])}\n")
    return "".join(out)


def stub_csv(qid: str, line: str) -> str:
    """The reference's stub result of an id without a pipeline
    (bespoke_tpch/query_impl.cpp:82-103): both fields quoted, an embedded
    double quote written as backslash + quote."""
    esc = lambda v: '"' + v.replace('"', '\\"') + '"'
    return f"query_id,params\n{esc(qid)},{esc(line)}\n"


def run_driver(data: pathlib.Path, run: pathlib.Path, requests):
    run.mkdir()
    return subprocess.run([str(B.DRIVER), str(data) + "/"], input="\n".join(requests) + "\n\n",
                          capture_output=True, text=True, cwd=run, timeout=120)


@pytest.mark.parametrize("orders", [1, 25_001])
def test_driver_answers_like_the_reference(tmp_path, orders):
    if not B.DRIVER.exists():
        pytest.skip("bolap_tpch not built (no Arrow C++)")
    tabs, cols = tables(11 + orders, orders)
    write(tabs, tmp_path / "data")
    b0, c0 = BRANDS[int(cols["brand"][0])], CONTAINERS[int(cols["cont"][0])]
    requests = ['1 90', '6 "1994-01-01" "0.06" "24"', '99 "GREEN" x', '1 "60"',
                '6 1995-01-01 0.05 25', '14 "1995-09-01"', f'17 "{b0}" "{c0}"',
                '17 "Brand#99" "SM BOX"', '18 250', '18 "<<NULL>>"', '14 1996-02-29',
                '3 "BUILDING" "1995-03-15"', '3 MACHINERY 1993-06-01', '3 "NOSUCH" "1995-03-15"']
    r = run_driver(tmp_path / "data", tmp_path / "run", requests)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    # every request gets a timing line and a result file, in request order;
    # an id without a pipeline gets the reference's stub file
    assert [ln.split(" | ")[0] for ln in lines] == [str(i + 1) for i in range(len(requests))]
    assert all(ln.split(" | ")[1].startswith("Execution ms: ") for ln in lines)
    run = tmp_path / "run"
    want = [want_q1(cols, 90), want_q6(cols, "1994-01-01", "0.06", "24"),
            stub_csv("99", '99 "GREEN" x'),
            want_q1(cols, 60), want_q6(cols, "1995-01-01", "0.05", "25"),
            want_q14(cols, "1995-09-01"), want_q17(cols, b0, c0),
            want_q17(cols, "Brand#99", "SM BOX"), want_q18(cols, "250"),
            want_q18(cols, "<<NULL>>"), want_q14(cols, "1996-02-29"),
            want_q3(cols, "BUILDING", "1995-03-15"), want_q3(cols, "MACHINERY", "1993-06-01"),
            want_q3(cols, "NOSUCH", "1995-03-15")]
    for i, w in enumerate(want):
        if w is not None:
            assert (run / f"result{i + 1}.csv").read_text() == w, requests[i]
    if orders > 1:
        assert want[8].count("\n") > 2  # some orders pass the Q18 threshold
        assert want[11].count("\n") > 100  # Q3 rows


def test_driver_lineitem_only_serves_q1_q6_and_rejects_q14(tmp_path):
    if not B.DRIVER.exists():
        pytest.skip("bolap_tpch not built (no Arrow C++)")
    tabs, cols = tables(5, 3000)
    write({"lineitem": tabs["lineitem"]}, tmp_path / "data")
    r = run_driver(tmp_path / "data", tmp_path / "ok", ["1 90", "6 1994-01-01 0.06 24"])
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "ok" / "result1.csv").read_text() == want_q1(cols, 90)
    r = run_driver(tmp_path / "data", tmp_path / "bad", ["14 1995-09-01"])
    assert r.returncode == 1 and "part.parquet" in r.stderr


def test_driver_reports_a_missing_file(tmp_path):
    if not B.DRIVER.exists():
        pytest.skip("bolap_tpch not built (no Arrow C++)")
    r = subprocess.run([str(B.DRIVER), str(tmp_path) + "/"], input="1 90\n\n",
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "lineitem.parquet" in r.stderr


# ------------------------------------------------- Q4, Q5, Q10, Q12, Q19
def idx(values, v) -> int:
    return values.index(v) if v in values else -1


def esc(v: str) -> str:
    return '"' + v.replace('"', '\\"') + '"'


def fixed_2(v: int) -> str:
    """append_fixed_2 (bespoke_tpch/query_q10.cpp:108-122)."""
    s = "-" if v < 0 else ""
    v = abs(v)
    return f"{s}{v // 100}.{v % 100:02d}"


def csv_q4(db, date):
    import oracle
    from tests import tpch_data as T
    c = oracle.q4(db, T.date_offset(date), T.add_months(date, 3), len(db["priority_dict"]))
    rows = sorted((db["priority_dict"][i], int(n)) for i, n in enumerate(c) if n)
    return "o_orderpriority,order_count\n" + "".join(f"{esc(p)},{n}\n" for p, n in rows)


def csv_q5(db, region, date):
    import oracle
    from tests import tpch_data as T
    out = "n_name,revenue\n"
    if region not in db["r_name"]:
        return out
    rk = db["r_name"].index(region)
    inreg = (np.array(db["n_regionkey"]) == rk).astype(np.uint8)
    rev = oracle.q5(db, inreg, T.date_offset(date), T.add_months(date, 12))
    rows = sorted(((-int(v), db["n_name"][k]) for k, v in enumerate(rev) if inreg[k] and v))
    return out + "".join(f"{esc(n)},{-r / 100.0:.2f}\n" for r, n in rows)


def csv_q10(db, r, date):
    import oracle
    from tests import tpch_data as T
    nls = len(db["ls_dict"])
    rc = idx(db["rf_dict"], "R") * nls
    if rc < 0:
        return "c_custkey,c_name,revenue,c_acctbal,n_name,c_address,c_phone,c_comment\n"
    rows = oracle.q10(db, rc, nls, T.date_offset(date), T.add_months(date, 3), 25)
    out = ["c_custkey,c_name,revenue,c_acctbal,n_name,c_address,c_phone,c_comment\n"]
    for ck, row, rev in rows.tolist():
        out.append(f"{ck},{esc(r['c_name'][row])},{fixed_2(rev)},{fixed_2(int(r['c_acctbal'][row]))},"
                   f"{esc(r['n_name'][int(r['c_nationkey'][row])])},{esc(r['c_address'][row])},"
                   f"{esc(r['c_phone'][row])},{esc(r['c_comment'][row])}\n")
    return "".join(out)


def csv_q12(db, m1, m2, date):
    import oracle
    from tests import tpch_data as T
    d = db["shipmode_dict"]
    c1, c2 = -1, -1
    for i, v in enumerate(d):
        if v == m1:
            c1 = i
        elif v == m2:
            c2 = i
    out = "l_shipmode,high_line_count,low_line_count\n"
    if c1 < 0 and c2 < 0:
        return out
    hc = [idx(db["priority_dict"], p) for p in ("1-URGENT", "2-HIGH")]
    got = oracle.q12(db, c1, c2, T.date_offset(date), T.add_months(date, 12), *hc)
    rows = []
    for i, v in enumerate(d):
        slot = 0 if i == c1 else 1 if i == c2 else -1
        if slot >= 0 and any(got[slot]):
            rows.append((v, *got[slot]))
    return out + "".join(f"{esc(m)},{h},{lo}\n" for m, h, lo in sorted(rows))


def csv_q19(db, qtys, brands):
    import oracle
    from tests.test_tpch2_oracle import q19_groups
    q = [None if x == "<<NULL>>" else float(x) for x in qtys]
    b = [None if x == "<<NULL>>" else x for x in brands]
    g = q19_groups(db, b, q)
    rev = 0
    if any(x[0] for x in g):
        rev = oracle.q19(db, g, idx(db["shipmode_dict"], "AIR"), -1,
                         idx(db["instruct_dict"], "DELIVER IN PERSON"))
    return f"revenue\n{rev / 10000.0:.2f}\n"


@pytest.mark.parametrize("orders", [3, 40_001])
def test_driver_order_range_queries_q12_q19(tmp_path, orders):
    """Q4, Q5, Q10, Q12 and Q19 through load / build / query over all eight
    Parquet tables, byte for byte against the oracle over the reference
    layout and the reference's CSV writers restated in Python."""
    if not B.DRIVER.exists():
        pytest.skip("bolap_tpch not built (no Arrow C++)")
    from tests import tpch_data as T
    r = T.raw(17 + orders, orders, parts=800, custs=500, supps=50)
    write(T.parquet(r), tmp_path / "data")
    db = T.build(r)
    requests = ['4 "1995-01-01"', '4 1993-07-31', '5 "ASIA" "1994-01-01"', '5 EUROPE 1996-02-29',
                '5 "ATLANTIS" "1994-01-01"', '10 "1993-10-01"', '10 1994-11-30', '10 2010-01-01',
                '12 "MAIL" "SHIP" "1994-01-01"', '12 AIR NOSUCH 1995-01-01',
                '12 NOSUCH NOPE 1995-01-01', '19 "1" "10" "20" "Brand#12" "Brand#23" "Brand#34"',
                '19 "5" "<<NULL>>" "25" "Brand#11" "Brand#21" "Brand#55"',
                '19 "<<NULL>>" "<<NULL>>" "<<NULL>>" "Brand#11" "Brand#21" "Brand#55"',
                '23 "15" "BRASS" "EUROPE"']
    res = run_driver(tmp_path / "data", tmp_path / "run", requests)
    assert res.returncode == 0, res.stderr
    assert [ln.split(" | ")[0] for ln in res.stdout.splitlines()] == \
        [str(i + 1) for i in range(len(requests))]
    want = [csv_q4(db, "1995-01-01"), csv_q4(db, "1993-07-31"),
            csv_q5(db, "ASIA", "1994-01-01"), csv_q5(db, "EUROPE", "1996-02-29"),
            csv_q5(db, "ATLANTIS", "1994-01-01"), csv_q10(db, r, "1993-10-01"),
            csv_q10(db, r, "1994-11-30"), csv_q10(db, r, "2010-01-01"),
            csv_q12(db, "MAIL", "SHIP", "1994-01-01"), csv_q12(db, "AIR", "NOSUCH", "1995-01-01"),
            csv_q12(db, "NOSUCH", "NOPE", "1995-01-01"),
            csv_q19(db, ("1", "10", "20"), ("Brand#12", "Brand#23", "Brand#34")),
            csv_q19(db, ("5", "<<NULL>>", "25"), ("Brand#11", "Brand#21", "Brand#55")),
            csv_q19(db, ("<<NULL>>",) * 3, ("Brand#11", "Brand#21", "Brand#55")),
            stub_csv("23", '23 "15" "BRASS" "EUROPE"')]
    run = tmp_path / "run"
    for i, w in enumerate(want):
        assert (run / f"result{i + 1}.csv").read_text() == w, requests[i]
    if orders > 3:
        assert want[0].count("\n") == 6 and want[5].count("\n") > 50
        assert want[2].count("\n") > 2 and want[8].count("\n") == 3
        assert want[11] != "revenue\n0.00\n"


def csv_q21(db, r, nation):
    import oracle
    out = "s_name,numwait\n"
    if nation == "<<NULL>>" or nation not in r["n_name"] or "F" not in db["status_dict"]:
        return out
    n = int(db["s_suppkey"].max()) + 1
    w = oracle.q21(db, db["status_dict"].index("F"), r["n_name"].index(nation), n)
    rows = sorted((-int(c), r["s_name"][k - 1]) for k, c in enumerate(w.tolist()) if c)
    return out + "".join(f"{esc(name)},{-c}\n" for c, name in rows)


def csv_q22(db, codes):
    import oracle
    numeric, extra = [], []
    for c in codes:
        if c == "<<NULL>>" or len(c) < 2:
            continue
        if c[:2].isdigit():
            if int(c[:2]) not in numeric:
                numeric.append(int(c[:2]))
        elif c[:2] not in extra:
            extra.append(c[:2])
    out = "cntrycode,numcust,totacctbal\n"
    if not numeric and not extra:
        return out
    num, ext = oracle.q22(db, numeric, [(ord(e[0]) << 8) | ord(e[1]) for e in extra])
    rows = [(f"{c:02d}", int(num[c, 0]), int(num[c, 1])) for c in numeric if num[c, 0]]
    rows += [(e, int(ext[i, 0]), int(ext[i, 1])) for i, e in enumerate(extra) if ext[i, 0]]
    return out + "".join(f"{esc(c)},{n},{t / 100.0:.2f}\n" for c, n, t in sorted(rows))


@pytest.mark.parametrize("orders", [3, 40_001])
def test_driver_q21_q22(tmp_path, orders):
    if not B.DRIVER.exists():
        pytest.skip("bolap_tpch not built (no Arrow C++)")
    from tests import tpch_data as T
    r = T.raw(23 + orders, orders, parts=800, custs=3000, supps=60)
    write(T.parquet(r), tmp_path / "data")
    db = T.build(r)
    q22a = ["13", "31", "23", "29", "30", "18", "17"]
    q22b = ["AB", "20", "7", "<<NULL>>", "20", "AB", "34"]
    q = lambda v: " ".join(f'"{x}"' for x in v)
    requests = ['21 "SAUDI ARABIA"', '21 GERMANY', '21 "ATLANTIS"', '21 "<<NULL>>"',
                f"22 {q(q22a)}", f"22 {q(q22b)}", f"22 {q(['<<NULL>>'] * 7)}"]
    res = run_driver(tmp_path / "data", tmp_path / "run", requests)
    assert res.returncode == 0, res.stderr
    want = [csv_q21(db, r, "SAUDI ARABIA"), csv_q21(db, r, "GERMANY"), csv_q21(db, r, "ATLANTIS"),
            csv_q21(db, r, "<<NULL>>"), csv_q22(db, q22a), csv_q22(db, q22b),
            csv_q22(db, ["<<NULL>>"] * 7)]
    for i, w in enumerate(want):
        assert (tmp_path / "run" / f"result{i + 1}.csv").read_text() == w, requests[i]
    if orders > 3:
        assert want[0].count("\n") > 1 and want[4].count("\n") > 3 and want[5].count("\n") > 2


Q7_DATES = None


def q7_dates():
    from tests import tpch_data as T
    return (T.date_offset("1995-01-01"), T.date_offset("1996-12-31"),
            T.date_offset("1996-01-01"))


def csv_q7(db, r, a, b):
    import oracle
    out = "supp_nation,cust_nation,l_year,revenue\n"
    if a not in r["n_name"] or b not in r["n_name"]:
        return out
    n1, n2 = r["n_name"].index(a), r["n_name"].index(b)
    rev = oracle.q7(db, n1, n2, *q7_dates())
    keys = [n1, n2] if n1 != n2 else [n1]
    rows = []
    for si, sk in enumerate(keys):
        for ci, ck in enumerate(keys):
            if len(keys) == 2 and si == ci:
                continue
            for y in (0, 1):
                if rev[si, ci, y]:
                    rows.append((r["n_name"][sk], r["n_name"][ck], 1995 + y, int(rev[si, ci, y])))
    return out + "".join(f"{esc(s)},{esc(c)},{y},{v / 100.0:.2f}\n" for s, c, y, v in sorted(rows))


def csv_q8(db, r, nation, region, ptype):
    import oracle
    out = "o_year,mkt_share\n"
    if nation not in r["n_name"] or region not in r["r_name"]:
        return out
    inreg = (np.array(r["n_regionkey"]) == r["r_name"].index(region)).astype(np.uint8)
    tc = idx(db["type_dict"], ptype)
    tot, nat = oracle.q8(db, inreg, r["n_name"].index(nation), tc, *q7_dates())
    return out + "".join(f"{1995 + y},{nat[y] / tot[y]:.2f}\n" for y in (0, 1) if tot[y])


def csv_q9(db, r, color):
    import oracle
    p_ok = np.array([color in name for name in r["p_name"]], np.uint8)
    y0, tot, used = oracle.q9(db, p_ok, 25)
    rows = [(r["n_name"][k], -(y0 + y), int(tot[k, y])) for k in range(25)
            for y in range(tot.shape[1]) if used[k, y]]
    return "nation,o_year,sum_profit\n" + "".join(
        f"{esc(n)},{-y},{v / 100.0:.2f}\n" for n, y, v in sorted(rows))


@pytest.mark.parametrize("orders", [3, 40_001])
def test_driver_q7_q8_q9(tmp_path, orders):
    if not B.DRIVER.exists():
        pytest.skip("bolap_tpch not built (no Arrow C++)")
    from tests import tpch_data as T
    r = T.raw(29 + orders, orders, parts=900, custs=600, supps=60)
    write(T.parquet(r), tmp_path / "data")
    db = T.build(r)
    requests = ['7 "FRANCE" "GERMANY"', '7 GERMANY FRANCE', '7 "JAPAN" "JAPAN"',
                '7 "ATLANTIS" "FRANCE"', '8 "BRAZIL" "AMERICA" "ECONOMY ANODIZED STEEL"',
                '8 "CHINA" "ASIA" "PROMO BRUSHED TIN"', '8 "CHINA" "ASIA" "NO SUCH TYPE"',
                '8 "CHINA" "NOWHERE" "PROMO BRUSHED TIN"', '9 "green"', '9 "puff"', '9 "zzz"']
    res = run_driver(tmp_path / "data", tmp_path / "run", requests)
    assert res.returncode == 0, res.stderr
    want = [csv_q7(db, r, "FRANCE", "GERMANY"), csv_q7(db, r, "GERMANY", "FRANCE"),
            csv_q7(db, r, "JAPAN", "JAPAN"), csv_q7(db, r, "ATLANTIS", "FRANCE"),
            csv_q8(db, r, "BRAZIL", "AMERICA", "ECONOMY ANODIZED STEEL"),
            csv_q8(db, r, "CHINA", "ASIA", "PROMO BRUSHED TIN"),
            csv_q8(db, r, "CHINA", "ASIA", "NO SUCH TYPE"),
            csv_q8(db, r, "CHINA", "NOWHERE", "PROMO BRUSHED TIN"),
            csv_q9(db, r, "green"), csv_q9(db, r, "puff"), csv_q9(db, r, "zzz")]
    for i, w in enumerate(want):
        assert (tmp_path / "run" / f"result{i + 1}.csv").read_text() == w, requests[i]
    if orders > 3:
        assert want[0].count("\n") == 5 and want[4].count("\n") == 3
        assert want[8].count("\n") > 100


def csv_q13(db, w1, w2):
    import oracle
    mx, dist = oracle.q13(db, w1, w2)
    rows = sorted(((int(d), k) for k, d in enumerate(dist.tolist()) if d), reverse=True)
    return "c_count,custdist\n" + "".join(f"{k},{d}\n" for d, k in rows)


def csv_q16(db, brand, tprefix, sizes):
    import oracle
    out = "p_brand,p_type,p_size,supplier_cnt\n"
    mask, any_ = 0, False
    for s_ in sizes:
        if s_ == "<<NULL>>" or not s_.isdigit() and s_ != "":
            continue
        any_ = True
        n = int(s_) if s_ else 0
        if n <= 50:
            mask |= 1 << n
    if not any_:
        return out
    bb = idx(db["brand_dict"], brand)
    tb = np.array([t.startswith(tprefix) for t in db["type_dict"]], np.uint8)
    counts = oracle.q16(db, bb, tb, mask, int(db["p_partkey"].max()) + 1,
                        int(db["s_suppkey"].max()) + 1)
    nt = len(db["type_dict"])
    rows = []
    for g in np.nonzero(counts)[0].tolist():
        b, rest = divmod(g, nt * 51)
        rows.append((-int(counts[g]), db["brand_dict"][b], db["type_dict"][rest // 51], rest % 51))
    return out + "".join(f"{esc(b)},{esc(t)},{s_},{-c}\n" for c, b, t, s_ in sorted(rows))


# ------------------------------------------------- Q2, Q11, Q15, Q20
def _nation_of_supplier(db, r):
    """n_name of every supplier row ("" when its nation key is outside the
    nation rows, as query_q2.cpp:201-209 leaves the view empty)."""
    names = [""] * len(r["n_name"])
    for k, n in zip(r["n_nationkey"].tolist(), r["n_name"]):
        if 0 <= k < len(names):
            names[k] = n
    return [names[k] if 0 <= k < len(names) else "" for k in db["s_nationkey"].tolist()]


def csv_q2(db, r, size, suffix, region):
    import oracle
    out = "s_acctbal,s_name,n_name,p_partkey,p_mfgr,s_address,s_phone,s_comment\n"
    if region not in r["r_name"]:
        return out
    reg = r["r_name"].index(region)
    in_region = {k for k, rk in zip(r["n_nationkey"].tolist(), r["n_regionkey"].tolist())
                 if rk == reg and 0 <= k < len(r["n_name"])}
    s_ok = np.array([k in in_region for k in db["s_nationkey"].tolist()], np.uint8)
    t_ok = np.array([t.endswith(suffix) for t in db["type_dict"]], np.uint8)
    pairs = oracle.q2(db, t_ok, int(size), s_ok)
    nname = _nation_of_supplier(db, r)
    rows = []
    for prow, srow in pairs.tolist():
        rows.append((-int(db["s_acctbal"][srow]), nname[srow], r["s_name"][srow], prow + 1,
                     db["mfgr_dict"][db["p_mfgr"][prow]], r["s_address"][srow], r["s_phone"][srow],
                     r["s_comment"][srow]))
    rows.sort(key=lambda x: x[:4])
    return out + "".join(f"{-a / 100.0:.2f},{esc(sn)},{esc(nn)},{pk},{esc(mf)},{esc(ad)},"
                         f"{esc(ph)},{esc(cm)}\n" for a, nn, sn, pk, mf, ad, ph, cm in rows)


def csv_q11(db, r, nation, fraction):
    import oracle
    out = "ps_partkey,value\n"
    if nation not in r["n_name"]:
        return out
    pk, val, _, _ = oracle.q11(db, r["n_nationkey"][r["n_name"].index(nation)], float(fraction))
    return out + "".join(f"{k},{v / 100.0:.2f}\n" for k, v in zip(pk.tolist(), val.tolist()))


def csv_q15(db, r, date):
    import oracle
    from tests import tpch_data as T
    rows, _ = oracle.q15(db, T.date_offset(date), T.add_months(date, 3))
    return "s_suppkey,s_name,s_address,s_phone,total_revenue\n" + "".join(
        f"{k},{esc(r['s_name'][row])},{esc(r['s_address'][row])},{esc(r['s_phone'][row])},"
        f"{v / 100.0:.2f}\n" for k, row, v in rows)


def csv_q20(db, r, color, date, nation):
    import oracle
    from tests import tpch_data as T
    out = "s_name,s_address\n"
    if "<<NULL>>" in (color, date, nation) or nation not in r["n_name"]:
        return out
    nk = int(r["n_nationkey"][r["n_name"].index(nation)])
    p_ok = np.array([n.startswith(color) for n in r["p_name"]], np.uint8)
    q, _, _ = oracle.q20(db, p_ok, T.date_offset(date), T.add_months(date, 12), nk)
    rows = sorted((r["s_name"][i], r["s_address"][i]) for i, (k, n) in
                  enumerate(zip(db["s_suppkey"].tolist(), db["s_nationkey"].tolist()))
                  if 0 <= k < len(q) and q[k] and n == nk)
    return out + "".join(f"{esc(a)},{esc(b)}\n" for a, b in rows)


@pytest.mark.parametrize("orders", [3, 40_001])
def test_driver_q13_q16(tmp_path, orders):
    if not B.DRIVER.exists():
        pytest.skip("bolap_tpch not built (no Arrow C++)")
    from tests import tpch_data as T
    r = T.raw(37 + orders, orders, parts=900, custs=600, supps=60)
    write(T.parquet(r), tmp_path / "data")
    db = T.build(r)
    s1 = ["49", "14", "23", "45", "19", "3", "36", "9"]
    s2 = ["1", "2", "<<NULL>>", "x7", "3", "4", "5", "6"]
    q = lambda v: " ".join(f'"{x}"' for x in v)
    requests = ['13 "special" "requests"', '13 pending deposits', '13 "x" ""',
                f'16 "Brand#45" "MEDIUM POLISHED" {q(s1)}', f'16 "Brand#11" "PROMO" {q(s2)}',
                f'16 "Brand#11" "PROMO" {q(["<<NULL>>"] * 8)}']
    res = run_driver(tmp_path / "data", tmp_path / "run", requests)
    assert res.returncode == 0, res.stderr
    want = [csv_q13(db, "special", "requests"), csv_q13(db, "pending", "deposits"),
            csv_q13(db, "x", ""), csv_q16(db, "Brand#45", "MEDIUM POLISHED", s1),
            csv_q16(db, "Brand#11", "PROMO", s2), csv_q16(db, "Brand#11", "PROMO", ["<<NULL>>"] * 8)]
    for i, w in enumerate(want):
        assert (tmp_path / "run" / f"result{i + 1}.csv").read_text() == w, requests[i]
    if orders > 3:
        assert want[0].count("\n") > 5 and want[3].count("\n") > 20


@pytest.mark.parametrize("orders", [3, 40_001])
def test_driver_q2_q11_q15_q20(tmp_path, orders):
    """The supplier-side pipelines through load / build / query, against the
    oracle plus Python restatements of the reference writers."""
    if not B.DRIVER.exists():
        pytest.skip("bolap_tpch not built (no Arrow C++)")
    from tests import tpch_data as T
    r = T.raw(43 + orders, orders, parts=900, custs=600, supps=120)
    write(T.parquet(r), tmp_path / "data")
    db = T.build(r)
    requests = ['2 "15" "BRASS" "EUROPE"', '2 3 TIN ASIA', '2 "7" "" "AFRICA"',
                '2 "15" "BRASS" "ATLANTIS"', '11 "GERMANY" "0.0001"', '11 PERU 0.003',
                '11 "ATLANTIS" "0.0001"', '11 "FRANCE" "-1"', '15 "1996-01-01" "0"',
                '15 1993-11-30 7', '15 "2010-01-01" "0"', '20 "forest" "1994-01-01" "CANADA"',
                '20 "" "1995-01-01" "GERMANY"', '20 "g" "1996-02-29" "ATLANTIS"',
                '20 "<<NULL>>" "1994-01-01" "CANADA"', '20 "b" "1993-01-01" "CHINA"']
    res = run_driver(tmp_path / "data", tmp_path / "run", requests)
    assert res.returncode == 0, res.stderr
    want = [csv_q2(db, r, "15", "BRASS", "EUROPE"), csv_q2(db, r, "3", "TIN", "ASIA"),
            csv_q2(db, r, "7", "", "AFRICA"), csv_q2(db, r, "15", "BRASS", "ATLANTIS"),
            csv_q11(db, r, "GERMANY", "0.0001"), csv_q11(db, r, "PERU", "0.003"),
            csv_q11(db, r, "ATLANTIS", "0.0001"), csv_q11(db, r, "FRANCE", "-1"),
            csv_q15(db, r, "1996-01-01"), csv_q15(db, r, "1993-11-30"),
            csv_q15(db, r, "2010-01-01"), csv_q20(db, r, "forest", "1994-01-01", "CANADA"),
            csv_q20(db, r, "", "1995-01-01", "GERMANY"),
            csv_q20(db, r, "g", "1996-02-29", "ATLANTIS"),
            csv_q20(db, r, "<<NULL>>", "1994-01-01", "CANADA"),
            csv_q20(db, r, "b", "1993-01-01", "CHINA")]
    for i, w in enumerate(want):
        assert (tmp_path / "run" / f"result{i + 1}.csv").read_text() == w, requests[i]
    if orders > 3:
        assert all(want[i].count("\n") > 2 for i in (0, 4, 7, 12))
        assert want[8].count("\n") == 2


@pytest.mark.parametrize("orders", [3, 20_001])
def test_driver_matches_the_reference_pipelines(tmp_path, orders):
    """Every served id through bolap_tpch (Parquet -> device mirror -> GPU
    pipelines -> CSV) against the reference's own run_qN + write_qN_csv
    (oracle/_ref, compiled from its sources) over the same tables, byte for
    byte.  Ids the driver does not serve get the reference's stub file."""
    if not B.DRIVER.exists():
        pytest.skip("bolap_tpch not built (no Arrow C++)")
    from tests import tpch_data as T
    from tests import tpch_ref as R
    if not R.available():
        pytest.skip("oracle/_ref not built")
    from tests.test_ref_tpch import Q16_S1, Q22_A, q22_defined, quoted
    r = T.raw(61 + orders, orders, parts=700, custs=901, supps=60)
    write(T.parquet(r), tmp_path / "data")
    requests = ['1 90', '1 "3"', '3 "BUILDING" "1995-03-15"', '3 "MACHINERY" "1993-01-01"',
                '4 "1995-01-01"', '5 "ASIA" "1994-01-01"', '6 "1994-01-01" "0.06" "24"',
                '6 "1997-01-01" "0.01" "40"', '7 "FRANCE" "GERMANY"',
                '8 "BRAZIL" "AMERICA" "ECONOMY ANODIZED STEEL"', '9 "green"', '10 "1993-10-01"',
                '12 "MAIL" "SHIP" "1994-01-01"', '13 "special" "requests"', '14 "1995-09-01"',
                f'16 "Brand#45" "MEDIUM POLISHED" {quoted(Q16_S1)}', '17 "Brand#23" "MED BOX"',
                '17 "Brand#11" "SM CASE"', '18 "200"', '18 "300"',
                '19 "1" "10" "20" "Brand#12" "Brand#23" "Brand#34"', '21 "SAUDI ARABIA"',
                f"22 {quoted(Q22_A)}", '2 "15" "BRASS" "EUROPE"', '11 "GERMANY" "0.0001"',
                '15 "1996-01-01" "0"', '20 "forest" "1994-01-01" "CANADA"',
                '20 "" "1995-01-01" "GERMANY"', '99 "no such query"']
    served = [q for q in requests if q.split()[0] != "99"]
    if not q22_defined(r):
        served = [q for q in served if not q.startswith("22 ")]
    res = run_driver(tmp_path / "data", tmp_path / "run", requests)
    assert res.returncode == 0, res.stderr
    got = {q: (tmp_path / "run" / f"result{i + 1}.csv").read_text()
           for i, q in enumerate(requests)}
    want = R.RefTpch(r).run(served)
    for q, w in zip(served, want):
        assert got[q] == w, q
    for q in requests:
        if q not in served and not q.startswith("22 "):
            assert got[q] == stub_csv(q.split()[0], q), q
    if orders > 3:
        assert sum(w.count("\n") > 1 for w in want) >= len(served) - 4

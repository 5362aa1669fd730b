"""Q4, Q5, Q10, Q12 and Q19 on the GPU (tpch_orders.hip) bit-exact against
the oracle (oracle/tpch2.c) over the same reference-layout columns: small,
medium and SF10-sized databases (15 M orders, ~60 M lineitems), lineitem in
order-key order and shuffled (the unsorted branches), and the edge cases the
reference handles (empty windows, absent dictionary codes, duplicate part
keys, lineitem keys outside the part and order maps)."""
import numpy as np
import pytest
import torch

import oracle
from tests import tpch_data as T
from tests.test_tpch2_oracle import q19_groups

pytestmark = pytest.mark.gpu

U8_CODES = ("l_shipmode", "l_shipinstruct", "o_orderpriority", "o_orderstatus", "p_type")


def upload(db: dict) -> dict:
    t = {}
    for k, v in db.items():
        if not isinstance(v, np.ndarray):
            t[k] = v
            continue
        a = v.astype(np.uint8) if k in U8_CODES or k in ("p_brand", "p_container") else v
        if a.dtype == np.uint32:
            a = a.view(np.int32)
        elif a.dtype == np.uint16:
            a = a.view(np.int16)
        t[k] = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t


def shuffled(r: dict, seed: int) -> dict:
    """The same database with lineitem rows shuffled within blocks of 64."""
    r2 = dict(r)
    n = len(r["l_orderkey"])
    rng = np.random.default_rng(seed)
    perm = np.argsort(np.arange(n) // 64 + rng.random(n), kind="stable")
    for k in list(r):
        if k.startswith("l_") and isinstance(r[k], np.ndarray) and len(r[k]) == n:
            r2[k] = r[k][perm]
    return r2


CASES = {"small": (5, 4000, 800, 300, 40), "medium": (6, 300_000, 20_000, 30_000, 2000)}


@pytest.fixture(scope="module", params=["small", "medium", "small-unsorted"])
def db(request):
    name = request.param.split("-")[0]
    seed, orders, parts, custs, supps = CASES[name]
    r = T.raw(seed, orders, parts=parts, custs=custs, supps=supps, strings=False)
    if request.param.endswith("unsorted"):
        r = shuffled(r, seed)
    h = T.build(r)
    if request.param.endswith("unsorted"):
        assert h["l_sorted"] == 0
    return r, h, upload(h)


WINDOWS = [("1993-07-01", 3), ("1995-01-01", 12), ("1998-07-01", 3), ("2010-01-01", 3),
           ("1990-01-01", 3)]


@pytest.mark.parametrize("date,months", WINDOWS)
def test_q4(engine, db, date, months):
    r, h, t = db
    lo, hi = T.date_offset(date), T.add_months(date, months)
    n = len(h["priority_dict"])
    assert engine.q4(t, n, lo, hi) == oracle.q4(h, lo, hi, n).tolist()


@pytest.mark.parametrize("date,months", WINDOWS)
@pytest.mark.parametrize("region", [2, 3])
def test_q5(engine, db, date, months, region):
    r, h, t = db
    lo, hi = T.date_offset(date), T.add_months(date, months)
    inreg = (np.array(h["n_regionkey"]) == region).astype(np.uint8)
    assert engine.q5(t, inreg, lo, hi) == oracle.q5(h, inreg, lo, hi).tolist()


@pytest.mark.parametrize("date,months", WINDOWS)
def test_q10(engine, db, date, months):
    r, h, t = db
    lo, hi = T.date_offset(date), T.add_months(date, months)
    nls = len(h["ls_dict"])
    rc = h["rf_dict"].index("R") * nls
    got = engine.q10(t, rc, nls, lo, hi, 25)
    want = oracle.q10(h, rc, nls, lo, hi, 25)
    assert got == [tuple(x) for x in want.tolist()]
    # no 'R' in the dictionary: no rows
    assert engine.q10(t, -1, nls, lo, hi, 25) == []


def test_q10_customers_without_rows_or_nation(engine, db):
    r, h, t = db
    h2 = dict(h)
    # drop every tenth customer row and give some a nation key out of range
    keep = np.arange(len(h["c_custkey"])) % 10 != 3
    h2["c_custkey"] = np.ascontiguousarray(h["c_custkey"][keep])
    nk = h["c_nationkey"][keep].copy()
    nk[::7] = 30
    h2["c_nationkey"] = nk
    t2 = dict(t, c_custkey=torch.from_numpy(h2["c_custkey"]).cuda(),
              c_nationkey=torch.from_numpy(nk).cuda())
    nls = len(h["ls_dict"])
    rc = h["rf_dict"].index("R") * nls
    lo, hi = T.date_offset("1994-01-01"), T.add_months("1994-01-01", 12)
    got = engine.q10(t2, rc, nls, lo, hi, 25)
    want = oracle.q10(h2, rc, nls, lo, hi, 25)
    assert got == [tuple(x) for x in want.tolist()] and len(got) > 0


@pytest.mark.parametrize("modes,date", [(("MAIL", "SHIP"), "1994-01-01"),
                                        (("AIR", "TRUCK"), "1997-01-01"),
                                        (("FOB", "NOSUCH"), "1995-01-01"),
                                        (("NOSUCH", "NOPE"), "1995-01-01")])
def test_q12(engine, db, modes, date):
    r, h, t = db
    lo, hi = T.date_offset(date), T.add_months(date, 12)
    codes = [h["shipmode_dict"].index(m) if m in h["shipmode_dict"] else -1 for m in modes]
    hc = [h["priority_dict"].index(p) for p in ("1-URGENT", "2-HIGH")]
    got = engine.q12(t, codes[0], codes[1], lo, hi, *hc)
    assert got == oracle.q12(h, codes[0], codes[1], lo, hi, *hc)


@pytest.mark.parametrize("brands,qtys", [(("Brand#12", "Brand#23", "Brand#34"), (1, 10, 20)),
                                         (("Brand#11", None, "Brand#55"), (5, 3, 25)),
                                         (("Brand#99", "Brand#21", "Brand#21"), (1, 12, 14)),
                                         ((None, None, None), (1, 2, 3))])
def test_q19(engine, db, brands, qtys):
    r, h, t = db
    g = q19_groups(h, brands, qtys)
    m1, m2 = h["shipmode_dict"].index("AIR"), h["shipmode_dict"].index("REG AIR")
    ins = h["instruct_dict"].index("DELIVER IN PERSON")
    want = oracle.q19(h, g, m1, m2, ins)
    assert engine.q19(t, g, m1, m2, ins) == want
    assert engine.q19(t, g, m1, -1, ins) == oracle.q19(h, g, m1, -1, ins)
    assert engine.q19(t, g, -1, -1, ins) == 0


def test_q19_duplicate_part_keys_and_keys_outside_the_parts(engine, db):
    r, h, t = db
    h2 = dict(h)
    n = len(h["p_partkey"])
    # duplicate the first 100 part keys with other brands (last row wins),
    # and point some lineitems past the largest part key
    h2["p_partkey"] = np.concatenate([h["p_partkey"], h["p_partkey"][:100]])
    h2["p_brand"] = np.concatenate([h["p_brand"], (h["p_brand"][:100] + 1) % 25]).astype(
        np.uint16)
    h2["p_container"] = np.concatenate([h["p_container"], h["p_container"][:100]])
    h2["p_size"] = np.concatenate([h["p_size"], h["p_size"][:100]])
    lk = h["l_partkey"].copy()
    lk[::97] = n + 5
    h2["l_partkey"] = lk
    t2 = upload(h2)
    g = q19_groups(h, ("Brand#12", "Brand#23", "Brand#34"), (1, 10, 20))
    m1, m2 = h["shipmode_dict"].index("AIR"), h["shipmode_dict"].index("REG AIR")
    ins = h["instruct_dict"].index("DELIVER IN PERSON")
    assert engine.q19(t2, g, m1, m2, ins) == oracle.q19(h2, g, m1, m2, ins)
    # int16 brand / container codes give the same answer
    t3 = dict(t2, p_brand=t2["p_brand"].to(torch.int16), p_container=t2["p_container"].to(torch.int16))
    assert engine.q19(t3, g, m1, m2, ins) == oracle.q19(h2, g, m1, m2, ins)


@pytest.mark.parametrize("nation", [3, 12, 24, 99])
def test_q21(engine, db, nation):
    r, h, t = db
    n = int(h["s_suppkey"].max()) + 1
    f = h["status_dict"].index("F") if "F" in h["status_dict"] else -1
    assert engine.q21(t, f, nation, n) == oracle.q21(h, f, nation, n).tolist()


def test_q21_duplicate_supplier_rows_and_absent_status(engine, db):
    r, h, t = db
    h2 = dict(h)
    # the last row of a duplicated key decides its nation
    h2["s_suppkey"] = np.concatenate([h["s_suppkey"], h["s_suppkey"][:5]])
    h2["s_nationkey"] = np.concatenate([h["s_nationkey"], (h["s_nationkey"][:5] + 1) % 25])
    t2 = dict(t, s_suppkey=torch.from_numpy(h2["s_suppkey"]).cuda(),
              s_nationkey=torch.from_numpy(h2["s_nationkey"].astype(np.int32)).cuda())
    n = int(h["s_suppkey"].max()) + 1
    f = h["status_dict"].index("F")
    for nation in range(25):
        assert engine.q21(t2, f, nation, n) == oracle.q21(h2, f, nation, n).tolist()
    assert engine.q21(t, -1, 3, n) == [0] * n          # no "F" status
    assert engine.q21(t, f, 3, n // 2) == oracle.q21(h, f, 3, n // 2).tolist()


@pytest.mark.parametrize("codes", [["13", "31", "23", "29", "30", "18", "17"], ["AB", "20", "7"],
                                   ["AB"], ["99"]])
def test_q22(engine, db, codes):
    r, h, t = db
    numeric = [int(c) for c in codes if len(c) >= 2 and c[:2].isdigit()]
    extra = [(ord(c[0]) << 8) | ord(c[1]) for c in codes if len(c) >= 2 and not c[:2].isdigit()]
    num, ext = oracle.q22(h, numeric, extra)
    assert engine.q22(t, numeric, extra) == (num.tolist(), ext.tolist())


Q7_DATES = (T.date_offset("1995-01-01"), T.date_offset("1996-12-31"), T.date_offset("1996-01-01"))


@pytest.mark.parametrize("n1,n2", [(6, 7), (7, 6), (12, 18), (3, 3), (3, 40)])
def test_q7(engine, db, n1, n2):
    r, h, t = db
    assert engine.q7(t, n1, n2, *Q7_DATES) == oracle.q7(h, n1, n2, *Q7_DATES).reshape(-1).tolist()


@pytest.mark.parametrize("nation,region,ptype", [(8, 2, "ECONOMY ANODIZED STEEL"),
                                                 (3, 1, "PROMO BRUSHED TIN"), (3, 1, "NOSUCH")])
def test_q8(engine, db, nation, region, ptype):
    r, h, t = db
    inreg = (np.array(h["n_regionkey"]) == region).astype(np.uint8)
    tc = h["type_dict"].index(ptype) if ptype in h["type_dict"] else -1
    tot, nat = oracle.q8(h, inreg, nation, tc, *Q7_DATES)
    assert engine.q8(t, inreg, nation, tc, *Q7_DATES) == (tot.tolist(), nat.tolist())


@pytest.mark.parametrize("color", ["green", "puff", "tan"])
def test_q9(engine, db, color):
    r, h, t = db
    p_ok = (h["p_name_words"] == T.COLORS.index(color)).any(axis=1).astype(np.uint8)
    y0, tot, used = oracle.q9(h, p_ok, 25)
    gy, gtot, gused = engine.q9(t, torch.from_numpy(p_ok).cuda(), 25)
    assert gy == y0 and np.array_equal(gtot, tot) and np.array_equal(gused, used)


@pytest.mark.parametrize("order", ["sorted", "shuffled"])
def test_q9_partsupp_duplicates_and_order(engine, db, order):
    """partsupp with duplicate (part, supplier) pairs of different cost (the
    first row in row order wins), keys past the part range and negative keys,
    kept sorted by partkey (contiguous key ranges) or shuffled (the CSR)."""
    r, h, t = db
    rng = np.random.default_rng(9)
    n = len(h["ps_partkey"])
    dup = rng.choice(n, n // 5, replace=False)
    past = int(h["p_partkey"].max()) + 7
    pk = np.concatenate([h["ps_partkey"], h["ps_partkey"][dup], [-4, past]]).astype(np.int32)
    sk = np.concatenate([h["ps_suppkey"], h["ps_suppkey"][dup], [1, 1]]).astype(np.int32)
    cost = np.concatenate([h["ps_supplycost"], rng.integers(1, 10**5, len(dup)),
                           [5, 5]]).astype(np.int32)
    perm = (np.argsort(pk, kind="stable") if order == "sorted"
            else rng.permutation(len(pk)))
    h2 = dict(h, ps_partkey=pk[perm], ps_suppkey=sk[perm], ps_supplycost=cost[perm])
    t2 = dict(t)
    for k in ("ps_partkey", "ps_suppkey", "ps_supplycost"):
        t2[k] = torch.from_numpy(np.ascontiguousarray(h2[k])).cuda()
    p_ok = (h["p_name_words"] == T.COLORS.index("green")).any(axis=1).astype(np.uint8)
    y0, tot, used = oracle.q9(h2, p_ok, 25)
    gy, gtot, gused = engine.q9(t2, torch.from_numpy(p_ok).cuda(), 25)
    assert used.any()
    assert gy == y0 and np.array_equal(gtot, tot) and np.array_equal(gused, used)


@pytest.mark.parametrize("w1,w2", [("special", "requests"), ("pending", "deposits"),
                                   ("express", ""), ("x", "x"), ("", "")])
def test_q13(engine, db, w1, w2):
    r, h, t = db
    mx, dist = oracle.q13(h, w1, w2)
    masks = engine.string_masks(t["o_comment_chars"], t["o_comment_offsets"])
    assert engine.q13(t, w1, w2) == (mx, dist.tolist())
    assert engine.q13(t, w1, w2, masks=masks) == (mx, dist.tolist())


def test_q13_comment_edges(engine):
    """Comments built to hit the search's edges: words at the first and last
    byte, word2 overlapping word1 or only before it, repeated prefixes
    ("spespecial"), empty comments, and comments of 80..300 bytes (past the
    register window, searched in memory), at every alignment."""
    rng = np.random.default_rng(13)
    parts = ["special", "requests", "spe", "spespecial", "requestsrequests", "x", "", " ",
             "specia", "equests", "specialrequests", "requests special", "pending", "deposits"]
    comments = []
    for i in range(6000):
        k = int(rng.integers(0, 6))
        s = "".join(parts[int(j)] + (" " if rng.random() < 0.5 else "")
                    for j in rng.integers(0, len(parts), k))
        if i % 7 == 0:
            s = "z" * int(rng.integers(60, 290)) + s
        elif i % 11 == 0:
            s = s + "q" * int(rng.integers(60, 290))
        comments.append(s.encode())
    lens = np.array([len(c) for c in comments])
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    chars = np.frombuffer(b"".join(comments) + b"\0", np.uint8)[:-1].copy()
    ck = rng.integers(1, 400, len(comments)).astype(np.int32)
    h = {"o_custkey_raw": ck, "o_comment_chars": chars, "o_comment_offsets": off,
         "c_custkey": np.arange(1, 401, dtype=np.int32)}
    assert (lens > 100).sum() > 500 and (lens == 0).sum() > 10
    t = {k: torch.from_numpy(v.view(np.int32) if v.dtype == np.uint32 else v).cuda()
         for k, v in h.items()}
    masks = engine.string_masks(t["o_comment_chars"], t["o_comment_offsets"])
    for w1, w2 in [("special", "requests"), ("spe", "special"), ("requests", "special"),
                   ("x", ""), ("", "x"), ("", ""), ("special", ""), ("s", "s"),
                   ("specialrequests", "requests"), ("z" * 60, "special")]:
        want = oracle.q13(h, w1, w2)
        assert engine.q13(t, w1, w2) == (want[0], want[1].tolist()), (w1, w2)
        assert engine.q13(t, w1, w2, masks=masks) == (want[0], want[1].tolist()), (w1, w2)


def test_string_masks_match_the_builder(engine, db):
    r, h, t = db
    a, b = engine.string_masks(t["o_comment_chars"], t["o_comment_offsets"])
    a, b = a.cpu().numpy().view(np.uint32), b.cpu().numpy().view(np.uint64)
    ch, off = h["o_comment_chars"].tobytes(), h["o_comment_offsets"]
    for i in range(0, len(off) - 1, max(1, (len(off) - 1) // 500)):
        s_ = ch[off[i]:off[i + 1]]
        al = 0
        bg = 0
        for j, c in enumerate(s_):
            if 97 <= c <= 122:
                al |= 1 << (c - 97)
            elif 65 <= c <= 90:
                al |= 1 << (c - 65)
            if j:
                bg |= 1 << ((s_[j - 1] * 131 + c) & 63)
        assert int(a[i]) == al and int(b[i]) == bg, i


@pytest.mark.parametrize("brand,tprefix,sizes", [("Brand#45", "MEDIUM POLISHED", [49, 14, 23, 45,
                                                                                  19, 3, 36, 9]),
                                                 ("Brand#11", "PROMO", [1, 2, 3]),
                                                 ("Brand#99", "ZZZ", [50, 0])])
def test_q16(engine, db, brand, tprefix, sizes):
    r, h, t = db
    bb = h["brand_dict"].index(brand) if brand in h["brand_dict"] else -1
    tb = np.array([x.startswith(tprefix) for x in h["type_dict"]], np.uint8)
    mask = sum(1 << s_ for s_ in sizes)
    npk, nsk = int(h["p_partkey"].max()) + 1, int(h["s_suppkey"].max()) + 1
    want = oracle.q16(h, bb, tb, mask, npk, nsk)
    assert np.array_equal(engine.q16(t, len(h["brand_dict"]), bb, tb, mask, npk, nsk), want)


@pytest.fixture(scope="module")
def sf10():
    """15,000,000 orders, about 60,000,000 lineitems, 2,000,000 parts,
    1,500,000 customers, 100,000 suppliers."""
    r = T.raw(42, 15_000_000, parts=2_000_000, custs=1_500_000, supps=100_000, strings=False)
    h = T.build(r)
    del r
    return h, upload(h)


def test_sf10_all_five(engine, sf10):
    h, t = sf10
    assert len(h["l_orderkey"]) > 55_000_000
    lo, hi = T.date_offset("1994-01-01"), T.add_months("1994-01-01", 3)
    print("q4", flush=True)
    assert engine.q4(t, 5, lo, hi) == oracle.q4(h, lo, hi, 5).tolist()
    print("q5", flush=True)
    inreg = (np.array(h["n_regionkey"]) == 2).astype(np.uint8)
    assert engine.q5(t, inreg, lo, T.add_months("1994-01-01", 12)) == \
        oracle.q5(h, inreg, lo, T.add_months("1994-01-01", 12)).tolist()
    print("q10", flush=True)
    nls = len(h["ls_dict"])
    rc = h["rf_dict"].index("R") * nls
    got = engine.q10(t, rc, nls, lo, hi, 25)
    want = oracle.q10(h, rc, nls, lo, hi, 25)
    assert len(got) > 100_000
    assert got == [tuple(x) for x in want.tolist()]
    print("q12", flush=True)
    m = [h["shipmode_dict"].index(x) for x in ("MAIL", "SHIP")]
    hc = [h["priority_dict"].index(p) for p in ("1-URGENT", "2-HIGH")]
    yr = T.add_months("1994-01-01", 12)
    assert engine.q12(t, m[0], m[1], lo, yr, *hc) == oracle.q12(h, m[0], m[1], lo, yr, *hc)
    print("q19", flush=True)
    g = q19_groups(h, ("Brand#12", "Brand#23", "Brand#34"), (1, 10, 20))
    m1, m2 = h["shipmode_dict"].index("AIR"), h["shipmode_dict"].index("REG AIR")
    ins = h["instruct_dict"].index("DELIVER IN PERSON")
    want19 = oracle.q19(h, g, m1, m2, ins)
    assert want19 > 0
    assert engine.q19(t, g, m1, m2, ins) == want19
    print("q7 q8 q9", flush=True)
    assert engine.q7(t, 6, 7, *Q7_DATES) == oracle.q7(h, 6, 7, *Q7_DATES).reshape(-1).tolist()
    inreg = (np.array(h["n_regionkey"]) == 1).astype(np.uint8)
    tc = h["type_dict"].index("ECONOMY ANODIZED STEEL")
    tot, nat = oracle.q8(h, inreg, 2, tc, *Q7_DATES)
    assert tot.sum() > 0
    assert engine.q8(t, inreg, 2, tc, *Q7_DATES) == (tot.tolist(), nat.tolist())
    p_ok = (h["p_name_words"] == T.COLORS.index("green")).any(axis=1).astype(np.uint8)
    y0, tot9, used9 = oracle.q9(h, p_ok, 25)
    gy, gtot, gused = engine.q9(t, torch.from_numpy(p_ok).cuda(), 25)
    assert gy == y0 and np.array_equal(gtot, tot9) and np.array_equal(gused, used9)
    print("q13 q16", flush=True)
    mx, dist = oracle.q13(h, "special", "requests")
    masks = engine.string_masks(t["o_comment_chars"], t["o_comment_offsets"])
    assert engine.q13(t, "special", "requests", masks=masks) == (mx, dist.tolist())
    tb = np.array([x.startswith("MEDIUM POLISHED") for x in h["type_dict"]], np.uint8)
    mask = sum(1 << s_ for s_ in [49, 14, 23, 45, 19, 3, 36, 9])
    npk, nsk = int(h["p_partkey"].max()) + 1, int(h["s_suppkey"].max()) + 1
    want16 = oracle.q16(h, h["brand_dict"].index("Brand#45"), tb, mask, npk, nsk)
    assert want16.sum() > 0
    assert np.array_equal(engine.q16(t, len(h["brand_dict"]), h["brand_dict"].index("Brand#45"),
                                     tb, mask, npk, nsk), want16)
    print("q21", flush=True)
    n = int(h["s_suppkey"].max()) + 1
    w21 = oracle.q21(h, h["status_dict"].index("F"), 7, n)
    assert w21.sum() > 0
    assert engine.q21(t, h["status_dict"].index("F"), 7, n) == w21.tolist()
    print("q22", flush=True)
    num, ext = oracle.q22(h, [13, 31, 23, 29, 30, 18, 17], [(ord("A") << 8) | ord("B")])
    assert num[:, 0].sum() > 0
    assert engine.q22(t, [13, 31, 23, 29, 30, 18, 17], [(ord("A") << 8) | ord("B")]) == \
        (num.tolist(), ext.tolist())
    print("q2 q11 q15 q20", flush=True)
    nation_ok = (np.array(h["n_regionkey"]) == 3).astype(np.uint8)
    t_ok = np.array([x.endswith("BRASS") for x in h["type_dict"]], np.uint8)
    want2 = oracle.q2(h, t_ok, 15, nation_ok[h["s_nationkey"]])
    assert len(want2) > 1000
    np.testing.assert_array_equal(engine.q2(t, t_ok, 15, nation_ok), want2)
    nk = int(h["s_suppkey"].max()) + 1
    pk11, val11, tot11 = engine.q11(t, 7, nk, 0.0001 / 10)
    w11 = oracle.q11(h, 7, 0.0001 / 10)
    assert len(w11[0]) > 100 and tot11 == w11[2]
    assert np.array_equal(pk11, w11[0]) and np.array_equal(val11, w11[1])
    lo15, hi15 = T.date_offset("1996-01-01"), T.add_months("1996-01-01", 3)
    assert engine.q15(t, lo15, hi15, nk) == oracle.q15(h, lo15, hi15)
    p_ok = np.isin(h["p_name_words"][:, 0], [T.COLORS.index("forest")]).astype(np.uint8)
    smap = np.full(nk, -1, np.int16)
    smap[h["s_suppkey"]] = h["s_nationkey"]
    lo20, hi20 = T.date_offset("1994-01-01"), T.add_months("1994-01-01", 12)
    want20, n20, pairs20 = oracle.q20(h, p_ok, lo20, hi20, 3)
    assert n20 > 10
    for hint in (0, 64, 1 << 16):   # exact, short (rerun), generous
        q20, st20 = engine.q20(t, torch.from_numpy(p_ok).cuda(), int(h["p_partkey"].max()),
                               lo20, hi20, 3, torch.from_numpy(smap).cuda(), nk - 1,
                               capacity_hint=hint)
        assert np.array_equal(q20, want20) and (st20.qualified, st20.pairs) == (n20, pairs20)
        assert st20.fallbacks == (1 if hint == 64 else 0)

"""A seeded TPC-H-shaped database for the order-range, Q12 and Q19 tests.

`raw(seed, orders, ...)` draws the eight tables in TPC-H's value domains
(sparse order keys, nation/region keys, 1-7 lineitems per order in order-key
order, receiptdate = shipdate + 1..30, dictionary strings).  Dates are day
offsets from 1992-01-01 and decimals integers in hundredths, the units the
reference builder produces (bespoke_tpch/builder_impl.cpp:226-338).

`build(raw)` lays the columns out as the reference's build() does
(builder_impl.cpp:447-999): first-appearance dictionary codes, orders sorted
by date (stable), the orderkey -> row map (last row wins), lineitem ranges,
truncated discounted_price, the returnflag_linestatus code, packed
commit_receipt and the denormalised supplier / customer nation keys.  It is
written with numpy from that description; the GPU path builds the same
layout with its own kernels.

`parquet(raw)` gives the tables as the reference's Parquet inputs
(decimal(15,2), date32, strings) for the driver test.
"""
from __future__ import annotations

import datetime

import numpy as np

BASE = datetime.date(1992, 1, 1)
NATIONS = [("ALGERIA", 0), ("ARGENTINA", 1), ("BRAZIL", 1), ("CANADA", 1), ("EGYPT", 4),
           ("ETHIOPIA", 0), ("FRANCE", 3), ("GERMANY", 3), ("INDIA", 2), ("INDONESIA", 2),
           ("IRAN", 4), ("IRAQ", 4), ("JAPAN", 2), ("JORDAN", 4), ("KENYA", 0), ("MOROCCO", 0),
           ("MOZAMBIQUE", 0), ("PERU", 1), ("CHINA", 2), ("ROMANIA", 3), ("SAUDI ARABIA", 4),
           ("VIETNAM", 2), ("RUSSIA", 3), ("UNITED KINGDOM", 3), ("UNITED STATES", 1)]
REGIONS = ["AFRICA", "AMERICA", "ASIA", "EUROPE", "MIDDLE EAST"]
SEGMENTS = ["BUILDING", "AUTOMOBILE", "MACHINERY", "HOUSEHOLD", "FURNITURE"]
PRIORITIES = ["1-URGENT", "2-HIGH", "3-MEDIUM", "4-NOT SPECIFIED", "5-LOW"]
SHIPMODES = ["REG AIR", "AIR", "RAIL", "SHIP", "TRUCK", "MAIL", "FOB"]
INSTRUCTS = ["DELIVER IN PERSON", "COLLECT COD", "NONE", "TAKE BACK RETURN"]
BRANDS = [f"Brand#{i}{j}" for i in range(1, 6) for j in range(1, 6)]
CONTAINERS = [f"{a} {b}" for a in ("SM", "LG", "MED", "JUMBO", "WRAP")
              for b in ("CASE", "BOX", "BAG", "JAR", "PKG", "PACK", "CAN", "DRUM")]
TYPES = [f"{a} {b} {c}" for a in ("STANDARD", "SMALL", "MEDIUM", "LARGE", "ECONOMY", "PROMO")
         for b in ("ANODIZED", "BURNISHED", "PLATED", "POLISHED", "BRUSHED")
         for c in ("TIN", "NICKEL", "BRASS", "STEEL", "COPPER")]
MFGRS = [f"Manufacturer#{i}" for i in range(1, 6)]
CURRENT = 1263  # 1995-06-17: returnflag / linestatus cut-over
# part-name words; no word contains another, so "name contains colour" is
# "colour is one of the name's words"
COLORS = ["almond", "azure", "blush", "chiffon", "coral", "cyan", "forest", "ghost", "green",
          "honeydew", "khaki", "lavender", "magenta", "navy", "olive", "orchid", "peru", "plum",
          "puff", "sienna", "tan", "thistle", "violet", "wheat"]


# comment words (Q13 searches them, Q16 bans "Customer ... Complaints")
WORDS = ["special", "requests", "packages", "pending", "unusual", "express", "accounts",
         "deposits", "furiously", "quickly", "ironic", "final", "regular", "blithely", "carefully",
         "slyly", "bold", "even", "sleep", "haggle", "nag", "wake", "use", "boost", "x"]


def comments(rng, n: int, words_per: int = 6):
    """n comments of `words_per` words each (space separated, one trailing
    space) as (chars uint8, offsets uint32 of n + 1)."""
    wb = [w.encode() + b" " for w in WORDS]
    width = max(len(w) for w in wb)
    table = np.zeros((len(wb), width), np.uint8)
    wlen = np.array([len(w) for w in wb], np.int64)
    for i, w in enumerate(wb):
        table[i, :len(w)] = np.frombuffer(w, np.uint8)
    ids = rng.integers(0, len(WORDS), (n, words_per)).astype(np.int32)
    flat = ids.reshape(-1)
    L = wlen[flat]
    clen = L.reshape(n, words_per).sum(axis=1)
    offsets = np.zeros(n + 1, np.int64)
    np.cumsum(clen, out=offsets[1:])
    rep_ids = np.repeat(flat, L)
    starts = np.repeat(np.cumsum(L) - L, L)
    within = (np.arange(len(rep_ids), dtype=np.int64) - starts).astype(np.int32)
    chars = table[rep_ids, within]
    return chars, offsets.astype(np.uint32)


def okeys(n: int) -> np.ndarray:
    """TPC-H's sparse order keys: 8 of every 32."""
    i = np.arange(n, dtype=np.int64)
    return ((i // 8) * 32 + (i % 8) + 1).astype(np.int32)


def raw(seed: int, orders: int, parts: int = 2000, custs: int = 300, supps: int = 100,
        strings: bool = True) -> dict:
    rng = np.random.default_rng(seed)
    r = {"orders_n": orders, "parts": parts, "custs": custs, "supps": supps}
    # nation / region
    r["n_nationkey"] = np.arange(25, dtype=np.int32)
    r["n_name"] = [n for n, _ in NATIONS]
    r["n_regionkey"] = np.array([k for _, k in NATIONS], np.int32)
    r["r_regionkey"] = np.arange(5, dtype=np.int32)
    r["r_name"] = list(REGIONS)
    # customer
    r["c_custkey"] = np.arange(1, custs + 1, dtype=np.int32)
    r["c_nationkey"] = rng.integers(0, 25, custs).astype(np.int32)
    r["c_acctbal"] = rng.integers(-99_999, 999_999, custs).astype(np.int64)
    r["c_mktsegment"] = rng.integers(0, 5, custs)
    # phone prefixes: TPC-H's country code 10 + nation, except every 17th
    # customer ("AB-...", no digits) and every 31st (a one-character phone)
    code = (10 + r["c_nationkey"]).astype(np.uint8)
    pair = ((ord("0") + code // 10).astype(np.int32) << 8) | (ord("0") + code % 10)
    i = np.arange(custs)
    odd, short = i % 17 == 5, i % 31 == 7
    code[odd | short] = 255
    pair[odd] = (ord("A") << 8) | ord("B")
    pair[short] = -1
    r["c_phone_code"], r["c_phone_pair"] = code, pair.astype(np.int32)
    # supplier
    r["s_suppkey"] = np.arange(1, supps + 1, dtype=np.int32)
    r["s_nationkey"] = rng.integers(0, 25, supps).astype(np.int32)
    r["s_name"] = [f"Supplier#{k:09d}" for k in r["s_suppkey"]]
    # every seventh supplier has "Customer ... Complaints" in its comment,
    # every eleventh only "Complaints ... Customer" (not a ban)
    r["s_comment"] = ["blithely Customer slyly Complaints even" if k % 7 == 3 else
                      "Complaints about Customer" if k % 11 == 4 else "carefully final deposits"
                      for k in r["s_suppkey"].tolist()]
    # part
    r["p_partkey"] = np.arange(1, parts + 1, dtype=np.int32)
    r["p_brand"] = rng.integers(0, len(BRANDS), parts)
    r["p_container"] = rng.integers(0, len(CONTAINERS), parts)
    r["p_size"] = rng.integers(1, 51, parts).astype(np.int32)
    r["p_type"] = rng.integers(0, len(TYPES), parts)
    r["p_name_words"] = rng.integers(0, len(COLORS), (parts, 5))
    # partsupp: four distinct suppliers per part (spread over the suppliers)
    pk = np.repeat(r["p_partkey"].astype(np.int64), 4)
    j = np.tile(np.arange(4), parts)
    r["ps_partkey"] = pk.astype(np.int32)
    r["ps_suppkey"] = ((pk + j * max(1, supps // 4)) % supps + 1).astype(np.int32)
    r["ps_supplycost"] = rng.integers(100, 100_001, 4 * parts).astype(np.int32)
    # orders (created in key order)
    r["o_orderkey"] = okeys(orders)
    # as in TPC-H, no order has a customer key divisible by 3
    ck = rng.integers(1, custs + 1, orders)
    ck = np.where(ck % 3 == 0, np.maximum(ck - 1, 1), ck)
    r["o_custkey"] = ck.astype(np.int32)
    r["o_orderdate"] = rng.integers(0, 2406, orders).astype(np.int32)
    r["o_orderpriority"] = rng.integers(0, 5, orders)
    r["o_totalprice"] = rng.integers(100_000, 50_000_000, orders).astype(np.int64)
    r["o_shippriority"] = np.zeros(orders, np.int32)
    r["o_comment_chars"], r["o_comment_offsets"] = comments(rng, orders)
    # lineitem, in order-key order
    per = rng.integers(1, 8, orders)
    n = int(per.sum())
    oi = np.repeat(np.arange(orders), per)
    r["l_orderkey"] = r["o_orderkey"][oi]
    r["l_partkey"] = rng.integers(1, parts + 1, n).astype(np.int32)
    # each lineitem's supplier is one of its part's four (partsupp)
    r["l_suppkey"] = r["ps_suppkey"][(r["l_partkey"].astype(np.int64) - 1) * 4 +
                                     rng.integers(0, 4, n)]
    qty = rng.integers(1, 51, n)
    r["l_quantity"] = (qty * 100).astype(np.int64)                       # hundredths
    r["l_extendedprice"] = (qty * rng.integers(90_000, 200_001, n) // 100).astype(np.int64)
    r["l_discount"] = rng.integers(0, 11, n).astype(np.int64)
    r["l_tax"] = rng.integers(0, 9, n).astype(np.int64)
    od = r["o_orderdate"][oi]
    ship = od + rng.integers(1, 122, n)
    commit = od + rng.integers(30, 91, n)
    receipt = ship + rng.integers(1, 31, n)
    r["l_shipdate"], r["l_commitdate"], r["l_receiptdate"] = (
        ship.astype(np.int32), commit.astype(np.int32), receipt.astype(np.int32))
    rflag = np.where(receipt <= CURRENT, np.where(rng.random(n) < 0.5, 0, 1), 2)  # R, A, N
    r["l_returnflag"] = rflag                      # values index ["R", "A", "N"]
    r["l_linestatus"] = np.where(ship > CURRENT, 0, 1)   # ["O", "F"]
    # orderstatus from the lines' status: F all F, O all O, else P
    ls = r["l_linestatus"]
    n_f = np.bincount(oi, weights=(ls == 1), minlength=orders)
    r["o_orderstatus"] = np.where(n_f == per, 0, np.where(n_f == 0, 1, 2))  # F, O, P
    r["l_shipinstruct"] = rng.integers(0, 4, n)
    r["l_shipmode"] = rng.integers(0, 7, n)
    r["rf_values"], r["ls_values"] = ["R", "A", "N"], ["O", "F"]
    # columns of Q2, Q11, Q15 and Q20, drawn from their own stream so the
    # columns above do not move: available quantities (small ones too, so
    # Q20's availqty * 200 > sum sits on its boundary now and then), supplier
    # balances with repeats (Q2's sort falls through to nation and name),
    # addresses with a quote, manufacturers
    r2 = np.random.default_rng(seed + 7919)
    r["ps_availqty"] = np.where(r2.random(4 * parts) < 0.3, r2.integers(1, 40, 4 * parts),
                                r2.integers(1, 10_000, 4 * parts)).astype(np.int32)
    r["s_acctbal"] = (r2.integers(-999, 9_999, supps) * 100 +
                      np.where(r2.random(supps) < 0.5, 0, r2.integers(0, 100, supps))).astype(
        np.int64)
    r["p_mfgr"] = r2.integers(0, 5, parts)
    if strings:
        r["s_address"] = [f'{k * 37 % 1000} "Elm" St, {k}' if k % 5 == 0 else f"road {k * 911 % 997}"
                          for k in r["s_suppkey"].tolist()]
        r["s_phone"] = [f"{10 + nk}-{k % 900 + 100}-{k * 7 % 9000 + 1000}"
                        for k, nk in zip(r["s_suppkey"].tolist(), r["s_nationkey"].tolist())]
        ch, off = r["o_comment_chars"].tobytes(), r["o_comment_offsets"]
        r["o_comment"] = [ch[off[i]:off[i + 1]].decode() for i in range(orders)]
        r["p_name"] = [" ".join(COLORS[w] for w in row) for row in r["p_name_words"].tolist()]
        r["c_name"] = [f"Customer#{k:09d}" for k in r["c_custkey"]]
        if custs > 1:
            r["c_name"][0], r["c_name"][1] = 'He said "hi"', "plain,comma"
        r["c_address"] = [f"addr {k * 7919 % 100003}" for k in r["c_custkey"]]
        r["c_phone"] = [("7" if pr < 0 else chr(pr >> 8) + chr(pr & 255)) +
                        ("" if pr < 0 else f"-{k % 900 + 100}-{k % 9000 + 1000}")
                        for k, pr in zip(r["c_custkey"], r["c_phone_pair"])]
        r["c_comment"] = [f'note "{k}" about, things' for k in r["c_custkey"]]
    return r


def first_appearance(codes: np.ndarray, values: list) -> tuple[np.ndarray, list]:
    """Re-codes value indices in first-appearance order (the reference's
    dictionary builder, builder_impl.cpp:204-224)."""
    _, first = np.unique(codes, return_index=True)
    order = [int(c) for c in codes[np.sort(first)]]
    remap = np.full(max(len(values), 1), -1, np.int64)
    for new, old in enumerate(order):
        remap[old] = new
    return remap[codes].astype(np.uint16), [values[o] for o in order]


def build(r: dict) -> dict:
    """The reference build() layout of the columns the queries read."""
    db = {}
    li_n = len(r["l_orderkey"])
    # lineitem
    db["l_orderkey"] = r["l_orderkey"].astype(np.int32)
    db["l_partkey"] = r["l_partkey"].astype(np.int32)
    db["l_suppkey"] = r["l_suppkey"].astype(np.int32)
    db["l_quantity"] = r["l_quantity"].astype(np.int16)
    db["l_extendedprice"] = r["l_extendedprice"].astype(np.int32)
    db["l_discount"] = r["l_discount"].astype(np.uint8)
    db["l_shipdate"] = r["l_shipdate"].astype(np.int16)
    p = db["l_extendedprice"].astype(np.int64)
    db["l_disc_price"] = (p * (100 - db["l_discount"].astype(np.int64)) // 100).astype(np.int32)
    db["l_commit_receipt"] = ((r["l_receiptdate"].astype(np.uint32) & 0xffff) << 16) | \
        (r["l_commitdate"].astype(np.uint32) & 0xffff)
    rf, db["rf_dict"] = first_appearance(r["l_returnflag"], r["rf_values"])
    ls, db["ls_dict"] = first_appearance(r["l_linestatus"], r["ls_values"])
    db["l_code"] = (rf.astype(np.int64) * len(db["ls_dict"]) + ls).astype(np.uint8)
    db["l_shipmode"], db["shipmode_dict"] = first_appearance(r["l_shipmode"], SHIPMODES)
    db["l_shipinstruct"], db["instruct_dict"] = first_appearance(r["l_shipinstruct"], INSTRUCTS)
    ok = db["l_orderkey"]
    db["l_sorted"] = int(li_n < 2 or bool(np.all(ok[1:] >= ok[:-1])))
    # supplier nation per lineitem (builder_impl.cpp:838-853)
    smax = int(r["s_suppkey"].max()) if len(r["s_suppkey"]) else -1
    s_by_key = np.full(smax + 1, -1, np.int16)
    s_by_key[r["s_suppkey"]] = r["s_nationkey"]
    sk = db["l_suppkey"]
    inb = (sk >= 0) & (sk <= smax)
    db["l_supp_nation"] = np.where(inb, s_by_key[np.clip(sk, 0, max(smax, 0))], -1).astype(
        np.int16)
    # orders, sorted by date (stable)
    perm = np.argsort(r["o_orderdate"], kind="stable")
    db["o_orderkey"] = r["o_orderkey"][perm].astype(np.int32)
    db["o_custkey"] = r["o_custkey"][perm].astype(np.int32)
    db["o_orderdate"] = r["o_orderdate"][perm].astype(np.int16)
    pr, db["priority_dict"] = first_appearance(r["o_orderpriority"], PRIORITIES)
    db["o_orderpriority"] = pr[perm]
    st, db["status_dict"] = first_appearance(r["o_orderstatus"], ["F", "O", "P"])
    db["o_orderstatus"] = st[perm]
    cmax = int(r["c_custkey"].max()) if len(r["c_custkey"]) else -1
    c_by_key = np.full(cmax + 1, -1, np.int16)
    c_by_key[r["c_custkey"]] = r["c_nationkey"]
    ck = db["o_custkey"]
    inb = (ck >= 0) & (ck <= cmax)
    db["o_cust_nation"] = np.where(inb, c_by_key[np.clip(ck, 0, max(cmax, 0))], -1).astype(
        np.int16)
    # orderkey -> row (last row wins) and lineitem ranges
    kmax = int(db["o_orderkey"].max()) if len(db["o_orderkey"]) else -1
    m = np.full(kmax + 1, -1, np.int32)
    m[db["o_orderkey"]] = np.arange(len(db["o_orderkey"]), dtype=np.int32)
    db["order_map"] = m
    start = np.zeros(len(db["o_orderkey"]), np.uint32)
    end = np.zeros(len(db["o_orderkey"]), np.uint32)
    if li_n and db["l_sorted"]:
        # runs of equal keys (the reference's sorted branch, :934-961)
        keys, first, count = np.unique(ok, return_index=True, return_counts=True)
        inb = (keys >= 0) & (keys <= kmax)
        row = np.where(inb, m[np.clip(keys, 0, max(kmax, 0))], -1)
        got = row >= 0
        start[row[got]] = first[got]
        end[row[got]] = (first + count)[got]
    elif li_n:
        inb = (ok >= 0) & (ok <= kmax)
        row = np.where(inb, m[np.clip(ok, 0, max(kmax, 0))], -1)
        idx = np.arange(li_n)
        has = row >= 0
        if has.any():
            lo = np.full(len(start), np.iinfo(np.int64).max)
            hi = np.full(len(start), -1)
            np.minimum.at(lo, row[has], idx[has])
            np.maximum.at(hi, row[has], idx[has])
            got = hi >= 0
            start[got] = lo[got]
            end[got] = hi[got] + 1
    db["r_start"], db["r_end"] = start, end
    # customer, part, nation
    db["c_custkey"] = r["c_custkey"].astype(np.int32)
    db["c_nationkey"] = r["c_nationkey"].astype(np.int32)
    db["c_acctbal"] = r["c_acctbal"].astype(np.int32)
    db["c_phone_code"] = r["c_phone_code"].astype(np.uint8)
    db["c_phone_pair"] = r["c_phone_pair"].astype(np.int32)
    db["s_suppkey"] = r["s_suppkey"].astype(np.int32)
    db["s_nationkey"] = r["s_nationkey"].astype(np.int32)
    db["o_custkey_raw"] = r["o_custkey"].astype(np.int32)
    # order comments stay in the table's row order (Q13 reads them with the
    # raw order custkeys; the reference reorders both with the date sort)
    db["o_comment_chars"], db["o_comment_offsets"] = r["o_comment_chars"], r["o_comment_offsets"]
    db["s_banned"] = np.array(["Customer" in c and "Complaints" in c[c.index("Customer") + 8:]
                               for c in r["s_comment"]], np.uint8)
    db["p_partkey"] = r["p_partkey"].astype(np.int32)
    db["p_brand"], db["brand_dict"] = first_appearance(r["p_brand"], BRANDS)
    db["p_container"], db["container_dict"] = first_appearance(r["p_container"], CONTAINERS)
    db["p_size"] = r["p_size"].astype(np.int32)
    db["p_type"], db["type_dict"] = first_appearance(r["p_type"], TYPES)
    db["p_name_words"] = r["p_name_words"]
    db["ps_partkey"] = r["ps_partkey"]
    db["ps_suppkey"] = r["ps_suppkey"]
    db["ps_supplycost"] = r["ps_supplycost"]
    db["ps_availqty"] = r["ps_availqty"].astype(np.int32)
    db["ps_value"] = (r["ps_supplycost"].astype(np.int64) * r["ps_availqty"]).astype(np.int32)
    db["s_acctbal"] = r["s_acctbal"].astype(np.int32)
    db["p_mfgr"], db["mfgr_dict"] = first_appearance(r["p_mfgr"], MFGRS)
    db["n_nationkey"] = r["n_nationkey"]
    db["n_name"] = r["n_name"]
    db["n_regionkey"] = r["n_regionkey"]
    db["r_name"] = r["r_name"]
    return db


def parquet(r: dict) -> dict:
    """The eight tables as the reference's Parquet inputs."""
    import decimal

    import pyarrow as pa
    D = decimal.Decimal
    dec = lambda v: pa.array([D(int(c)).scaleb(-2) for c in v], pa.decimal128(15, 2))
    day = lambda v: pa.array([BASE + datetime.timedelta(days=int(s)) for s in v], pa.date32())
    strs = lambda codes, vals: pa.array([vals[int(c)] for c in codes])
    i32 = lambda v: pa.array(np.asarray(v, np.int32))
    li = pa.table({
        "l_orderkey": i32(r["l_orderkey"]), "l_partkey": i32(r["l_partkey"]),
        "l_suppkey": i32(r["l_suppkey"]),
        "l_quantity": dec(r["l_quantity"]), "l_extendedprice": dec(r["l_extendedprice"]),
        "l_discount": dec(r["l_discount"]), "l_tax": dec(r["l_tax"]),
        "l_returnflag": strs(r["l_returnflag"], r["rf_values"]),
        "l_linestatus": strs(r["l_linestatus"], r["ls_values"]),
        "l_shipdate": day(r["l_shipdate"]), "l_commitdate": day(r["l_commitdate"]),
        "l_receiptdate": day(r["l_receiptdate"]),
        "l_shipinstruct": strs(r["l_shipinstruct"], INSTRUCTS),
        "l_shipmode": strs(r["l_shipmode"], SHIPMODES)})
    orders = pa.table({
        "o_orderkey": i32(r["o_orderkey"]), "o_custkey": i32(r["o_custkey"]),
        "o_orderdate": day(r["o_orderdate"]), "o_totalprice": dec(r["o_totalprice"]),
        "o_shippriority": i32(r["o_shippriority"]),
        "o_orderpriority": strs(r["o_orderpriority"], PRIORITIES),
        "o_orderstatus": strs(r["o_orderstatus"], ["F", "O", "P"]),
        "o_comment": pa.array(r["o_comment"])})
    cust = pa.table({
        "c_custkey": i32(r["c_custkey"]), "c_name": pa.array(r["c_name"]),
        "c_nationkey": i32(r["c_nationkey"]), "c_acctbal": dec(r["c_acctbal"]),
        "c_mktsegment": strs(r["c_mktsegment"], SEGMENTS),
        "c_address": pa.array(r["c_address"]), "c_phone": pa.array(r["c_phone"]),
        "c_comment": pa.array(r["c_comment"])})
    part = pa.table({
        "p_partkey": i32(r["p_partkey"]), "p_brand": strs(r["p_brand"], BRANDS),
        "p_container": strs(r["p_container"], CONTAINERS), "p_size": i32(r["p_size"]),
        "p_type": strs(r["p_type"], TYPES), "p_name": pa.array(r["p_name"]),
        "p_mfgr": strs(r["p_mfgr"], MFGRS)})
    partsupp = pa.table({"ps_partkey": i32(r["ps_partkey"]), "ps_suppkey": i32(r["ps_suppkey"]),
                         "ps_supplycost": dec(r["ps_supplycost"]),
                         "ps_availqty": i32(r["ps_availqty"])})
    supp = pa.table({"s_suppkey": i32(r["s_suppkey"]), "s_nationkey": i32(r["s_nationkey"]),
                     "s_name": pa.array(r["s_name"]), "s_comment": pa.array(r["s_comment"]),
                     "s_acctbal": dec(r["s_acctbal"]), "s_address": pa.array(r["s_address"]),
                     "s_phone": pa.array(r["s_phone"])})
    nation = pa.table({"n_nationkey": i32(r["n_nationkey"]), "n_name": pa.array(r["n_name"]),
                       "n_regionkey": i32(r["n_regionkey"])})
    region = pa.table({"r_regionkey": i32(r["r_regionkey"]), "r_name": pa.array(r["r_name"])})
    return {"lineitem": li, "orders": orders, "customer": cust, "part": part, "partsupp": partsupp,
            "supplier": supp, "nation": nation, "region": region}


def date_offset(s: str) -> int:
    return (datetime.date.fromisoformat(s) - BASE).days


def add_months(s: str, months: int) -> int:
    """The reference's month add with the day clamped to the month's end
    (query_q4.cpp:40-60), as a day offset."""
    d = datetime.date.fromisoformat(s)
    mi = d.month - 1 + months
    y, mi = d.year + mi // 12, mi % 12
    import calendar
    day = min(d.day, calendar.monthrange(y, mi + 1)[1])
    return (datetime.date(y, mi + 1, day) - BASE).days

"""The reference's CEB pipelines (oracle/_ref/libref_ceb.so) against pandas.

Pins the shim that fills the reference's Database (tests/ceb_ref.py,
oracle/ref_ceb.cpp) and the reading of the reference that the GPU driver
restates: COUNT(*) of 1a and 2a is the SQL count over the seeded IMDB-shaped
tables with their nulls (a NULL never matches).  An IN list of 2a whose values
were split at a comma matches the column value joined back with ", "
(query_q2a.cpp:75-121); 1a looks each value up alone.
Skipped where the reference could not be built.
"""
from __future__ import annotations

import numpy as np
import pandas as pd
import pytest

from tests import ceb_data as CD
from tests import ceb_ref as CR

pytestmark = pytest.mark.skipif(not CR.available(), reason="oracle/_ref not built")


def frame(t, table):
    cols = {}
    for c, v in t[table].items():
        if isinstance(v, tuple):
            s = pd.Series(v[0], dtype="Int64")
            if v[1] is not None:
                s[v[1]] = pd.NA
            cols[c] = s
        else:
            cols[c] = pd.Series(v, dtype="object")
    return pd.DataFrame(cols)


def sql_count(t, q, with_keyword, info_keys=("INFO1v", "INFO2v")):
    """The SQL COUNT(*) of 1a / 2a: a NULL never equals anything.  info_keys
    name the IN-list values as the query resolves them."""
    kt, rt, nm = frame(t, "kind_type"), frame(t, "role_type"), frame(t, "name")
    ti, mi, ci = frame(t, "title"), frame(t, "movie_info"), frame(t, "cast_info")
    kinds = set(kt.loc[kt["kind"].isin(q["KIND"]), "id"].dropna())
    roles = set(rt.loc[rt["role"].isin(q["ROLE"]), "id"].dropna())
    people = set(nm.loc[nm["gender"].isin(q["GENDER"]), "id"].dropna())
    types = set(frame(t, "info_type")["id"].dropna())
    y = ti["production_year"]
    tt = ti[ti["id"].notna() & y.notna() & (y >= q["YEAR2"]) & (y <= q["YEAR1"]) &
            ti["kind_id"].isin(kinds)]
    movies = set(tt["id"])

    def per_movie(df, keep):
        d = df[df["movie_id"].isin(movies) & keep]
        return d.groupby("movie_id").size()

    def mi_count(ids, vals):
        keep = mi["info_type_id"].isin(set(ids) & types) & mi["info"].isin(vals)
        return per_movie(mi, keep)

    parts = [mi_count(q["ID1"], q[info_keys[0]]), mi_count(q["ID2"], q[info_keys[1]]),
             per_movie(ci, ci["role_id"].isin(roles) & ci["person_id"].isin(people))]
    if with_keyword:
        mk = frame(t, "movie_keyword")
        kws = set(frame(t, "keyword")["id"].dropna())
        parts.append(per_movie(mk, mk["keyword_id"].isin(kws)))
    total = 0
    for m in tt["id"]:
        p = 1
        for s in parts:
            p *= int(s.get(m, 0))
        total += p
    return total


def line(qid, q):
    lst = lambda v: "(" + ", ".join("'" + str(x).replace("'", "''") + "'" for x in v) + ")"
    return (f"{qid} {lst(q['ID1'])} {lst(q['ID2'])} {lst(q['INFO1'])} {lst(q['INFO2'])} "
            f"{lst(q['KIND'])} {lst(q['ROLE'])} {lst(q['GENDER'])} "
            f"\"{q['YEAR1']}\" \"{q['YEAR2']}\"")


QUERIES = [
    dict(ID1=[3], ID2=[8], INFO1=["Drama", "Comedy"], INFO1v=["Drama", "Comedy"],
         INFO2=["USA", "UK"], INFO2v=["USA", "UK"], KIND=["movie", "tv movie"],
         ROLE=["actor", "actress"], GENDER=["m"], YEAR1=2015, YEAR2=1980),
    dict(ID1=[3, 4], ID2=[8], INFO1=["Horror", "English"], INFO1v=["Horror", "English"],
         INFO2=["Germany", "Austria"], INFO2v=["Germany, Austria"],
         KIND=["movie", "episode", "nope"], ROLE=["director", "writer", "producer"],
         GENDER=["f", "m"], YEAR1=2019, YEAR2=1950),
]


@pytest.mark.parametrize("k", range(len(QUERIES)))
def test_reference_counts_equal_pandas(k):
    t = CD.tables(7, 3000)
    q = QUERIES[k]
    ref = CR.RefCeb(t)
    got1a, got2a = ref.run([line("1a", q), line("2a", q)])
    want2a = sql_count(t, q, with_keyword=True)
    want_sql_1a = sql_count(t, q, with_keyword=False, info_keys=("INFO1", "INFO2"))
    assert want2a > 0 and want_sql_1a > 0
    assert got2a == f"count_star()\n{want2a}\n"
    assert got1a == f"count_star()\n{want_sql_1a}\n"


def test_reference_stub_and_errors():
    t = CD.tables(8, 200)
    ref = CR.RefCeb(t)
    assert ref.run(["99x 1"]) == [None]
    with pytest.raises(RuntimeError, match="IN list"):
        ref.run(["1a 3"])

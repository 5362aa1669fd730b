"""A seeded IMDB-shaped database (the CEB schema) with nulls.

`tables(seed, movies, ...)` draws title, kind_type, role_type, info_type,
name, keyword, movie_info, cast_info and movie_keyword in the reference's
column names and types (int32 ids, strings), as Python/numpy columns:
`{table: {column: (values, null_mask or None)}}` with string columns as lists
of str.  Nulls are drawn in the columns the reference's scans check
(title.kind_id / production_year, name.gender, movie_info.movie_id /
info_type_id / info, cast_info.person_id / role_id, movie_keyword ids), and
some fact rows point at dimension ids that do not exist.  movie_info holds
values with a comma ("Germany, Austria") so that an IN list split at the
comma exercises the reference's re-join of such values.

`parquet(t)` gives the tables as pyarrow Tables for the driver.
"""
from __future__ import annotations

import numpy as np

KINDS = ["movie", "tv series", "tv movie", "video movie", "tv mini series", "video game",
         "episode"]
ROLES = ["actor", "actress", "producer", "writer", "cinematographer", "composer",
         "costume designer", "director", "editor", "miscellaneous crew",
         "production designer", "guest"]
GENRES = ["Drama", "Comedy", "Horror", "Action", "Sci-Fi", "Romance", "Thriller"]
COUNTRIES = ["USA", "UK", "France", "Germany", "Austria", "Germany, Austria", "Italy,Spain",
             "Italy", "Spain", "Japan"]
LANGS = ["English", "French", "German", "Japanese", "Italian"]
INFO_TYPES = {3: GENRES, 8: COUNTRIES, 4: LANGS, 16: [f"{y}" for y in range(1990, 2000)]}


def _nulls(rng, n, p):
    return rng.random(n) < p if p > 0 else None


def tables(seed: int, movies: int, people: int = 3000, mi_per: float = 4.0,
           ci_per: float = 6.0, mk_per: float = 3.0, keywords: int = 400) -> dict:
    rng = np.random.default_rng(seed)
    t = {}
    i32 = lambda a: np.asarray(a, np.int32)
    # dimensions
    t["kind_type"] = {"id": (i32(np.arange(1, len(KINDS) + 2)), None),
                      "kind": KINDS + [None]}
    t["role_type"] = {"id": (i32(np.arange(1, len(ROLES) + 1)), None), "role": list(ROLES)}
    it_ids = np.arange(1, 114)
    t["info_type"] = {"id": (i32(it_ids), None), "info": [f"info type {i}" for i in it_ids]}
    t["keyword"] = {"id": (i32(np.arange(1, keywords + 1)), None),
                    "keyword": [f"kw-{i % (keywords // 2)}" for i in range(1, keywords + 1)]}
    # title: ids 1..movies in shuffled row order
    ids = rng.permutation(np.arange(1, movies + 1))
    t["title"] = {"id": (i32(ids), None),
                  "kind_id": (i32(rng.integers(1, len(KINDS) + 2, movies)),
                              _nulls(rng, movies, 0.03)),
                  "production_year": (i32(rng.integers(1950, 2020, movies)),
                                      _nulls(rng, movies, 0.08)),
                  "title": [f"Title {i % (movies // 3 + 1)}" for i in ids.tolist()]}
    # name: ids 1..people shuffled
    pid = rng.permutation(np.arange(1, people + 1))
    g = rng.integers(0, 3, people)
    t["name"] = {"id": (i32(pid), None),
                 "name": [f"Person {i}" for i in pid.tolist()],
                 "gender": [None if x == 2 else "mf"[x] for x in g.tolist()],
                 "name_pcode_cf": [f"P{i % 50}" for i in pid.tolist()]}

    def movie_ids(n):   # skewed towards low ids; a few past the last title
        m = 1 + np.floor(movies * rng.random(n) ** 2).astype(np.int64)
        m[rng.random(n) < 0.005] = movies + 5
        return i32(m)

    # movie_info
    n_mi = int(movies * mi_per)
    types = np.array(list(INFO_TYPES) + [1, 2, 7], np.int32)
    ty = types[rng.integers(0, len(types), n_mi)]
    info = []
    for k in ty.tolist():
        vocab = INFO_TYPES.get(k, ["x", "y", "z"])
        info.append(vocab[int(rng.integers(0, len(vocab)))])
    for i in np.nonzero(rng.random(n_mi) < 0.02)[0].tolist():
        info[i] = None
    t["movie_info"] = {"movie_id": (movie_ids(n_mi), _nulls(rng, n_mi, 0.02)),
                       "info_type_id": (ty, _nulls(rng, n_mi, 0.02)),
                       "info": info}
    # cast_info
    n_ci = int(movies * ci_per)
    person = rng.integers(1, people + 20, n_ci)          # some past the last person
    t["cast_info"] = {"movie_id": (movie_ids(n_ci), _nulls(rng, n_ci, 0.01)),
                      "person_id": (i32(person), _nulls(rng, n_ci, 0.01)),
                      "role_id": (i32(rng.integers(1, len(ROLES) + 1, n_ci)),
                                  _nulls(rng, n_ci, 0.03))}
    # movie_keyword
    n_mk = int(movies * mk_per)
    t["movie_keyword"] = {"movie_id": (movie_ids(n_mk), _nulls(rng, n_mk, 0.01)),
                          "keyword_id": (i32(rng.integers(1, keywords + 30, n_mk)),
                                         _nulls(rng, n_mk, 0.01))}
    return t


def column(t: dict, table: str, col: str):
    """(values, null mask) of an int column, or the list of a string column."""
    return t[table][col]


def parquet(t: dict) -> dict:
    import pyarrow as pa
    out = {}
    for name, cols in t.items():
        arrays = {}
        for c, v in cols.items():
            if isinstance(v, tuple):
                vals, nul = v
                arrays[c] = pa.array(vals, pa.int32(), mask=nul)
            else:
                arrays[c] = pa.array(v, pa.string())
        out[name] = pa.table(arrays)
    return out

"""CEB star queries on the GPU: one at a time and batched (hipGraph replay,
autotuned), bit-exact against the oracle."""
import re

import numpy as np
import pytest

import oracle as O
from bespokeolap_artifacts_amd import ceb as CB
from bespokeolap_artifacts_amd import native as N

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(engine):
    return CB.CebSchema(engine, seed=3, scale=0.01), O.gen_ceb(3, 0.01)


def host(t):
    return t.cpu().numpy()


def like_regex_mask(chars: bytes, offsets: np.ndarray, pattern: str) -> np.ndarray:
    """ILIKE as a regular expression, independent of the library: '%' any
    run, '_' any one character, everything else literal, ASCII case folded,
    whole string.  It agrees with the reference's matcher (query_q10a.cpp:
    43-209) on text without '%' or '_', which the dictionaries here hold."""
    key = (id(chars), pattern)
    if key not in _LIKE_CACHE:
        texts = _TEXTS.get(id(chars))
        if texts is None:
            texts = [chars[offsets[i]:offsets[i + 1]].decode() for i in range(len(offsets) - 1)]
            assert not any("%" in x or "_" in x for x in texts)
            _TEXTS[id(chars)] = texts
        rx = re.compile("".join(".*" if c == "%" else "." if c == "_" else re.escape(c)
                                for c in pattern), re.DOTALL | re.IGNORECASE | re.ASCII)
        _LIKE_CACHE[key] = np.fromiter((rx.fullmatch(x) is not None for x in texts), np.uint8,
                                       len(texts))
    return _LIKE_CACHE[key]


_TEXTS: dict = {}
_LIKE_CACHE: dict = {}


def oracle_masks(schema, q):
    """The oracle's masks from like_regex_mask; the library's own (the
    device bitmaps' source) must agree with them."""
    m = {}
    for key in ("title", "info", "info_type", "keyword"):
        if key in q:
            chars, offs = schema.dicts[key]
            mask = like_regex_mask(chars, offs, q[key])
            np.testing.assert_array_equal(schema.bits(key, q[key])[2], mask)
            if key in ("info_type", "keyword"):  # over ids: id = code + 1
                mask = np.concatenate([[0], mask]).astype(np.uint8)
            m[key] = mask
    return m


def test_tables_match_oracle(small):
    s, t = small
    for c in ("year", "kind", "code"):
        np.testing.assert_array_equal(host(s.title[c]), t["title"][c])
    for name, cols in (("ci", ("movie", "role")), ("mi", ("movie", "type", "info")),
                       ("mk", ("movie", "keyword"))):
        for c in cols:
            np.testing.assert_array_equal(host(getattr(s, name)[c]), t[name][c])


def test_single_queries_match_oracle(engine, small):
    s, t = small
    qs = CB.random_queries(s, 48, seed=5)
    for q in qs:
        got = CB.run_one(engine, s, CB.bind(s, q))
        assert got == O.ceb_run(t, q, oracle_masks(s, q)), q


def test_batch_equals_single_and_replays(engine, small):
    s, t = small
    qs = CB.random_queries(s, 64, seed=6)
    bound = [CB.bind(s, q) for q in qs]
    single = [CB.run_one(engine, s, b) for b in bound]
    for rep in range(3):  # graph replays, with eager work in between
        assert CB.run_batch(engine, s, bound) == single
        CB.run_one(engine, s, bound[rep])
    engine.tune_set("ceb.graph", 2)
    try:
        assert CB.run_batch(engine, s, bound) == single
    finally:
        engine.tune_set("ceb.graph", 0)


def test_autotune_keeps_results(engine, small):
    s, t = small
    qs = [CB.bind(s, q) for q in CB.random_queries(s, 40, seed=7)]
    before = CB.run_batch(engine, s, qs)
    ms = CB.autotune(engine, s, qs)
    assert ms > 0
    assert engine.tune_get("ceb.blocks_per_cu") in (2, 4, 8, 16)
    assert CB.run_batch(engine, s, qs) == before


def test_empty_and_min_edge_cases(engine, small):
    s, t = small
    q = {"tables": N.CEB_CI, "agg": N.CEB_MIN_YEAR, "year_lo": 2030, "year_hi": 2040,
         "kind_mask": 0xFE, "role_mask": 0x1FFE}
    r = CB.run_one(engine, s, CB.bind(s, q))
    assert r == {"count": 0, "min_year": 2**63 - 1, "title_pass": 0, "joined_movies": 0}
    # a role mask that admits nothing: movies pass the title but join nothing
    q2 = dict(q, year_lo=1880, year_hi=2025, role_mask=1 << 20)
    r2 = CB.run_one(engine, s, CB.bind(s, q2))
    assert r2["title_pass"] > 0 and r2["count"] == 0 and r2["joined_movies"] == 0
    assert r2 == O.ceb_run(t, q2, {})
    with pytest.raises(N.BolapError):
        CB.run_one(engine, s, CB.bind(s, dict(q, tables=0)))


def test_full_size_batch_matches_single_and_oracle(engine):
    s = CB.CebSchema(engine, seed=42, scale=1.0)
    qs = CB.random_queries(s, 512, seed=42)
    bound = [CB.bind(s, q) for q in qs]
    batch = CB.run_batch(engine, s, bound)
    for i in range(0, 512, 8):
        assert CB.run_one(engine, s, bound[i]) == batch[i]
    t = O.gen_ceb(42, 1.0)
    for i in (0, 1, 2, 3, 5, 8, 13, 21):
        assert batch[i] == O.ceb_run(t, qs[i], oracle_masks(s, qs[i])), qs[i]


@pytest.mark.parametrize("qchunk", [32, 64, 0])
def test_batch_across_query_words_and_chunks(engine, small, qchunk):
    """A batch whose 32-query words use disjoint tables (a whole word skips a
    table), a ragged last word, and fact passes over 1, 2 or all words at a
    time: every result equals the one-query run."""
    s, t = small
    qs = CB.random_queries(s, 75, seed=11)
    for i, q in enumerate(qs):  # word 0: movie_keyword only; word 1: no movie_keyword
        if i < 32:
            q.update(tables=N.CEB_MK, keyword=q.get("keyword") or "%a%")
        elif i < 64:
            q["tables"] = q["tables"] & ~N.CEB_MK or N.CEB_CI
            q.pop("keyword", None)
    bound = [CB.bind(s, q) for q in qs]
    single = [CB.run_one(engine, s, b) for b in bound]
    old = engine.tune_get("ceb.qchunk")
    engine.tune_set("ceb.qchunk", qchunk or len(qs))
    try:
        assert CB.run_batch(engine, s, bound) == single
    finally:
        engine.tune_set("ceb.qchunk", old)
    for i in (0, 31, 32, 63, 64, 74):
        assert single[i] == O.ceb_run(t, qs[i], oracle_masks(s, qs[i])), qs[i]


def test_batch_beyond_one_staged_chunk(engine, small):
    """600 queries: more than the 512 whose predicates one workgroup stages in
    LDS, so the title and fact stages walk two query chunks (and a ragged last
    32-query word)."""
    s, t = small
    qs = CB.random_queries(s, 600, seed=13)
    bound = [CB.bind(s, q) for q in qs]
    batch = CB.run_batch(engine, s, bound)
    for i in list(range(0, 600, 37)) + [511, 512, 513, 575, 576, 599]:
        assert CB.run_one(engine, s, bound[i]) == batch[i], i
    for i in (0, 512, 599):
        assert batch[i] == O.ceb_run(t, qs[i], oracle_masks(s, qs[i])), qs[i]

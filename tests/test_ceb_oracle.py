"""CEB star queries: the oracle against materialised pandas joins, and the
library's host ILIKE against a regular-expression restatement (CPU only)."""
import re

import numpy as np
import pandas as pd
import pytest

import oracle as O
from bespokeolap_artifacts_amd import ceb as CB
from bespokeolap_artifacts_amd import native as N


def regex_ilike(pattern: str, s: str) -> bool:
    rx = "".join(".*" if c == "%" else "." if c == "_" else re.escape(c) for c in pattern)
    return re.fullmatch(rx, s, flags=re.IGNORECASE | re.DOTALL | re.ASCII) is not None


def lib_ilike(strings: list[str], pattern: str) -> np.ndarray:
    enc = [s.encode() for s in strings]
    offs = np.zeros(len(enc) + 1, np.int64)
    offs[1:] = np.cumsum([len(e) for e in enc])
    return CB.ilike_mask(N.load(), b"".join(enc), offs, pattern)


# Texts without a literal '%': there the reference's matcher is plain SQL ILIKE.
EDGE = ["", "a", "A", "abc", "ABC", "aXc", "a_c", "xxabcxx", "abcabc", "ab", "bc",
        "Love Story 12", "love", "a.c", "a(b)c", "_", "__", "ab_cd"]
PATTERNS = ["%", "", "_", "__", "a", "abc", "ABC", "a%", "%c", "%b%", "a_c", "a%c", "%a%c%",
            "_b_", "%%", "a%%c", "%abc", "abc%", "%love%", "LOVE%", "%story _2", "a.c",
            "a(b)c", "%\\%", "ab%cd", "%_%", "_%_"]


@pytest.mark.parametrize("pattern", PATTERNS)
def test_ilike_matches_regex_restatement(pattern):
    got = lib_ilike(EDGE, pattern)
    want = np.array([regex_ilike(pattern, s) for s in EDGE], np.uint8)
    np.testing.assert_array_equal(got, want)


def test_ilike_literal_percent_follows_the_reference():
    # Read off bespoke_ceb/query_q10a.cpp: 'a%' is a prefix test, so 'a%c'
    # matches; the general scan ('_' present) consumes the pattern '%' as a
    # literal against a '%' in the text first, so 'a%_d' does not match
    # 'a%xyd' although SQL ILIKE would.
    assert lib_ilike(["a%c"], "a%").tolist() == [1]
    assert lib_ilike(["a%xyd"], "a%_d").tolist() == [0]
    assert lib_ilike(["abxyd"], "a%_d").tolist() == [1]
    assert lib_ilike(["a%c"], "%\\%").tolist() == [0]


def test_ilike_on_a_generated_dictionary():
    chars, offs = CB.make_dictionary(3000, 5)
    strings = [chars[offs[i]:offs[i + 1]].decode() for i in range(3000)]
    for p in CB.patterns(np.random.default_rng(2), 30):
        got = CB.ilike_mask(N.load(), chars, offs, p)
        want = np.array([regex_ilike(p, s) for s in strings], np.uint8)
        np.testing.assert_array_equal(got, want, err_msg=p)


def test_mask_to_words_bit_layout():
    m = np.zeros(100, np.uint8)
    m[[0, 31, 32, 99]] = 1
    w = CB.mask_to_words(m, 1)
    bits = [i for i in range(len(w) * 32) if (int(w[i // 32]) >> (i % 32)) & 1]
    assert bits == [1, 32, 33, 100]


@pytest.fixture(scope="module")
def tiny():
    return O.gen_ceb(7, 0.0002)


def host_masks(t, q, rng):
    masks = {}
    if rng.random() < 0.5:
        masks["title"] = (rng.random(CB.DICT_SIZES["title"]) < 0.3).astype(np.uint8)
    if q["tables"] & N.CEB_MI:
        masks["info"] = (rng.random(CB.DICT_SIZES["info"]) < 0.5).astype(np.uint8)
        masks["info_type"] = (rng.random(101) < 0.5).astype(np.uint8)
    if q["tables"] & N.CEB_MK:
        masks["keyword"] = (rng.random(t["n"]["keyword"] + 1) < 0.4).astype(np.uint8)
    return masks


def pandas_ceb(t, q, masks):
    """The query as a materialised pandas join, counted and minimised."""
    tt = pd.DataFrame({"movie": np.arange(1, len(t["title"]["year"]) + 1),
                       "year": t["title"]["year"], "kind": t["title"]["kind"],
                       "code": t["title"]["code"]})
    ok = (tt.year >= q["year_lo"]) & (tt.year <= q["year_hi"]) & \
        tt.kind.map(lambda k: bool((q["kind_mask"] >> k) & 1))
    if "title" in masks:
        ok &= masks["title"][tt.code.to_numpy()].astype(bool)
    tt = tt[ok]
    title_pass = len(tt)
    j = tt[["movie", "year"]]
    if q["tables"] & N.CEB_CI:
        ci = pd.DataFrame(t["ci"])
        ci = ci[ci.role.map(lambda r: bool((q["role_mask"] >> r) & 1))]
        j = j.merge(ci[["movie"]], on="movie")
    if q["tables"] & N.CEB_MI:
        mi = pd.DataFrame(t["mi"])
        if "info_type" in masks:
            mi = mi[masks["info_type"][mi["type"].to_numpy()].astype(bool)]
        if "info" in masks:
            mi = mi[masks["info"][mi["info"].to_numpy()].astype(bool)]
        j = j.merge(mi[["movie"]], on="movie")
    if q["tables"] & N.CEB_MK:
        mk = pd.DataFrame(t["mk"])
        if "keyword" in masks:
            mk = mk[masks["keyword"][mk.keyword.to_numpy()].astype(bool)]
        j = j.merge(mk[["movie"]], on="movie")
    return {"count": len(j), "min_year": int(j.year.min()) if len(j) else 2**63 - 1,
            "title_pass": title_pass, "joined_movies": int(j.movie.nunique())}


def test_oracle_matches_materialised_joins(tiny):
    rng = np.random.default_rng(0)
    for i in range(24):
        tables, _ = CB.TEMPLATES[i % 8]
        lo = int(rng.integers(1880, 2000))
        q = {"tables": tables, "year_lo": lo, "year_hi": lo + int(rng.integers(5, 100)),
             "kind_mask": CB._rand_mask(rng, 1, 7), "role_mask": CB._rand_mask(rng, 1, 12)}
        masks = host_masks(tiny, q, rng)
        assert O.ceb_run(tiny, q, masks) == pandas_ceb(tiny, q, masks)


def test_planted_mistake_is_caught(tiny):
    rng = np.random.default_rng(1)
    q = {"tables": N.CEB_CI | N.CEB_MK, "year_lo": 1880, "year_hi": 2025,
         "kind_mask": 0xFE, "role_mask": 0x1FFE}
    O.set_mutation(8)
    try:
        got = O.ceb_run(tiny, q, {})
    finally:
        O.set_mutation(0)
    assert got["count"] != pandas_ceb(tiny, q, {})["count"]

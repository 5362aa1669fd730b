"""CEB-style batched workload: six seeded tables in HBM and parameterised
star queries over movie ids.

Tables (sizes of the benchmark; every column int32, fact tables sorted by
movie_id as the reference builder orders them,
bespoke_ceb/builder_impl.cpp:432-452):

    cast_info      36,000,000  movie_id (Zipf 1.2), role_id U{1..12}
    movie_info     15,000,000  movie_id (Zipf 1.2), info_type_id U{1..100},
                               info (code into a 200,000-string dictionary)
    movie_keyword   4,500,000  movie_id (Zipf 1.2), keyword_id (Zipf 1.2)
    title           2,500,000  id 1.., production_year U{1880..2025},
                               kind_id U{1..7}, title (code into 200,000 strings)
    keyword           130,000  id 1.., keyword (code = id - 1)
    info_type             100  id 1.., info (code = id - 1)

String predicates are ILIKE patterns evaluated on the host over the string
dictionaries (bolap_ilike_mask) and handed to the device as bitmaps, once
per distinct pattern.  Eight templates join title with 1..5 of the other
tables and aggregate COUNT(*) or MIN(production_year).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native as N
from .workload import coprime_mul, zipf_quantiles

SIZES = {"cast_info": 36_000_000, "movie_info": 15_000_000, "movie_keyword": 4_500_000,
         "title": 2_500_000, "keyword": 130_000, "info_type": 100}
DICT_SIZES = {"title": 200_000, "info": 200_000, "keyword": 130_000, "info_type": 100}
ZIPF_S = 1.2
YEAR_LO, YEAR_HI = 1880, 2025

# Column streams of the seeded generators.
S_CI_MOVIE, S_CI_ROLE = 0x400, 0x401
S_MI_MOVIE, S_MI_TYPE, S_MI_INFO = 0x410, 0x411, 0x412
S_MK_MOVIE, S_MK_KW = 0x420, 0x421
S_T_YEAR, S_T_KIND, S_T_CODE = 0x430, 0x431, 0x432

WORDS = ("love murder war family night city dark blood king queen dream ghost river "
         "storm fire ice star moon sun road house game secret story money power girl "
         "boy man woman friend death life time world heart soul song dance music school "
         "police hero zombie alien space robot island ocean mountain desert forest "
         "winter summer spring autumn gold silver iron stone glass shadow light").split()

# Templates: (tables, aggregate).  Joins: title-mk-k 2, title-ci 1,
# title-mi-it 2, title-ci-mk-k 3, title-mi-it-mk-k 4, all 5, title-ci-mi-it 3,
# title-mk-k 2.
TEMPLATES = [
    (N.CEB_MK, N.CEB_COUNT),
    (N.CEB_CI, N.CEB_COUNT),
    (N.CEB_MI, N.CEB_MIN_YEAR),
    (N.CEB_CI | N.CEB_MK, N.CEB_COUNT),
    (N.CEB_MI | N.CEB_MK, N.CEB_COUNT),
    (N.CEB_CI | N.CEB_MI | N.CEB_MK, N.CEB_COUNT),
    (N.CEB_CI | N.CEB_MI, N.CEB_MIN_YEAR),
    (N.CEB_MK, N.CEB_MIN_YEAR),
]


def make_dictionary(n: int, seed: int) -> tuple[bytes, np.ndarray]:
    """n strings of 2-3 words and a number, mixed case, as (chars, offsets)."""
    rng = np.random.default_rng(seed)
    k = rng.integers(2, 4, n)
    w = rng.integers(0, len(WORDS), (n, 3))
    caps = rng.random((n, 3)) < 0.3
    num = rng.integers(0, 1000, n)
    out = []
    for i in range(n):
        parts = []
        for j in range(int(k[i])):
            t = WORDS[int(w[i, j])]
            parts.append(t.capitalize() if caps[i, j] else t)
        out.append(" ".join(parts) + f" {int(num[i])}")
    enc = [s.encode() for s in out]
    offsets = np.zeros(n + 1, np.int64)
    offsets[1:] = np.cumsum([len(e) for e in enc])
    return b"".join(enc), offsets


def ilike_mask(lib, chars: bytes, offsets: np.ndarray, pattern: str) -> np.ndarray:
    """uint8 mask over the dictionary (bolap_ilike_mask, host code)."""
    n = len(offsets) - 1
    mask = np.zeros(max(1, n), np.uint8)
    N.check(lib.bolap_ilike_mask(chars, offsets.ctypes.data_as(C.POINTER(C.c_int64)), n,
                                 pattern.encode(), mask.ctypes.data_as(C.POINTER(C.c_uint8))))
    return mask[:n]


def mask_to_words(mask: np.ndarray, offset: int = 0) -> np.ndarray:
    """Bitmap words (uint32, bit v of word v // 32) of a 0/1 mask, shifted so
    that mask index i becomes bit i + offset."""
    n = len(mask) + offset
    bits = np.zeros((n + 31) // 32 * 32, np.uint8)
    bits[offset:offset + len(mask)] = mask
    return np.packbits(bits.reshape(-1, 8), axis=1, bitorder="little").reshape(-1).view(np.uint32)


def patterns(rng: np.random.Generator, k: int) -> list[str]:
    """ILIKE patterns: contains, prefix, suffix, and single-character wildcards."""
    out = []
    for _ in range(k):
        w = WORDS[int(rng.integers(0, len(WORDS)))]
        kind = int(rng.integers(0, 5))
        if kind == 0:
            out.append(f"%{w}%")
        elif kind == 1:
            out.append(f"{w}%")
        elif kind == 2:
            out.append(f"%{w} {int(rng.integers(0, 10))}%")
        elif kind == 3:
            i = int(rng.integers(0, len(w)))
            out.append(f"%{w[:i]}_{w[i + 1:]}%")
        else:
            w2 = WORDS[int(rng.integers(0, len(WORDS)))]
            out.append(f"%{w}%{w2}%")
    return out


class CebSchema:
    """The six tables in HBM plus host dictionaries and cached pattern bitmaps."""

    def __init__(self, engine, seed: int = 42, scale: float = 1.0):
        import torch
        self.engine = engine
        self.seed = int(seed)
        lib, ctx = engine.lib, engine.ctx
        dev = f"cuda:{engine.device}"
        n = {k: max(1, int(v * scale)) if k != "info_type" else v for k, v in SIZES.items()}
        self.n = n
        nt = n["title"]
        self.dicts = {k: make_dictionary(v if k != "keyword" else n["keyword"],
                                         self.seed + 17 * (i + 1))
                      for i, (k, v) in enumerate(DICT_SIZES.items())}

        def col(rows):
            return torch.empty(rows, dtype=torch.int32, device=dev)

        def uni(rows, stream, lo, hi):
            t = col(rows)
            N.check(lib.bolap_gen_uniform_i32(ctx, self.seed, stream, rows, lo, hi, _p(t)))
            return t

        def sorted_zipf(rows, stream, keys):
            q = torch.from_numpy(zipf_quantiles(keys, ZIPF_S)).to(dev)
            t = col(rows)
            N.check(lib.bolap_gen_zipf_i32(ctx, self.seed, stream, rows, _p(q), keys,
                                           coprime_mul(keys), _p(t)))
            k64 = t.to(torch.int64)
            v = torch.zeros(rows, dtype=torch.int32, device=dev)
            engine.radix_sort(k64, v, key_bits=max(1, int(keys).bit_length()))
            return k64.to(torch.int32)

        self.title = {"year": uni(nt, S_T_YEAR, YEAR_LO, YEAR_HI),
                      "kind": uni(nt, S_T_KIND, 1, 7),
                      "code": uni(nt, S_T_CODE, 0, DICT_SIZES["title"] - 1)}
        self.ci = {"movie": sorted_zipf(n["cast_info"], S_CI_MOVIE, nt),
                   "role": uni(n["cast_info"], S_CI_ROLE, 1, 12)}
        self.mi = {"movie": sorted_zipf(n["movie_info"], S_MI_MOVIE, nt),
                   "type": uni(n["movie_info"], S_MI_TYPE, 1, n["info_type"]),
                   "info": uni(n["movie_info"], S_MI_INFO, 0, DICT_SIZES["info"] - 1)}
        q = torch.from_numpy(zipf_quantiles(n["keyword"], ZIPF_S)).to(dev)
        kw = col(n["movie_keyword"])
        N.check(lib.bolap_gen_zipf_i32(ctx, self.seed, S_MK_KW, n["movie_keyword"], _p(q),
                                       n["keyword"], coprime_mul(n["keyword"]), _p(kw)))
        self.mk = {"movie": sorted_zipf(n["movie_keyword"], S_MK_MOVIE, nt), "keyword": kw}
        engine.sync()
        self._bits = {}
        self.db = N.CebDb(nt, _p(self.title["year"]), _p(self.title["kind"]),
                          _p(self.title["code"]),
                          n["cast_info"], _p(self.ci["movie"]), _p(self.ci["role"]),
                          n["movie_info"], _p(self.mi["movie"]), _p(self.mi["type"]),
                          _p(self.mi["info"]),
                          n["movie_keyword"], _p(self.mk["movie"]), _p(self.mk["keyword"]))

    def bits(self, dict_name: str, pattern: str):
        """Device bitmap of a pattern over a dictionary, cached per pattern.
        keyword and info_type bitmaps are over ids (code + 1)."""
        import torch
        key = (dict_name, pattern)
        if key not in self._bits:
            chars, offs = self.dicts[dict_name]
            mask = ilike_mask(self.engine.lib, chars, offs, pattern)
            shift = 1 if dict_name in ("keyword", "info_type") else 0
            words = mask_to_words(mask, shift)
            t = torch.from_numpy(words.view(np.int32).copy()).to(f"cuda:{self.engine.device}")
            self._bits[key] = (t, len(mask) + shift, mask)
        return self._bits[key]


def _p(t):
    return t.data_ptr()


def random_queries(schema: CebSchema, n: int, seed: int) -> list[dict]:
    """n queries drawn from the eight templates with random constants."""
    rng = np.random.default_rng(seed)
    pools = {d: patterns(np.random.default_rng(seed + i), 48)
             for i, d in enumerate(("title", "info", "keyword", "info_type"))}
    out = []
    for _ in range(n):
        t = int(rng.integers(0, len(TEMPLATES)))
        tables, agg = TEMPLATES[t]
        lo = int(rng.integers(YEAR_LO, YEAR_HI + 1))
        width = int(rng.integers(1, YEAR_HI - YEAR_LO + 2))
        q = {"template": t, "tables": tables, "agg": agg, "year_lo": lo,
             "year_hi": min(YEAR_HI, lo + width - 1),
             "kind_mask": _rand_mask(rng, 1, 7), "role_mask": _rand_mask(rng, 1, 12)}
        if rng.random() < 0.3:
            q["title"] = str(rng.choice(pools["title"]))
        if tables & N.CEB_MI:
            if rng.random() < 0.5:
                q["info"] = str(rng.choice(pools["info"]))
            if rng.random() < 0.7:
                q["info_type"] = f"%{WORDS[int(rng.integers(0, len(WORDS)))]}%"
        if tables & N.CEB_MK:
            q["keyword"] = str(rng.choice(pools["keyword"]))
        out.append(q)
    return out


def _rand_mask(rng, lo: int, hi: int) -> int:
    k = int(rng.integers(1, hi - lo + 2))
    vals = rng.choice(np.arange(lo, hi + 1), k, replace=False)
    m = 0
    for v in vals:
        m |= 1 << int(v)
    return m


def bind(schema: CebSchema, q: dict) -> N.CebQuery:
    s = N.CebQuery()
    s.tables, s.agg = q["tables"], q["agg"]
    s.year_lo, s.year_hi = q["year_lo"], q["year_hi"]
    s.kind_mask, s.role_mask = q["kind_mask"], q["role_mask"]
    for field, dname in (("title", "title"), ("info", "info"), ("type", "info_type"),
                         ("kw", "keyword")):
        pat = q.get(dname)
        if pat is None:
            continue
        t, nbits, _ = schema.bits(dname, pat)
        setattr(s, f"{field}_bits", t.data_ptr())
        setattr(s, f"{field}_bits_n", nbits)
    return s


def result_dict(r: N.CebResult) -> dict:
    return {"count": (int(r.count_hi) << 64) | int(r.count_lo), "min_year": int(r.min_year),
            "title_pass": int(r.title_pass), "joined_movies": int(r.joined_movies)}


def run_one(engine, schema: CebSchema, q: N.CebQuery) -> dict:
    r = N.CebResult()
    N.check(engine.lib.bolap_ceb_run(engine.ctx, C.byref(schema.db), C.byref(q), C.byref(r)))
    return result_dict(r)


def run_batch(engine, schema: CebSchema, qs: list) -> list[dict]:
    arr = (N.CebQuery * len(qs))(*qs)
    out = (N.CebResult * len(qs))()
    N.check(engine.lib.bolap_ceb_run_batch(engine.ctx, C.byref(schema.db), arr, len(qs), out))
    return [result_dict(r) for r in out]


def autotune(engine, schema: CebSchema, qs: list) -> float:
    arr = (N.CebQuery * len(qs))(*qs)
    best = C.c_double(0)
    N.check(engine.lib.bolap_ceb_autotune(engine.ctx, C.byref(schema.db), arr, len(qs),
                                          C.byref(best)))
    return best.value


class StringPool:
    """Host string pool of libbolap (bolap_string_pool_*): strings -> int32 ids in
    first-appearance order, -1 for nulls, as the reference builder's StringPool
    (bespoke_ceb/builder_impl.hpp:14-89) assigns them to every string column."""

    def __init__(self, lib):
        self.lib = lib
        self.h = C.c_void_p()
        N.check(lib.bolap_string_pool_create(C.byref(self.h)))

    def intern(self, strings, valid=None) -> np.ndarray:
        enc = [s.encode() if isinstance(s, str) else bytes(s) for s in strings]
        chars = np.frombuffer(b"".join(enc) or b"\0", np.uint8)
        offs = np.zeros(len(enc) + 1, np.int64)
        offs[1:] = np.cumsum([len(e) for e in enc])
        ids = np.zeros(max(1, len(enc)), np.int32)
        v = None if valid is None else np.ascontiguousarray(valid, np.uint8)
        N.check(self.lib.bolap_string_pool_intern(self.h, chars.ctypes.data, offs.ctypes.data,
                                                  len(enc),
                                                  None if v is None else v.ctypes.data,
                                                  ids.ctypes.data))
        return ids[:len(enc)]

    def export(self, lower: bool = False) -> tuple[bytes, np.ndarray]:
        n, c = C.c_int64(), C.c_int64()
        N.check(self.lib.bolap_string_pool_size(self.h, C.byref(n), C.byref(c)))
        chars = np.zeros(max(1, c.value), np.uint8)
        offs = np.zeros(n.value + 1, np.int64)
        N.check(self.lib.bolap_string_pool_export(self.h, int(lower), chars.ctypes.data,
                                                  offs.ctypes.data))
        return chars[:c.value].tobytes(), offs

    def close(self):
        if self.h:
            self.lib.bolap_string_pool_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()

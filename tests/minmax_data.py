"""Seeded inputs of the grouped MIN/MAX (the shape of bespoke_ceb/query_q10a.cpp's
result aggregation) and an independent pandas formulation of its result."""
import numpy as np
import pandas as pd


def gen(seed: int, pairs: int, items: int, groups: int, keys: int, max_range: int = 6,
        drop: float = 0.1, null: float = 0.2, year_lo: int = 1880, year_hi: int = 2025):
    """Join rows (group, item), per-item CSR ranges of key values, years with nulls."""
    rng = np.random.default_rng(seed)
    count = rng.integers(0, max_range + 1, items).astype(np.int32)
    offset = np.zeros(items, np.int64)
    if items:
        offset[1:] = np.cumsum(count[:-1])
    values = int(count.astype(np.int64).sum())
    range_value = rng.integers(0, keys, values).astype(np.int32)
    has = (rng.random(items) >= null).astype(np.uint8)
    year = rng.integers(year_lo, year_hi + 1, items).astype(np.int32)
    year[has == 0] = rng.integers(-5, 5, int((has == 0).sum()))  # ignored when null
    group = rng.integers(0, groups, pairs).astype(np.int32)
    group[rng.random(pairs) < drop] = -1
    item = rng.integers(0, max(items, 1), pairs).astype(np.int32)
    return {"pair_group": group, "pair_item": item, "item_offset": offset.astype(np.int32),
            "item_count": count, "item_has": has, "item_value": year,
            "range_value": range_value, "groups": groups, "keys": keys}


def pandas_minmax(d) -> np.ndarray:
    """(group, key, has, min, max) rows in (group, key) order, by expanding the
    join rows and a pandas group-by: has = any row with a value, min/max over
    the rows with a value, 0/0 when none."""
    keep = d["pair_group"] >= 0
    g = d["pair_group"][keep].astype(np.int64)
    it = d["pair_item"][keep].astype(np.int64)
    cnt = d["item_count"][it].astype(np.int64)
    rep_g = np.repeat(g, cnt)
    rep_it = np.repeat(it, cnt)
    starts = np.repeat(d["item_offset"][it].astype(np.int64), cnt)
    within = np.arange(len(rep_it)) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    key = d["range_value"][starts + within]
    has = d["item_has"][rep_it].astype(bool) if d["item_has"] is not None else \
        np.ones(len(rep_it), bool)
    df = pd.DataFrame({"group": rep_g, "key": key, "has": has,
                       "v": np.where(has, d["item_value"][rep_it], np.nan)})
    out = df.groupby(["group", "key"], sort=True).agg(has=("has", "any"), min=("v", "min"),
                                                     max=("v", "max")).reset_index()
    res = np.zeros(len(out), [("group", "<i4"), ("key", "<i4"), ("has", "<i4"),
                              ("min", "<i4"), ("max", "<i4")])
    res["group"], res["key"] = out["group"], out["key"]
    res["has"] = out["has"].astype(np.int32)
    res["min"] = out["min"].fillna(0).astype(np.int64)
    res["max"] = out["max"].fillna(0).astype(np.int64)
    return res

"""Autotuners of the join probe and the radix digit, and the knob table's
persistence: tuning picks a configuration among the candidates and never
changes a result."""
import numpy as np
import pytest

import oracle
from bespokeolap_artifacts_amd import native as N

pytestmark = pytest.mark.gpu

SEED = 42


def test_join_autotune_keeps_the_result(engine):
    d = engine.gen_dim(SEED, 200_000)
    f = engine.gen_fact(SEED, 3_000_001, 200_000)
    preds = [engine.pred(d["cat"], N.OP_RANGE, 7, 7),
             engine.pred(d["cont"], N.OP_IN_SET, values=[3, 17, 29])]
    a = engine.join_sum_args(d["key"], preds, f["fk"], f["price"], f["disc"])
    before = engine.join_sum_run(a)
    try:
        ms = engine.join_autotune(a, reps=2)
        assert ms > 0
        assert 1 <= engine.tune_get("join.probe_variant") <= 5
        assert 6 <= engine.tune_get("join.lds_bloom_bits") <= 8
        assert 1 <= engine.tune_get("join.blocks_per_cu") <= 2
        after = engine.join_sum_run(a)
        for k in ("revenue", "build_rows", "matches"):
            assert after[k] == before[k], k
    finally:
        for k in ("join.probe_variant", "join.lds_bloom_bits", "join.blocks_per_cu"):
            engine.tune_set(k, 0)


@pytest.mark.parametrize("n,bits", [(1_000_003, 26), (300_000, 40)])
def test_sort_autotune_picks_a_width_and_sorts_exactly(engine, n, bits):
    import torch
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64)
    vals = np.arange(n, dtype=np.uint32)
    k = torch.from_numpy(keys.view(np.int64)).cuda()
    v = torch.from_numpy(vals.view(np.int32)).cuda()
    try:
        ms = engine.sort_autotune(k, v, bits, reps=2)
        assert ms > 0
        assert 5 <= engine.tune_get("radix.digit_bits") <= 8
        # the caller's arrays are left as they were
        assert np.array_equal(k.cpu().numpy().view(np.uint64), keys)
        assert np.array_equal(v.cpu().numpy().view(np.uint32), vals)
        engine.radix_sort(k, v, bits)
        ok, ov = oracle.radix_sort_u64(keys, vals, bits)
        assert np.array_equal(k.cpu().numpy().view(np.uint64), ok)
        assert np.array_equal(v.cpu().numpy().view(np.uint32), ov)
    finally:
        engine.tune_set("radix.digit_bits", 0)


@pytest.mark.parametrize("width", [5, 6, 7, 8])
def test_forced_digit_widths_sort_exactly(engine, width):
    import torch
    n, bits = 500_001, 33
    rng = np.random.default_rng(width)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64)
    vals = rng.integers(0, 1 << 31, n).astype(np.uint32)
    k = torch.from_numpy(keys.view(np.int64)).cuda()
    v = torch.from_numpy(vals.view(np.int32)).cuda()
    engine.tune_set("radix.digit_bits", width)
    try:
        engine.radix_sort(k, v, bits)
    finally:
        engine.tune_set("radix.digit_bits", 0)
    ok, ov = oracle.radix_sort_u64(keys, vals, bits)
    assert np.array_equal(k.cpu().numpy().view(np.uint64), ok)
    assert np.array_equal(v.cpu().numpy().view(np.uint32), ov)


def test_knob_table_round_trip(engine, tmp_path):
    path = tmp_path / "knobs.txt"
    engine.tune_set("radix.digit_bits", 7)
    engine.tune_set("join.probe_variant", 4)
    try:
        engine.tune_save(path)
        text = path.read_text()
        assert "radix.digit_bits 7" in text and "join.probe_variant 4" in text
        engine.tune_set("radix.digit_bits", 5)
        engine.tune_set("join.probe_variant", 1)
        engine.tune_load(path)
        assert engine.tune_get("radix.digit_bits") == 7
        assert engine.tune_get("join.probe_variant") == 4
        bad = tmp_path / "bad.txt"
        bad.write_text("radix.digit_bits 6\nthis is not a knob line\n")
        with pytest.raises(N.BolapError):
            engine.tune_load(bad)
        assert engine.tune_get("radix.digit_bits") == 7  # unchanged by a bad file
        with pytest.raises(N.BolapError):
            engine.tune_load(tmp_path / "missing.txt")
    finally:
        engine.tune_set("radix.digit_bits", 0)
        engine.tune_set("join.probe_variant", 0)

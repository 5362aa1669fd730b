"""ILIKE over a dictionary in device memory (bolap_ilike_mask_device, used by
CEB 10a's name filter) against the host form (bolap_ilike_mask, the
reference's classification and matching, query_q10a.cpp:43-209) on random
strings and the pattern forms: match-all, whole string, prefix, suffix,
substring, '_' and '%' anywhere, literal '%' and '_' in the text, case."""
import ctypes as C

import numpy as np
import pytest
import torch

from bespokeolap_artifacts_amd import native as N

pytestmark = pytest.mark.gpu

PATTERNS = ["%", "%%", "", "a", "ab%", "%ab", "%ab%", "a%b", "a_c", "_", "%a_%", "%%a%%",
            "ABC", "%b%c%", "x%", "%_%", "a%%", "%%b", "%a%b%c%", "100%", "%\\%", "_%_",
            "%son 1%", "person 1_"]


def words(rng, n):
    alpha = np.array(list("abcABC_%x 1"))
    out = []
    for _ in range(n):
        k = int(rng.integers(0, 12))
        out.append("".join(alpha[rng.integers(0, len(alpha), k)]))
    return out + ["", "%", "_", "a%b", "Person 12", "person 1x", "ab%x"]


def test_device_ilike_matches_host(engine):
    rng = np.random.default_rng(7)
    ws = words(rng, 20_000)
    raw = [w.encode() for w in ws]
    off = np.concatenate([[0], np.cumsum([len(b) for b in raw])]).astype(np.int64)
    chars = np.frombuffer(b"".join(raw) + b"\0", np.uint8)
    lib = N.load()
    d_chars = torch.from_numpy(chars.copy()).cuda()
    d_off = torch.from_numpy(off).cuda()
    d_mask = torch.zeros(len(ws), dtype=torch.uint8, device="cuda")
    host = np.zeros(len(ws), np.uint8)
    matched = 0
    for p in PATTERNS:
        N.check(lib.bolap_ilike_mask(chars.ctypes.data_as(C.c_char_p),
                                     off.ctypes.data_as(C.POINTER(C.c_int64)), len(ws),
                                     p.encode(), host.ctypes.data_as(C.POINTER(C.c_uint8))))
        N.check(lib.bolap_ilike_mask_device(engine.ctx, d_chars.data_ptr(), d_off.data_ptr(),
                                            len(ws), p.encode(), d_mask.data_ptr()))
        torch.cuda.synchronize()
        assert np.array_equal(d_mask.cpu().numpy(), host), p
        matched += 0 < host.sum() < len(ws)
    assert matched >= 15
    with pytest.raises(RuntimeError, match="pattern"):
        N.check(lib.bolap_ilike_mask_device(engine.ctx, d_chars.data_ptr(), d_off.data_ptr(),
                                            len(ws), b"a" * 300, d_mask.data_ptr()))

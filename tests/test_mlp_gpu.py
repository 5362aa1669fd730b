"""The estimator network on the GPU (BF16 MFMA) vs the CPU reference."""
import ctypes as C

import numpy as np
import pytest

from bespokeolap_artifacts_amd import native as N
from oracle import mlp

pytestmark = pytest.mark.gpu


def make(engine, seed=1):
    h = C.c_void_p()
    N.check(engine.lib.bolap_mlp_create(engine.ctx, seed, C.byref(h)))
    return h


def params_of(engine, h):
    n = engine.lib.bolap_mlp_num_params()
    buf = (C.c_float * n)()
    N.check(engine.lib.bolap_mlp_get_params(h, buf))
    return np.frombuffer(buf, np.float32).astype(np.float64).copy()


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def test_forward_matches_cpu_reference(engine):
    h = make(engine)
    p = params_of(engine, h)
    rng = np.random.default_rng(2)
    for n in (1, 17, 256, 1000):
        x = rng.normal(0, 1, (n, 256)).astype(np.float32)
        y = dev(np.zeros((n, 3)))
        N.check(engine.lib.bolap_mlp_predict(h, C.c_void_p(dev(x).data_ptr()), n,
                                             C.c_void_p(y.data_ptr())))
        got = y.cpu().numpy().astype(np.float64)
        want, _ = mlp.forward(p, x)
        err = np.abs(got - want).max() / (np.abs(want).max() + 1e-9)
        assert err < 1e-2, (n, err)
    engine.lib.bolap_mlp_destroy(h)


def test_gradients_match_cpu_reference(engine):
    h = make(engine, seed=7)
    p = params_of(engine, h)
    rng = np.random.default_rng(3)
    n = 64
    x = rng.normal(0, 1, (n, 256)).astype(np.float32)
    t = rng.normal(0, 1, (n, 3)).astype(np.float32)
    g = (C.c_float * len(p))()
    loss = C.c_float(0)
    xd, td = dev(x), dev(t)
    N.check(engine.lib.bolap_mlp_gradients(h, C.c_void_p(xd.data_ptr()),
                                           C.c_void_p(td.data_ptr()), n, g, C.byref(loss)))
    got = np.frombuffer(g, np.float32).astype(np.float64)
    want_loss, want = mlp.loss_and_grads(p, x, t)
    assert abs(loss.value - want_loss) <= 1e-3 * abs(want_loss)
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert rel < 2e-2, rel
    # per-layer check as well (a transposed or misplaced block fails here)
    i = 0
    for fin, fout in mlp.DIMS:
        for size in (fin * fout, fout):
            a, b = got[i:i + size], want[i:i + size]
            assert np.linalg.norm(a - b) <= 3e-2 * np.linalg.norm(b) + 1e-7
            i += size
    engine.lib.bolap_mlp_destroy(h)


def test_training_reduces_loss(engine):
    h = make(engine, seed=11)
    rng = np.random.default_rng(4)
    n = 2048
    x = rng.normal(0, 1, (n, 256)).astype(np.float32)
    wtrue = rng.normal(0, 0.1, (256, 3))
    t = np.tanh(x @ wtrue).astype(np.float32)
    xd, td = dev(x), dev(t)
    losses = []
    for step in range(300):
        loss = C.c_float(0)
        N.check(engine.lib.bolap_mlp_train_step(h, C.c_void_p(xd.data_ptr()),
                                                C.c_void_p(td.data_ptr()), n, 1e-3,
                                                C.byref(loss)))
        losses.append(loss.value)
    assert losses[-1] < 0.2 * losses[0], (losses[0], losses[-1])
    engine.lib.bolap_mlp_destroy(h)

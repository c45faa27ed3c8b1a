"""The float64 surrogate restatement (oracle/unet_ref.py) pinned by central
finite differences of its own loss, plus shape and parameter-count checks."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import unet_ref as U  # noqa: E402


def _volume(n=8, seed=0):
    rng = np.random.default_rng(seed)
    occ = (rng.random((n, n, n)) < 0.6).astype(np.uint8)
    occ[0] = occ[-1] = 0
    edt = (rng.integers(0, 10, size=occ.shape) * occ).astype(np.int32)
    lam2 = (rng.integers(1, 60, size=occ.shape) * occ).astype(np.uint32)
    return occ, edt, lam2


def test_param_count():
    # 27 * sum(Cin * Cout) + biases + head, the 0.68 M parameter network
    assert U.num_params() == 27 * 25120 + 512 + 34 == 678786


def test_forward_shapes():
    occ, edt, _ = _volume(16)
    th = U.init_params(7)
    out = U.forward(th, U.inputs(occ, edt))
    assert out.shape == (16, 16, 16, 2)


def test_gradients_match_finite_differences():
    occ, edt, lam2 = _volume(8, 1)
    rng = np.random.default_rng(2)
    th = U.init_params(3).astype(np.float64)
    # larger biases so that ReLUs are not all dead on random inputs
    th += rng.normal(0, 0.05, size=th.shape)
    L, g = U.gradients(th, occ, edt, lam2)
    assert np.isfinite(L) and g.shape == th.shape
    # the error target is a constant of the loss: freeze it at theta
    e = U.error_target(U.forward(th, U.inputs(occ, edt)), lam2)
    # parameters spread over every layer
    idx = np.unique(np.linspace(0, len(th) - 1, 60).astype(int))
    h = 1e-6
    for i in idx:
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (U.loss_only(tp, occ, edt, lam2, e) - U.loss_only(tm, occ, edt, lam2, e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])


def test_pool_up_adjoint():
    """<pool(x), y> = <x, pool_back(y)> and <up(x), y> = <x, up_back(y)>."""
    rng = np.random.default_rng(4)
    x = rng.normal(size=(8, 8, 8, 3))
    y = rng.normal(size=(4, 4, 4, 3))
    assert np.isclose((U.pool(x) * y).sum(), (x * U.pool_back(y)).sum())
    assert np.isclose((U.up(y) * x).sum(), (y * U.up_back(x)).sum())


def test_init_is_he_uniform():
    th = U.init_params(11)
    W, b = U.unpack(th)["L1a"]
    a = np.sqrt(6.0 / (27 * 16))
    assert np.all(np.abs(W) <= a) and np.all(b == 0)
    assert abs(W.std() - a / np.sqrt(3)) < 0.05 * a

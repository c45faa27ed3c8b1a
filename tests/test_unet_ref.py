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


def test_round_bf16_known_patterns():
    """Ties to even and carries into the exponent, at exact bit patterns."""
    two = 2.0
    cases = [
        (1.0, 1.0),
        (1.0 + two ** -8, 1.0),                      # tie, 1.0 is even
        (1.0 + 3 * two ** -8, 1.0 + two ** -6),      # tie, odd -> up
        (1.0 + two ** -8 + two ** -20, 1.0 + two ** -7),
        (1.0 + two ** -8 - two ** -20, 1.0),
        (-3.0, -3.0),
        (0.0, 0.0),
        (2.0 - two ** -9, 2.0),                      # carry into the exponent
        (255.5, 256.0),                              # 0x437F8000: tie, odd -> up
        (1e-30, np.float32(1e-30).view(np.uint32) & 0xFFFF0000),
    ]
    for x, want in cases:
        if isinstance(want, (int, np.integer)):
            u = int(want)
            if (np.float32(x).view(np.uint32) & 0xFFFF) > 0x8000:
                u += 0x10000
            want = float(np.uint32(u).view(np.float32))
        assert U.round_bf16(x) == want, (x, U.round_bf16(x), want)


def test_round_bf16_is_nearest_tie_even():
    """Against a direct choice between the two BF16 neighbours of each
    float32 (the truncation and the next BF16 up in magnitude): the nearer
    one, the even mantissa on a tie."""
    rng = np.random.default_rng(9)
    f = (rng.standard_normal(20000) * 10.0 ** rng.integers(-6, 6, 20000)).astype(np.float32)
    f[:64] = np.float32(1.0) + np.arange(64, dtype=np.float32) * np.float32(2.0 ** -13)
    u = f.view(np.uint32)
    lo = (u & 0xFFFF0000).view(np.float32).astype(np.float64)
    hi = ((u & 0xFFFF0000) + 0x10000).astype(np.uint32).view(np.float32).astype(np.float64)
    x = f.astype(np.float64)
    dl, dh = np.abs(x - lo), np.abs(hi - x)
    lo_even = ((u >> 16) & 1) == 0
    want = np.where(dl < dh, lo, np.where(dh < dl, hi, np.where(lo_even, lo, hi)))
    assert np.array_equal(U.round_bf16(x), want)


def test_bf16_variant_close_to_float64():
    """The BF16 variant follows the float64 path up to BF16 rounding: loss
    within 2 %, gradient cosine above 0.99 for the whole vector."""
    occ, edt, lam2 = _volume(8, 3)
    rng = np.random.default_rng(4)
    th = U.init_params(5).astype(np.float64) + rng.normal(0, 0.05, size=U.num_params())
    th = th.astype(np.float32).astype(np.float64)
    L, g = U.gradients(th, occ, edt, lam2)
    Lb, gb = U.gradients(th, occ, edt, lam2, bf16=True)
    assert abs(Lb - L) <= 0.02 * abs(L)
    assert g @ gb / (np.linalg.norm(g) * np.linalg.norm(gb)) > 0.99
    assert not np.array_equal(g, gb)

"""CPU float64 restatement of the surrogate network (test infrastructure: the
checker for the GPU kernels, never part of the product path).

The network is project-defined (no reference counterpart): a 4-level 3-D
U-Net mapping (occupancy, distance / 32) to (lambda / 32, log relative error)
per voxel.  Every definition below is the one the HIP kernels implement:

  conv      out[v] = sum_tap in[v + tap] @ W[tap] + b, zero outside the
            volume, tap index (dx+1)*9 + (dy+1)*3 + (dz+1); ReLU after every
            3^3 conv
  levels    L0a 2->16, L0b 16->16 | avg-pool 2^3 | L1a 16->32, L1b 32->32 |
            pool | L2a 32->48, L2b 48->48 | pool | L3a 48->64, L3b 64->64 |
            nearest up x2, concat [up, skip] | L4a 112->48, L4b 48->48 | up,
            concat | L5a 80->32, L5b 32->32 | up, concat | L6a 48->16,
            L6b 16->16 | head 1^3 16->2 (linear)
  inputs    x0 = [occ, sqrt(max(edt, 0)) / 32]; dims must be multiples of 8
  target    t = sqrt(lam2) / 32 on inside voxels
  loss      mean over inside voxels of ((p0 - t) / (t + 1/32))^2
            + 0.1 * (p1 - e)^2,  e = log(|p0 - t| / (t + 1/32) + 1e-4) held
            constant
  params    flat float32 vector: per conv W[27][Cin][Cout] then b[Cout], in
            layer order, then head W[16][2], b[2]

Parity of this restatement is pinned by central finite differences of the
loss in float64 (tests/test_unet_ref.py), not by the GPU path.

BF16 variant (forward / gradients with bf16=True): the same algebra with
every tensor the kernels store in BF16 rounded to the nearest BF16 (ties to
even) of its float32 value where they store it -- the network input, the
weights the convolutions read, each conv output after bias and ReLU, each
pooled tensor, the head's input gradient, each input gradient of a conv,
the up-sampling gradient sums and the pooled-gradient accumulations; sums
stay in float64 (the kernels accumulate in float32), the head output and
dL/d(out) in float32.  It reproduces the kernels' rounding points so their
gradients can be held to a percent-level bound.  round_bf16 is pinned by
exact bit patterns (tests/test_unet_ref.py); the algebra is the float64
path's.
"""
import numpy as np

LAYERS = [("L0a", 2, 16), ("L0b", 16, 16), ("L1a", 16, 32), ("L1b", 32, 32),
          ("L2a", 32, 48), ("L2b", 48, 48), ("L3a", 48, 64), ("L3b", 64, 64),
          ("L4a", 112, 48), ("L4b", 48, 48), ("L5a", 80, 32), ("L5b", 32, 32),
          ("L6a", 48, 16), ("L6b", 16, 16)]
HEAD = (16, 2)
SCALE = 32.0
EPS_T = 1.0 / SCALE
ERR_W = 0.1
ERR_FLOOR = 1e-4


def num_params():
    return sum(27 * ci * co + co for _, ci, co in LAYERS) + HEAD[0] * HEAD[1] + HEAD[1]


def unpack(theta):
    theta = np.asarray(theta, np.float64)
    out, o = {}, 0
    for name, ci, co in LAYERS:
        n = 27 * ci * co
        out[name] = (theta[o:o + n].reshape(27, ci, co), theta[o + n:o + n + co])
        o += n + co
    n = HEAD[0] * HEAD[1]
    out["head"] = (theta[o:o + n].reshape(HEAD), theta[o + n:o + n + HEAD[1]])
    return out


def pack(grads):
    parts = []
    for name, _, _ in LAYERS:
        parts += [grads[name][0].ravel(), grads[name][1].ravel()]
    parts += [grads["head"][0].ravel(), grads["head"][1].ravel()]
    return np.concatenate(parts)


TAPS = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]


def round_bf16(x):
    """Nearest bfloat16 of float32(x), ties to even, as float64 (finite
    values)."""
    f = np.ascontiguousarray(np.asarray(x, np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _ident(x):
    return x


def shift(x, d):
    """y[v] = x[v + d], zero outside."""
    y = np.zeros_like(x)
    X, Y, Z = x.shape[:3]
    sl_dst, sl_src = [], []
    for n, dd in zip((X, Y, Z), d):
        sl_dst.append(slice(max(0, -dd), n - max(0, dd)))
        sl_src.append(slice(max(0, dd), n - max(0, -dd)))
    y[tuple(sl_dst)] = x[tuple(sl_src)]
    return y


def conv(x, W, b):
    out = np.zeros(x.shape[:3] + (W.shape[2],))
    for t, d in enumerate(TAPS):
        out += shift(x, d) @ W[t]
    return out + b


def conv_back(x, W, dz):
    """dW, db, dx of conv given dL/d(out) = dz."""
    dW = np.zeros_like(W)
    dx = np.zeros_like(x)
    for t, d in enumerate(TAPS):
        xs = shift(x, d)
        dW[t] = np.einsum("xyzc,xyzo->co", xs, dz)
        dx += shift(dz @ W[t].T, tuple(-c for c in d))
    return dW, dz.sum(axis=(0, 1, 2)), dx


def pool(x):
    X, Y, Z, C = x.shape
    return x.reshape(X // 2, 2, Y // 2, 2, Z // 2, 2, C).mean(axis=(1, 3, 5))


def pool_back(dp):
    return np.repeat(np.repeat(np.repeat(dp, 2, 0), 2, 1), 2, 2) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_back(du):
    X, Y, Z, C = du.shape
    return du.reshape(X // 2, 2, Y // 2, 2, Z // 2, 2, C).sum(axis=(1, 3, 5))


def inputs(occ, edt):
    occ = np.asarray(occ)
    x = np.zeros(occ.shape + (2,))
    x[..., 0] = occ > 0
    x[..., 1] = np.sqrt(np.maximum(np.asarray(edt, np.float64), 0.0)) / SCALE
    return x


def forward(theta, x0, keep=False, bf16=False):
    rb = round_bf16 if bf16 else _ident
    P = unpack(theta)
    acts = {}
    x0 = rb(x0)

    def cr(name, x):
        W, b = P[name]
        y = rb(np.maximum(conv(x, rb(W), b), 0.0))
        acts[name] = (x, y)
        return y

    def pl(x):
        return rb(pool(x))

    b0 = cr("L0b", cr("L0a", x0))
    b1 = cr("L1b", cr("L1a", pl(b0)))
    b2 = cr("L2b", cr("L2a", pl(b1)))
    b3 = cr("L3b", cr("L3a", pl(b2)))
    d2 = cr("L4b", cr("L4a", np.concatenate([up(b3), b2], axis=3)))
    d1 = cr("L5b", cr("L5a", np.concatenate([up(d2), b1], axis=3)))
    d0 = cr("L6b", cr("L6a", np.concatenate([up(d1), b0], axis=3)))
    Wh, bh = P["head"]
    out = d0 @ Wh + bh
    if bf16:
        out = out.astype(np.float32).astype(np.float64)
    if keep:
        return out, acts, P
    return out


def error_target(out, lam2):
    t = np.sqrt(np.asarray(lam2, np.float64)) / SCALE
    return np.log(np.abs((out[..., 0] - t) / (t + EPS_T)) + ERR_FLOOR)


def loss_and_dout(out, occ, lam2, e=None):
    """Loss and dL/d(out); e (the error target) is held constant: pass it to
    evaluate the loss at a perturbed output with the same target."""
    inside = np.asarray(occ) > 0
    n = max(int(inside.sum()), 1)
    t = np.sqrt(np.asarray(lam2, np.float64)) / SCALE
    den = t + EPS_T
    r = (out[..., 0] - t) / den
    if e is None:
        e = error_target(out, lam2)
    q = out[..., 1] - e
    L = (np.where(inside, r * r + ERR_W * q * q, 0.0)).sum() / n
    dout = np.zeros_like(out)
    dout[..., 0] = np.where(inside, 2.0 * r / den, 0.0) / n
    dout[..., 1] = np.where(inside, 2.0 * ERR_W * q, 0.0) / n
    return L, dout


def gradients(theta, occ, edt, lam2, bf16=False):
    """(loss, flat float64 gradient) of the network on one volume; with
    bf16, the BF16 variant (the kernels' rounding points)."""
    rb = round_bf16 if bf16 else _ident
    x0 = inputs(occ, edt)
    out, acts, P = forward(theta, x0, keep=True, bf16=bf16)
    L, dout = loss_and_dout(out, occ, lam2)
    if bf16:
        dout = dout.astype(np.float32).astype(np.float64)
    G = {}
    Wh, _ = P["head"]
    d0 = acts["L6b"][1]
    G["head"] = (np.einsum("xyzc,xyzo->co", d0, dout), dout.sum(axis=(0, 1, 2)))
    dy = rb(dout @ Wh.T)

    def back(name, dy):
        x, y = acts[name]
        dz = dy * (y > 0)
        dW, db, dx = conv_back(x, rb(P[name][0]), dz)
        G[name] = (dW, db)
        return rb(dx)

    def split(dcat, c_up):
        return rb(up_back(dcat[..., :c_up])), dcat[..., c_up:]

    dcat0 = back("L6a", back("L6b", dy))
    dd1, db0 = split(dcat0, 32)
    dcat1 = back("L5a", back("L5b", dd1))
    dd2, db1 = split(dcat1, 48)
    dcat2 = back("L4a", back("L4b", dd2))
    db3, db2 = split(dcat2, 64)
    db2 = rb(db2 + pool_back(back("L3a", back("L3b", db3))))
    db1 = rb(db1 + pool_back(back("L2a", back("L2b", db2))))
    db0 = rb(db0 + pool_back(back("L1a", back("L1b", db1))))
    back("L0a", back("L0b", db0))
    return L, pack(G)


def loss_only(theta, occ, edt, lam2, e=None):
    return loss_and_dout(forward(theta, inputs(occ, edt)), occ, lam2, e)[0]


def splitmix64(x):
    """splitmix64 on uint64 numpy arrays (wrapping arithmetic)."""
    x = np.asarray(x, np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def init_params(seed):
    """He-uniform weights, zero biases, the head scaled by 0.1: parameter i
    of a weight block of fan-in F is (2u - 1) sqrt(6 / F), u = top 24 bits of
    splitmix64(seed * 2^32 + i) / 2^24 (i = index in the flat vector), all in
    float32 arithmetic."""
    theta = np.zeros(num_params(), np.float32)
    o = 0
    for _, ci, co in LAYERS + [("head", HEAD[0], HEAD[1])]:
        head = (ci, co) == HEAD
        taps = 1 if head else 27
        n = taps * ci * co
        a = np.float32(np.sqrt(np.float32(6.0) / np.float32(taps * ci)))
        if head:
            a = np.float32(a * np.float32(0.1))
        idx = np.arange(o, o + n, dtype=np.uint64) + np.uint64((seed & 0xFFFFFFFF) << 32)
        u = (splitmix64(idx) >> np.uint64(40)).astype(np.float32) * np.float32(1.0 / (1 << 24))
        theta[o:o + n] = (np.float32(2.0) * u - np.float32(1.0)) * a
        o += n + co
    return theta

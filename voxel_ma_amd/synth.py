"""Seeded synthetic volumes for the five workloads (inputs, not the product).

All volumes are uint8 occupancy, C order [x][y][z]; voxel (i, j, k) is
centred at integer coordinates (i, j, k), as in the reference Surfacer
(corners at +-0.5, include/surfacing.h:97-119).

  torus_c1     64^3 solid torus from a 64x32-segment triangle mesh (4,096
               triangles), R=20, r=8, centre (32,32,32), axis z; a voxel is
               inside iff its centre is inside the mesh (ray parity along +x;
               rays are offset by (dy, dz) = (1.3e-6, 2.9e-6) voxels so that no
               ray meets a mesh vertex or edge).
  spheres_c2   256^3 union of 32 balls, centres uniform integer in [0,256)^3,
               radii uniform integer in [8,40], |p-c|^2 <= r^2 exact; a 4-voxel
               empty margin is enforced.
  blurred noise  n^3 (512 in the benchmark) N(0,1) noise, separable
               Gaussian blur sigma=12 (97 taps = truncate at 4 sigma, periodic
               boundary), inside iff value > the exact 70th-percentile order
               statistic, then a 1-voxel empty shell.  Generated on the device
               by vc_synth_blurred_noise (include/vc_synth.h); blur_c3_host is
               its numpy restatement for small sizes.
  batch_c4     256 cubic volumes, L = round(32 * 6^u), u ~ U(0,1), each a union
               of 4-16 capsules/ellipsoids (semi-axes/radius U[4,30], uniform
               random rotation, centres inside the grid), 1-voxel margin.
"""
import math

import numpy as np

_DY, _DZ = 1.3e-6, 2.9e-6


def torus_mesh(R=20.0, r=8.0, nu=64, nv=32, centre=(32.0, 32.0, 32.0)):
    """Vertices (nu*nv, 3) and triangles (2*nu*nv, 3) of a torus around z."""
    u = 2.0 * np.pi * np.arange(nu) / nu
    v = 2.0 * np.pi * np.arange(nv) / nv
    U, V = np.meshgrid(u, v, indexing="ij")
    X = (R + r * np.cos(V)) * np.cos(U) + centre[0]
    Y = (R + r * np.cos(V)) * np.sin(U) + centre[1]
    Z = r * np.sin(V) + centre[2]
    verts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    tris = []
    for i in range(nu):
        for j in range(nv):
            a = i * nv + j
            b = ((i + 1) % nu) * nv + j
            c = ((i + 1) % nu) * nv + (j + 1) % nv
            d = i * nv + (j + 1) % nv
            tris.append((a, b, c))
            tris.append((a, c, d))
    return verts, np.array(tris, dtype=np.int64)


def voxelize_mesh(verts, tris, shape):
    """Inside test of voxel centres by ray parity along +x."""
    nx, ny, nz = shape
    crossings = [[[] for _ in range(nz)] for _ in range(ny)]
    P = verts[tris]  # (T, 3, 3)
    for t in range(len(tris)):
        a, b, c = P[t]
        ya, za = a[1], a[2]
        e1 = b - a
        e2 = c - a
        det = e1[1] * e2[2] - e1[2] * e2[1]
        if det == 0.0:
            continue
        ymin, ymax = min(a[1], b[1], c[1]), max(a[1], b[1], c[1])
        zmin, zmax = min(a[2], b[2], c[2]), max(a[2], b[2], c[2])
        for j in range(max(0, math.ceil(ymin - _DY)), min(ny - 1, math.floor(ymax - _DY)) + 1):
            for k in range(max(0, math.ceil(zmin - _DZ)), min(nz - 1, math.floor(zmax - _DZ)) + 1):
                py, pz = j + _DY - ya, k + _DZ - za
                s = (py * e2[2] - pz * e2[1]) / det
                w = (e1[1] * pz - e1[2] * py) / det
                if s < 0.0 or w < 0.0 or s + w > 1.0:
                    continue
                crossings[j][k].append(a[0] + s * e1[0] + w * e2[0])
    occ = np.zeros(shape, np.uint8)
    xs = np.arange(nx, dtype=np.float64)
    for j in range(ny):
        for k in range(nz):
            cr = crossings[j][k]
            if not cr:
                continue
            cr = np.sort(np.array(cr))
            n_right = len(cr) - np.searchsorted(cr, xs, side="right")
            occ[:, j, k] = (n_right % 2 == 1)
    return occ


def torus_c1():
    v, t = torus_mesh()
    return voxelize_mesh(v, t, (64, 64, 64))


def spheres_c2(n=256, count=32, rmin=8, rmax=40, seed=2, margin=4):
    rng = np.random.default_rng(seed)
    centres = rng.integers(0, n, size=(count, 3))
    radii = rng.integers(rmin, rmax + 1, size=count)
    occ = np.zeros((n, n, n), np.uint8)
    for (cx, cy, cz), r in zip(centres, radii):
        lo = np.maximum([cx - r, cy - r, cz - r], 0)
        hi = np.minimum([cx + r + 1, cy + r + 1, cz + r + 1], n)
        gx = np.arange(lo[0], hi[0])[:, None, None] - cx
        gy = np.arange(lo[1], hi[1])[None, :, None] - cy
        gz = np.arange(lo[2], hi[2])[None, None, :] - cz
        ball = gx * gx + gy * gy + gz * gz <= r * r
        occ[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] |= ball.astype(np.uint8)
    if margin:
        occ[:margin] = 0
        occ[-margin:] = 0
        occ[:, :margin] = 0
        occ[:, -margin:] = 0
        occ[:, :, :margin] = 0
        occ[:, :, -margin:] = 0
    return occ


def gaussian_taps(sigma=12.0, truncate=4.0):
    rad = int(truncate * sigma + 0.5)
    x = np.arange(-rad, rad + 1, dtype=np.float64)
    w = np.exp(-0.5 * (x / sigma) ** 2)
    return (w / w.sum()).astype(np.float32)


def blur_c3_host(n=64, sigma=12.0, fill=0.30, seed=3):
    """Host (numpy, float64) restatement of the device generator
    vc_synth_blurred_noise for small n: the same noise bits, blur taps and
    exact order-statistic threshold.  Blurred values can differ from the
    device's float32 sums in the last bits, so a few voxels at the threshold
    may differ."""
    N = n * n * n
    i = np.arange(N, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = np.uint64(seed) * np.uint64(0x632BE59BD9B4E019) + i
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        r = x ^ (x >> np.uint64(31))
    u1 = ((r >> np.uint64(40)).astype(np.float64) + 1.0) / 16777216.0
    u2 = ((r >> np.uint64(16)) & np.uint64(0xFFFFFF)).astype(np.float64) / 16777216.0
    v = (np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)).reshape(n, n, n)
    taps = gaussian_taps(sigma).astype(np.float64)
    rad = (len(taps) - 1) // 2
    for axis in (2, 1, 0):
        acc = np.zeros_like(v)
        for t, w in enumerate(taps):
            acc += w * np.roll(v, rad - t, axis=axis)
        v = acc
    k = int(math.ceil((1.0 - fill) * N))
    thr = np.partition(v.ravel(), k - 1)[k - 1]
    occ = (v > thr).astype(np.uint8)
    occ[0] = occ[-1] = 0
    occ[:, 0] = occ[:, -1] = 0
    occ[:, :, 0] = occ[:, :, -1] = 0
    return occ


def _rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


SOLID_WORDS = 16  # VC_SOLID_WORDS (include/vc_synth.h)


def solid_specs(L, rng):
    """Primitive records (k, 16) of one batched-solids volume of edge L: 4-16
    primitives, each an ellipsoid (semi-axes U[4,30]) or a capsule (radius and
    half-length U[4,30]) with probability 1/2, uniform random rotation, centre
    uniform in [0, L-1]^3.  Layout as vc_synth_solids."""
    k = int(rng.integers(4, 17))
    out = np.zeros((k, SOLID_WORDS), np.float64)
    for i in range(k):
        c = rng.uniform(0, L - 1, size=3)
        Rm = _rotation(rng)
        out[i, 1:4] = c
        out[i, 4:13] = Rm.ravel()
        if rng.random() < 0.5:
            out[i, 0] = 0.0
            out[i, 13:16] = rng.uniform(4, 30, size=3)
        else:
            out[i, 0] = 1.0
            out[i, 13] = rng.uniform(4, 30)
            out[i, 14] = rng.uniform(4, 30)
    return out


def solids_host(L, prims):
    """numpy restatement of vc_synth_solids for a cube of edge L: the same
    IEEE double operations in the same order (elementwise, no matmul)."""
    g = np.indices((L, L, L)).astype(np.float64)
    occ = np.zeros((L, L, L), bool)
    for r in prims:
        dx, dy, dz = g[0] - r[1], g[1] - r[2], g[2] - r[3]
        q = [(dx * r[4 + j] + dy * r[7 + j]) + dz * r[10 + j] for j in range(3)]
        if r[0] == 0.0:
            t0, t1, t2 = q[0] / r[13], q[1] / r[14], q[2] / r[15]
            occ |= ((t0 * t0 + t1 * t1) + t2 * t2) <= 1.0
        elif r[0] == 2.0:
            rho = np.sqrt(q[0] * q[0] + q[1] * q[1]) - r[13]
            occ |= (rho * rho + q[2] * q[2]) <= r[14] * r[14]
        else:
            h = r[14]
            e = q[0] - np.clip(q[0], -h, h)
            occ |= ((e * e + q[1] * q[1]) + q[2] * q[2]) <= r[13] * r[13]
    occ = occ.astype(np.uint8)
    occ[0] = occ[-1] = 0
    occ[:, 0] = occ[:, -1] = 0
    occ[:, :, 0] = occ[:, :, -1] = 0
    return occ


def training_specs(L, rng):
    """One training volume of edge L in the style of the torus and spheres
    workloads, rescaled: with probability 1/2 a union of 1-3 tori (major
    radius U[0.12, 0.3] L, minor U[0.04, 0.11] L, random rotation), else a
    union of 4-16 balls (radius U[0.05, 0.25] L); centres uniform in the
    middle 60% of the grid."""
    out = []
    lo, hi = 0.2 * (L - 1), 0.8 * (L - 1)
    if rng.random() < 0.5:
        for _ in range(int(rng.integers(1, 4))):
            r = np.zeros(SOLID_WORDS)
            r[0] = 2.0
            r[1:4] = rng.uniform(lo, hi, size=3)
            r[4:13] = _rotation(rng).ravel()
            r[13] = rng.uniform(0.12, 0.3) * L
            r[14] = rng.uniform(0.04, 0.11) * L
            out.append(r)
    else:
        for _ in range(int(rng.integers(4, 17))):
            r = np.zeros(SOLID_WORDS)
            r[1:4] = rng.uniform(lo, hi, size=3)
            r[4:13] = np.eye(3).ravel()
            r[13:16] = rng.uniform(0.05, 0.25) * L
            out.append(r)
    return np.array(out)


def batch_c4_specs(count=256, seed=4, lmin=32, ratio=6.0):
    """(L, primitive records) of each volume of the batched-solids workload:
    L = round(lmin * ratio^u), u ~ U(0,1)."""
    rng = np.random.default_rng(seed)
    sizes = [int(round(lmin * ratio ** rng.random())) for _ in range(count)]
    return [(L, solid_specs(L, rng)) for L in sizes]


def batch_c4(count=256, seed=4, lmin=32, ratio=6.0):
    """Host-generated volumes of batch_c4_specs (small counts in tests)."""
    return [solids_host(L, p) for L, p in batch_c4_specs(count, seed, lmin, ratio)]


DEFAULT_TT = (0.01, 0.02, 0.04)


def inside_size(occ):
    """Longest side of the bounding box of the inside voxel centres (float32),
    the voxel form of VoroInfo::getInsidePartSize (reference
    src/voroinfo.cpp:496)."""
    occ = np.asarray(occ)
    sizes = []
    for axis in range(3):
        other = tuple(a for a in range(3) if a != axis)
        idx = np.nonzero(occ.any(axis=other))[0]
        sizes.append(0 if len(idx) == 0 else int(idx[-1] - idx[0]))
    return np.float32(max(sizes))


def levels_from_tt(size, tt=DEFAULT_TT):
    """thresh = float(insideSize * (double)t) per -tt value, ascending, as the
    threshold loop of exportInsideVoroMesh (reference
    src/highlevelalgo.cpp:825-836)."""
    return sorted(float(np.float32(np.float64(np.float32(size)) * float(t))) for t in tt)

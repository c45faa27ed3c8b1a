"""CPU tests: pin the oracle against golden vectors, closed forms, brute force
and independent definitions (no GPU)."""
import itertools
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_envelope_matches_bruteforce(oracle):
    rng = np.random.default_rng(7)
    for _ in range(3000):
        n = int(rng.integers(1, 14))
        f = rng.integers(0, 8, size=n) * rng.integers(0, 2, size=n)
        f = np.where(rng.random(n) < 0.3, -1, f).astype(np.int64)
        D, a = oracle.envelope_1d(f)
        for j in range(n):
            cand = [(f[i] + (j - i) ** 2, i) for i in range(n) if f[i] >= 0]
            if not cand:
                assert D[j] == -1 and a[j] == -1
            else:
                best = min(cand)
                assert (D[j], a[j]) == best


@pytest.mark.parametrize("seed", range(6))
def test_edt_matches_bruteforce(oracle, seed):
    rng = np.random.default_rng(seed)
    for _ in range(25):
        shape = tuple(int(s) for s in rng.integers(1, 10, size=3))
        occ = (rng.random(shape) < rng.random()).astype(np.uint8)
        d, f = oracle.edt_ft(occ)
        db, fb = oracle.edt_brute(occ)
        assert np.array_equal(d, db) and np.array_equal(f, fb)


def test_edt_matches_scipy_golden(oracle):
    z = np.load(os.path.join(HERE, "golden", "edt_scipy.npz"))
    names = sorted({k.split("__")[0] for k in z.files})
    assert len(names) >= 4
    for name in names:
        occ, d2 = z[name + "__occ"], z[name + "__d2"]
        d, f = oracle.edt_ft(occ)
        assert np.array_equal(d, d2), name
        # the feature is a background voxel at exactly that distance
        idx = np.indices(occ.shape).reshape(3, -1)
        fi = np.stack(np.unravel_index(f.ravel(), occ.shape))
        assert (occ.ravel()[f.ravel()] == 0).all()
        assert np.array_equal(((idx - fi) ** 2).sum(0), d.ravel())


def test_edt_closed_forms(oracle):
    occ = np.ones((9, 7, 11), np.uint8)
    occ[4, 3, 5] = 0
    d, f = oracle.edt_ft(occ)
    g = np.indices(occ.shape)
    assert np.array_equal(d, (g[0] - 4) ** 2 + (g[1] - 3) ** 2 + (g[2] - 5) ** 2)
    assert (f == np.ravel_multi_index((4, 3, 5), occ.shape)).all()
    # half space: background plane x = 0
    occ = np.ones((12, 5, 6), np.uint8)
    occ[0] = 0
    d, _ = oracle.edt_ft(occ)
    assert np.array_equal(d, np.indices(occ.shape)[0] ** 2)
    # no background at all
    d, f = oracle.edt_ft(np.ones((3, 4, 5), np.uint8))
    assert (d == -1).all() and (f == -1).all()
    # ties resolve to the lowest linear index: two background voxels
    occ = np.ones((5, 1, 1), np.uint8)
    occ[0, 0, 0] = occ[4, 0, 0] = 0
    d, f = oracle.edt_ft(occ)
    assert f[2, 0, 0] == 0 and d[2, 0, 0] == 4


def test_lambda_closed_form(oracle):
    # slab of thickness 2 along z: features on both faces, lambda^2 = 3^2
    occ = np.zeros((6, 6, 8), np.uint8)
    occ[:, :, 3:5] = 1
    d, f = oracle.edt_ft(occ)
    lam = oracle.lambda2(occ, f)
    assert (lam[:, :, 3:5] >= 9).all() and (lam[occ == 0] == 0).all()
    assert lam[2, 2, 3] == 9 and lam[2, 2, 4] == 9


def _lambda2_numpy(occ, ft):
    """lambda^2 by array shifts: for each of the 6 face neighbours, the
    squared distance between the two voxels' feature coordinates where both
    are inside with a feature; the maximum over the 6 (0 elsewhere)."""
    nx, ny, nz = occ.shape
    ok = (occ > 0) & (ft >= 0)
    fx, fy, fz = np.unravel_index(np.where(ok, ft, 0), occ.shape)
    cx = np.stack([fx, fy, fz]).astype(np.int64)
    out = np.zeros(occ.shape, np.int64)
    for ax in range(3):
        for sgn in (-1, 1):
            sl_a = [slice(None)] * 3
            sl_b = [slice(None)] * 3
            if sgn > 0:
                sl_a[ax], sl_b[ax] = slice(0, -1), slice(1, None)
            else:
                sl_a[ax], sl_b[ax] = slice(1, None), slice(0, -1)
            a, b = tuple(sl_a), tuple(sl_b)
            both = ok[a] & ok[b]
            d = ((cx[(slice(None),) + a] - cx[(slice(None),) + b]) ** 2).sum(axis=0)
            out[a] = np.maximum(out[a], np.where(both, d, 0))
    return out


def test_lambda_against_brute_force_features(oracle):
    """lambda^2 of the oracle against an independent restatement: features
    from the brute-force nearest-background search (or_edt_brute, O(N * B))
    and the 6-neighbour maximum by array shifts, on random volumes of every
    fill (including a volume with no background, where every feature is
    -1 and lambda^2 = 0)."""
    rng = np.random.default_rng(41)
    for i in range(60):
        shape = tuple(int(s) for s in rng.integers(1, 16, size=3))
        occ = (rng.random(shape) < rng.choice([0.3, 0.6, 0.85, 0.97])).astype(np.uint8)
        if i == 0:
            occ[:] = 1
        db, fb = oracle.edt_brute(occ)
        d, f = oracle.edt_ft(occ)
        assert np.array_equal(f, fb), i
        assert np.array_equal(oracle.lambda2(occ, fb), _lambda2_numpy(occ, fb)), i
    # slab of even width 2k between two background planes: the middle pair's
    # features lie on opposite planes, (2k + 1)^2 apart; every other inside
    # voxel sees features one apart across y or z, or identical ones
    for k in (1, 2, 5):
        occ = np.zeros((2 * k + 2, 5, 6), np.uint8)
        occ[1:2 * k + 1] = 1
        d, f = oracle.edt_ft(occ)
        lam = oracle.lambda2(occ, f)
        assert (lam[k] == (2 * k + 1) ** 2).all() and (lam[k + 1] == (2 * k + 1) ** 2).all()
        rest = np.ones(occ.shape[0], bool)
        rest[[0, k, k + 1, 2 * k + 1]] = False
        assert set(np.unique(lam[rest])) <= {0, 1}


def test_level_thresholds(oracle):
    for l, T in [(1.5, 3), (2.0, 4), (0.0, 0), (-1.0, 0), (3.0000002, 10), (0.55, 1)]:
        assert oracle.level_thresholds([l])[0] == T, l
    # lambda < level  <=>  lam2 < T, checked against float64 sqrt
    for l in np.linspace(0.1, 40, 397).astype(np.float32):
        T = oracle.level_thresholds([l])[0]
        for m in range(max(0, T - 3), T + 3):
            assert (np.sqrt(np.float64(m)) < np.float64(l)) == (m < T)


def _cube_code(cells):
    """26-bit neighbourhood code from a set of (dx,dy,dz) offsets."""
    code, b = 0, 0
    for dx, dy, dz in itertools.product((-1, 0, 1), repeat=3):
        if (dx, dy, dz) == (0, 0, 0):
            continue
        if (dx, dy, dz) in cells:
            code |= 1 << b
        b += 1
    return code


def test_simple_point_known_cases(oracle):
    assert not oracle.is_simple(0)                       # isolated point
    assert not oracle.is_simple((1 << 26) - 1)            # interior point
    assert oracle.is_simple(_cube_code({(1, 0, 0)}))      # end of a segment
    assert not oracle.is_simple(_cube_code({(1, 0, 0), (-1, 0, 0)}))  # inside a curve
    half = {c for c in itertools.product((-1, 0, 1), repeat=3) if c[0] == -1}
    assert oracle.is_simple(_cube_code(half))             # on a flat face
    plane = {c for c in itertools.product((-1, 0, 1), repeat=3) if c[2] == 0 and c != (0, 0, 0)}
    assert not oracle.is_simple(_cube_code(plane))        # inside a 1-voxel sheet


def test_simple_point_three_definitions_agree(oracle, simple_lut_cpu):
    rng = np.random.default_rng(11)
    codes = list(rng.integers(0, 1 << 26, size=6000))
    # sparse and dense neighbourhoods are where the definitions differ most
    for k in (1, 2, 3, 23, 24, 25):
        for _ in range(300):
            bits = rng.choice(26, size=k, replace=False)
            c = int(np.bitwise_or.reduce(1 << bits))
            codes.append(c)
    for c in codes:
        c = int(c)
        a = oracle.is_simple(c)
        assert a == oracle.is_simple_topo(c), hex(c)
        assert a == bool((simple_lut_cpu[c >> 5] >> (c & 31)) & 1), hex(c)


def test_cube_components_match_reference_cycle_table(oracle):
    doc = json.load(open(os.path.join(HERE, "golden", "cube_components.json")))
    tbl = oracle.cube_components()
    assert list(tbl) == doc["counts"]


def test_cube_components_symmetric(oracle):
    """The table is invariant under the 48 cube symmetries, so the unknown
    corner order of the absent isosurface `coordmap` cannot change it."""
    tbl = oracle.cube_components()
    for perm in itertools.permutations(range(3)):
        for flips in itertools.product((0, 1), repeat=3):
            def mapc(i):
                b = [(i >> 2) & 1, (i >> 1) & 1, i & 1]
                b = [b[perm[a]] ^ flips[a] for a in range(3)]
                return (b[0] << 2) | (b[1] << 1) | b[2]
            for m in range(256):
                m2 = 0
                for i in range(8):
                    if (m >> (7 - i)) & 1:
                        m2 |= 1 << (7 - mapc(i))
                assert tbl[m] == tbl[m2]


def _ball(n, c, r):
    g = np.indices((n, n, n))
    return (((g - c) ** 2).sum(0) <= r * r).astype(np.uint8)


def test_surfacer_single_voxel_known_answer(oracle):
    """Hand-derived from the reference tables (include/surfacing.h:97-190):
    voxel (2,2,2) is first met from (1,2,2) through +x, then (2,1,2) through
    +y, then (2,2,1) through +z, then from itself."""
    occ = np.zeros((5, 5, 5), np.uint8)
    occ[2, 2, 2] = 1
    s, q, e = oracle.boundary(occ)
    exp_s = [(1.5, 1.5, 1.5), (1.5, 2.5, 1.5), (1.5, 2.5, 2.5), (1.5, 1.5, 2.5),
             (2.5, 1.5, 1.5), (2.5, 1.5, 2.5), (2.5, 2.5, 1.5), (2.5, 2.5, 2.5)]
    exp_q = [(0, 3, 2, 1), (5, 4, 0, 3), (4, 6, 1, 0), (4, 5, 7, 6), (7, 6, 1, 2), (5, 7, 2, 3)]
    assert np.array_equal(s, np.array(exp_s, np.float32))
    assert np.array_equal(q, np.array(exp_q, np.int32))
    assert (e["V"], e["E"], e["F"], e["T"], e["C"], e["euler"]) == (8, 12, 6, 1, 1, 1)


@pytest.mark.parametrize("case", ["ball", "torus", "shell", "blobs", "c1"])
def test_surfacer_euler_closed_forms(oracle, case):
    from voxel_ma_amd import synth
    if case == "ball":
        occ, chi = _ball(20, 10, 6), 1
    elif case == "torus":
        g = np.indices((48, 48, 48)).astype(float) - 23.5
        occ = ((np.sqrt(g[0] ** 2 + g[1] ** 2) - 14) ** 2 + g[2] ** 2 <= 25).astype(np.uint8)
        chi = 0
    elif case == "shell":
        occ, chi = _ball(24, 12, 9) - _ball(24, 12, 5), 2
    elif case == "blobs":
        occ = np.zeros((20, 20, 20), np.uint8)
        occ[2:6, 2:6, 2:6] = 1
        occ[10:15, 10:15, 3:9] = 1
        occ[3:5, 14:17, 10:12] = 1
        chi = 3
    else:
        occ, chi = synth.torus_c1(), 0
    s, q, e = oracle.boundary(occ)
    assert e["euler"] == chi
    assert e["euler"] == oracle.euler_voxels(occ)
    # sites are exactly the lattice corners whose 2x2x2 block is mixed
    p = np.pad(occ, 1)
    blocks = [p[a:a + occ.shape[0] + 1, b:b + occ.shape[1] + 1, c:c + occ.shape[2] + 1]
              for a in (0, 1) for b in (0, 1) for c in (0, 1)]
    sm = sum(b.astype(int) for b in blocks)
    mixed = np.argwhere((sm > 0) & (sm < 8)).astype(np.float32) - 0.5
    assert len(s) == len(mixed)
    assert np.array_equal(np.unique(s, axis=0), np.unique(mixed, axis=0))
    # quads = faces between voxels of different value (grid outside = empty)
    nf = sum(np.count_nonzero(np.diff(p.astype(int), axis=a)) for a in range(3))
    assert len(q) == nf == e["F"] - e["F_int"]
    assert e["T"] == occ.sum()


def test_voxel_topology_helpers(oracle):
    occ = _ball(16, 8, 5)
    assert oracle.betti(occ) == (1, 0, 0, 1)
    shell = _ball(20, 10, 8) - _ball(20, 10, 4)
    assert oracle.betti(shell) == (1, 0, 1, 2)


def test_thinning_preserves_topology_and_is_a_fixed_point(oracle, simple_lut_cpu):
    from voxel_ma_amd import synth
    occ = synth.torus_c1()
    d, f = oracle.edt_ft(occ)
    lam = oracle.lambda2(occ, f)
    levels = [1.5, 3.0, 6.0]
    lvl = oracle.levels(occ, lam, levels)
    lab, st = oracle.thin(occ, lvl, len(levels))
    b = oracle.betti(occ)
    for k in range(len(levels)):
        sk = ((lab == 1) | (lab > 2 + k)).astype(np.uint8)
        assert oracle.betti(sk) == b
        removed = (lab == 2 + k)
        assert removed.sum() == st.removed[k]
        assert (lvl[removed] <= k).all()
    # no eligible voxel of the final skeleton is simple
    alive = (lab == 1)
    p = np.pad(alive, 1)
    code = np.zeros(occ.shape, np.int64)
    bit = 0
    for dx, dy, dz in itertools.product((-1, 0, 1), repeat=3):
        if (dx, dy, dz) == (0, 0, 0):
            continue
        code |= p[1 + dx:1 + dx + occ.shape[0], 1 + dy:1 + dy + occ.shape[1],
                  1 + dz:1 + dz + occ.shape[2]].astype(np.int64) << bit
        bit += 1
    elig = alive & (lvl <= len(levels) - 1)
    c = code[elig]
    assert not ((simple_lut_cpu[c >> 5] >> (c & 31)) & 1).any()


def test_oracle_threads_do_not_change_results(oracle):
    """The OpenMP loops of the restatement (EDT rows, measure, levels, the
    voxels of one subfield pass) give the sequential results."""
    rng = np.random.default_rng(77)
    occ = (rng.random((30, 26, 34)) < 0.75).astype(np.uint8)
    levels = [1.5, 3.0]
    out = []
    n0 = oracle.set_threads(0)
    for n in (1, 4):
        oracle.set_threads(n)
        d, f = oracle.edt_ft(occ)
        lam = oracle.lambda2(occ, f)
        lvl = oracle.levels(occ, lam, levels)
        lab, st = oracle.thin(occ, lvl, len(levels))
        out.append((d, f, lam, lvl, lab, list(st.removed[:2]), st.tests))
    oracle.set_threads(n0)
    for a, b in zip(*out):
        assert np.array_equal(np.asarray(a), np.asarray(b))


def test_component_counts_match_scipy_label(oracle):
    """The Betti helpers' component counts against scipy.ndimage.label (an
    independent library routine): 26-components of the object, and
    6-components of the background with the outside of the grid counted as
    background (so every component touching the border is the one
    unbounded component); chi of a union of closed cubes is additive over
    disjoint pieces."""
    from scipy import ndimage
    rng = np.random.default_rng(31)
    s26 = np.ones((3, 3, 3), bool)
    s6 = ndimage.generate_binary_structure(3, 1)
    for i in range(40):
        shape = tuple(int(s) for s in rng.integers(1, 24, size=3))
        occ = (rng.random(shape) < rng.choice([0.2, 0.5, 0.7, 0.9])).astype(np.uint8)
        assert oracle.components26(occ) == ndimage.label(occ, structure=s26)[1], i
        bg = np.pad(occ == 0, 1, constant_values=True)
        assert oracle.bg_components6(occ) == ndimage.label(bg, structure=s6)[1], i
    # chi adds over pieces that share no lattice point
    a = (rng.random((9, 9, 9)) < 0.6).astype(np.uint8)
    b = (rng.random((9, 9, 9)) < 0.4).astype(np.uint8)
    both = np.zeros((9, 20, 9), np.uint8)
    both[:, :9] = a
    both[:, 11:] = b
    assert oracle.euler_voxels(both) == oracle.euler_voxels(a) + oracle.euler_voxels(b)


def test_simple_point_definitions_agree_on_all_codes(oracle, simple_lut_cpu):
    """Every one of the 2^26 neighbourhood codes: the BFS definition
    (T26 = 1 and T6 = 1 by per-cell adjacency lists) equals the
    bit-parallel definition that fills the table; every 61st code is also
    checked against the topological definition (b0 of the object, b0 of the
    background and chi of a padded 5x5x5 box unchanged when the centre is
    removed).  The table the GPU builds is compared with this one bit for
    bit in the GPU suite."""
    bad, n_simple, tbad, tchk = oracle.simple_exhaustive(topo_stride=61, topo_offset=17)
    assert bad == 0
    assert tchk == len(range(17, 1 << 26, 61)) and tbad == 0
    # the table holds exactly the simple codes the BFS definition counts
    pop = int(np.unpackbits(simple_lut_cpu.view(np.uint8)).sum())
    assert pop == n_simple

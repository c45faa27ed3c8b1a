"""Voronoi-side steps on plain arrays (include/vc_voro.h, host functions):
inside extraction and close-vertex merging against oracle/voro_ref.py, which
is pinned by hand-derived known answers of the reference's order."""
import os
import sys

import numpy as np

import voxel_ma_amd as V

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import voro_ref as R  # noqa: E402


def test_extract_known_answer_square():
    """One valid quad [0,1,2,3] with its sides stored as (0,1) (1,2) (2,3)
    (3,0): zig-zag gives (3,0,1), (3,1,2); the side (0,3) of the first
    triangle is not the stored pair (3,0), so it is appended (the reference
    dedupes exact pairs), then the diagonal (1,3)."""
    edges = [(0, 1), (1, 2), (2, 3), (3, 0)]
    faces = [[0, 1, 2, 3]]
    le = [1.0, 2.0, 3.0, 4.0]
    lf = [7.0]
    for impl in (R.extract_inside, V.voro_extract_inside):
        vi, eo, em, tr, ff, fm = impl(4, edges, faces, [1, 1, 1, 1], le, lf)
        assert list(vi) == [0, 1, 2, 3]
        assert [tuple(e) for e in eo] == [(0, 1), (1, 2), (2, 3), (3, 0), (0, 3), (1, 3)]
        assert list(em) == [1.0, 2.0, 3.0, 4.0, 7.0, 7.0]
        assert [tuple(t) for t in tr] == [(3, 0, 1), (3, 1, 2)]
        assert list(ff) == [0, 0] and list(fm) == [7.0, 7.0]


def test_extract_known_answer_invalid_vertex():
    """Vertex 1 invalid: edges touching it and the face drop out; the rest is
    renumbered 0->0, 2->1, 3->2."""
    edges = [(0, 1), (1, 2), (2, 3), (0, 3)]
    faces = [[0, 1, 2, 3]]
    for impl in (R.extract_inside, V.voro_extract_inside):
        vi, eo, em, tr, ff, fm = impl(4, edges, faces, [1, 0, 1, 1], [1, 2, 3, 4], [9])
        assert list(vi) == [0, 2, 3]
        assert [tuple(e) for e in eo] == [(1, 2), (0, 2)]
        assert list(em) == [3.0, 4.0] and len(tr) == 0


def _random_complex(rng, nv=40, nf=25):
    faces = []
    for _ in range(nf):
        k = int(rng.integers(3, 7))
        faces.append([int(x) for x in rng.choice(nv, size=k, replace=False)])
    # finalize-like edge list: given edges (some unnormalised) then sides
    edges, seen = [], set()
    for _ in range(10):
        a, b = (int(x) for x in rng.choice(nv, size=2, replace=False))
        if (a, b) not in seen:
            seen.add((a, b))
            edges.append((a, b))
    for f in faces:
        for j in range(len(f)):
            a, b = f[j], f[(j + 1) % len(f)]
            k = (min(a, b), max(a, b))
            if k not in seen:
                seen.add(k)
                edges.append(k)
    return nv, edges, faces


def test_extract_matches_oracle_random():
    rng = np.random.default_rng(3)
    for it in range(60):
        nv, edges, faces = _random_complex(rng)
        valid = (rng.random(nv) < 0.8).astype(np.uint8)
        le = rng.random(len(edges)).astype(np.float32)
        lf = rng.random(len(faces)).astype(np.float32)
        got = V.voro_extract_inside(nv, edges, faces, valid, le, lf)
        ref = R.extract_inside(nv, edges, faces, valid, le, lf)
        assert list(got[0]) == list(ref[0])
        assert [tuple(e) for e in got[1]] == [tuple(e) for e in ref[1]]
        assert np.array_equal(got[2], ref[2])
        assert [tuple(t) for t in got[3]] == [tuple(t) for t in ref[3]]
        assert list(got[4]) == list(ref[4]) and np.array_equal(got[5], ref[5])


def test_merge_known_answer_path():
    """Path 0 - 1 - 2 at x = 0, 0.5, 1.2, eps 1: seed 0 absorbs 1 (0.5 from
    the seed); 2 is 1.2 from the seed 0, so it seeds vertex 1.  Edge (0,1)
    collapses, (1,2) becomes (0,1)."""
    vts = [(0, 0, 0), (0.5, 0, 0), (1.2, 0, 0)]
    edges = [(0, 1), (1, 2)]
    for impl in (R.merge_close, V.voro_merge_close):
        o2n, nvts, faces, rem, ne, keep = impl(vts, edges, [], [1, 1, 1], 1.0, False)
        assert list(o2n) == [0, 0, 1] and list(nvts) == [0, 2]
        assert [tuple(e) for e in ne] == [(0, 1)] and list(keep) == [0, 1]


def test_merge_known_answer_triangle_to_edge():
    """Triangle 0,1,2 with 0 and 1 within eps: it degenerates to the edge
    (0,1) of the new numbering and is removed."""
    vts = [(0, 0, 0), (0.1, 0, 0), (0, 3, 0)]
    edges = [(0, 1), (1, 2), (0, 2)]
    for impl in (R.merge_close, V.voro_merge_close):
        o2n, nvts, faces, rem, ne, keep = impl(vts, edges, [[0, 1, 2]], [1, 1, 1], 0.5, False)
        assert list(o2n) == [0, 0, 1] and faces == [] and list(rem) == [1]
        assert [tuple(e) for e in ne] == [(0, 1)] and list(keep) == [0, 1, 1]


def test_merge_matches_oracle_random():
    rng = np.random.default_rng(5)
    for it in range(60):
        nv, edges, faces = _random_complex(rng, nv=30, nf=15)
        vts = rng.random((nv, 3)).astype(np.float32) * 3
        valid = (rng.random(nv) < 0.85).astype(np.uint8)
        eps = float(rng.uniform(0.2, 1.0))
        for only in (False, True):
            got = V.voro_merge_close(vts, edges, faces, valid, eps, only)
            ref = R.merge_close(vts, edges, faces, valid, eps, only)
            assert list(got[0]) == list(ref[0])
            assert list(got[1]) == list(ref[1])
            assert [list(f) for f in got[2]] == ref[2]
            assert list(got[3]) == ref[3]
            assert [tuple(e) for e in got[4]] == ref[4]
            assert list(got[5]) == ref[5]


def test_part_inside_known_answer():
    """Pentagon [0..4] with vertex 4 invalid: zig-zag gives (4,0,1), (4,1,3),
    (3,1,2); only (3,1,2) has three valid corners -> (3,1,2) in the new
    numbering 0..3, sides (1,3), (1,2), (2,3) sorted.  The invalid face still
    contributes its valid triangle (the reference tests triangles, not
    faces)."""
    for impl in (R.part_inside, V.voro_part_inside):
        vi, eo, tr = impl(5, [[0, 1, 2, 3, 4]], [1, 1, 1, 1, 0])
        assert list(vi) == [0, 1, 2, 3]
        assert [tuple(int(x) for x in t) for t in tr] == [(3, 1, 2)]
        assert [tuple(int(x) for x in e) for e in eo] == [(1, 2), (1, 3), (2, 3)]


def test_part_inside_matches_oracle_random():
    rng = np.random.default_rng(11)
    for it in range(60):
        nv, edges, faces = _random_complex(rng)
        valid = (rng.random(nv) < 0.8).astype(np.uint8)
        got = V.voro_part_inside(nv, faces, valid)
        ref = R.part_inside(nv, faces, valid)
        assert list(got[0]) == ref[0]
        assert [tuple(int(x) for x in e) for e in got[1]] == ref[1]
        assert [tuple(int(x) for x in t) for t in got[2]] == ref[2]


def test_radii_oracle_known_answer():
    """Vertex 0 lies in faces 0 and 1; face 1 comes last, so its first site
    (site 2 at (0, 0, 4)) decides: |(3, 4, 0) - (0, 0, 4)| = sqrt(41).
    Vertex 1 only in face 0 (first site 0 at the origin): |(3, 4, 12)| = 13.
    Vertex 3 is in no face: 0."""
    vts = [(3, 4, 0), (3, 4, 12), (0, 0, 0), (9, 9, 9)]
    faces = [[0, 1, 2], [2, 0, 1]]
    fs = [(0, 1), (2, 0)]
    sites = [(0, 0, 0), (1, 1, 1), (0, 0, 4)]
    r = R.radii(vts, [[0, 1, 2], [0, 2]], fs, sites)
    assert r[0] == np.float32(np.sqrt(np.float32(41))) and r[1] == 13.0
    assert r[2] == 4.0 and r[3] == 0.0
    r2 = R.radii(vts, faces, fs, sites)
    assert r2[1] == np.float32(np.sqrt(np.float32(9 + 16 + 64)))


# ---- discrete Voronoi diagram of the feature transform (oracle pinning)

def _brute_ft(occ):
    """Nearest background voxel of every voxel by exhaustive search: squared
    distance to every background voxel, first (lowest linear index) minimum;
    -1 when there is no background."""
    shape = occ.shape
    coords = np.indices(shape).reshape(3, -1).T
    bg = np.flatnonzero(occ.reshape(-1) == 0)
    if not len(bg):
        return np.full(shape, -1, np.int64)
    d = ((coords[:, None, :] - coords[bg][None, :, :]) ** 2).sum(-1)
    return bg[np.argmin(d, axis=1)].reshape(shape)


def _loop_voronoi(occ, ft):
    """Per-voxel loop restatement: faces as corner coordinate tuples, then
    vertices numbered by sorting the distinct corners."""
    nx, ny, nz = occ.shape
    faces, fsites = [], []
    for x in range(nx):
        for y in range(ny):
            for z in range(nz):
                if not occ[x, y, z]:
                    continue
                for a, (dx, dy, dz) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
                    q = (x + dx, y + dy, z + dz)
                    if q[a] >= occ.shape[a] or not occ[q] or ft[q] == ft[x, y, z]:
                        continue
                    b, c = (a + 1) % 3, (a + 2) % 3
                    quad = []
                    for db, dc in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        v = [x - 0.5, y - 0.5, z - 0.5]
                        v[a] += 1
                        v[b] += db
                        v[c] += dc
                        quad.append(tuple(v))
                    faces.append(quad)
                    fsites.append((int(ft[x, y, z]), int(ft[q])))
    corners = sorted({v for f in faces for v in f})
    vid = {v: i for i, v in enumerate(corners)}
    used = sorted({s for p in fsites for s in p})
    sid = {s: i for i, s in enumerate(used)}
    sites = [(s // (ny * nz), (s // nz) % ny, s % nz) for s in used]
    return (np.array(corners, np.float32).reshape(-1, 3),
            np.array([[vid[v] for v in f] for f in faces], np.int32).reshape(-1, 4),
            np.array([[sid[a], sid[b]] for a, b in fsites], np.int32).reshape(-1, 2),
            np.array(sites, np.float32).reshape(-1, 3))


def test_voronoi_from_ft_known_answer():
    """Two inside voxels (1,1,1), (1,1,2) in a 3 x 3 x 4 grid: each has five
    background face neighbours at distance 1, the lowest-index ones are
    (0,1,1) (index 5) and (0,1,2) (index 6), so the pair is split by one face
    at z = 1.5 with corners (x, y) = (0.5, 0.5), (1.5, 0.5), (1.5, 1.5),
    (0.5, 1.5); lattice points (1,1,2), (1,2,2), (2,1,2), (2,2,2) number
    0..3, so the quad is [0, 2, 3, 1]."""
    occ = np.zeros((3, 3, 4), np.uint8)
    occ[1, 1, 1] = occ[1, 1, 2] = 1
    ft = _brute_ft(occ)
    assert ft[1, 1, 1] == 5 and ft[1, 1, 2] == 6
    vts, faces, fs, sites = R.voronoi_from_ft(occ, ft)
    assert vts.tolist() == [[0.5, 0.5, 1.5], [0.5, 1.5, 1.5], [1.5, 0.5, 1.5],
                            [1.5, 1.5, 1.5]]
    assert faces.tolist() == [[0, 2, 3, 1]]
    assert fs.tolist() == [[0, 1]]
    assert sites.tolist() == [[0, 1, 1], [0, 1, 2]]


def test_voronoi_from_ft_matches_brute_force():
    """The oracle against a per-voxel loop on a brute-force feature
    transform, on small random volumes (both empty-face and dense cases)."""
    rng = np.random.default_rng(12)
    for i in range(12):
        shape = tuple(int(s) for s in rng.integers(1, 7, size=3))
        occ = (rng.random(shape) < rng.choice([0.3, 0.7, 0.9, 1.0])).astype(np.uint8)
        ft = _brute_ft(occ)
        got = R.voronoi_from_ft(occ, ft)
        want = _loop_voronoi(occ, ft)
        for g, w in zip(got, want):
            assert g.shape == w.shape and np.array_equal(g, w), (i, shape)
        # every face separates two different sites; every vertex is used
        if len(got[1]):
            assert np.all(got[2][:, 0] != got[2][:, 1])
            assert set(np.unique(got[1])) == set(range(len(got[0])))

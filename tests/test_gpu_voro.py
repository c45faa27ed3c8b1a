"""GPU parity of the Voronoi-side kernels (include/vc_voro.h): inside
tagging of points and the lambda measures of faces, edges and vertices,
bit-exact against oracle/voro_ref.py."""
import os
import sys

import numpy as np
import pytest

import voxel_ma_amd as V
from conftest import ROOT, say

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import voro_ref as R  # noqa: E402

pytestmark = pytest.mark.gpu


def test_tag_points_matches_oracle(vc):
    rng = np.random.default_rng(1)
    occ = (rng.random((13, 9, 17)) < 0.5).astype(np.uint8)
    vol = vc.upload(occ)
    pts = rng.uniform(-2, 19, size=(5000, 3)).astype(np.float32)
    # exact half-integers (round away from zero), boundary and odd values
    half = (rng.integers(-2, 18, size=(500, 3)) + 0.5).astype(np.float32)
    special = np.array([[-0.5, 0, 0], [-0.49999997, 0, 0], [12.5, 8.4, 16.4],
                        [np.nan, 1, 1], [np.inf, 1, 1], [3e9, 1, 1], [0, 0, 0]], np.float32)
    pts = np.concatenate([pts, half, special])
    say("tag points", len(pts))
    got = V.voro_tag_points(vol, pts)
    assert np.array_equal(got, R.tag_points(occ, pts))
    vol.release()


def _complex(rng, nv=300, nf=200, nsites=80):
    faces = []
    for _ in range(nf):
        k = int(rng.integers(3, 7))
        faces.append([int(x) for x in rng.choice(nv, size=k, replace=False)])
    edges, seen = [], set()
    for f in faces:
        for j in range(len(f)):
            a, b = f[j], f[(j + 1) % len(f)]
            k = (min(a, b), max(a, b))
            if k not in seen:
                seen.add(k)
                edges.append(k)
    fs = [tuple(int(x) for x in rng.choice(nsites, size=2, replace=False)) for _ in range(nf)]
    sites = (rng.random((nsites, 3)) * 40 - 5).astype(np.float32)
    return nv, edges, faces, fs, sites


def test_measures_match_oracle(vc):
    rng = np.random.default_rng(2)
    for it in range(5):
        nv, edges, faces, fs, sites = _complex(rng)
        fvalid = (rng.random(len(faces)) < 0.7).astype(np.uint8)
        lv_r, le_r, lf_r, ve, ef = R.measures(nv, edges, faces, fs, sites, fvalid)
        ve_off = np.cumsum([0] + [len(x) for x in ve])
        ef_off = np.cumsum([0] + [len(x) for x in ef])
        say("measures", it, nv, len(edges), len(faces))
        lv, le, lf = V.voro_measures(vc, nv, edges, fs, sites, fvalid, ve_off,
                                     [i for x in ve for i in x], ef_off,
                                     [i for x in ef for i in x])
        assert np.array_equal(lf, lf_r)
        assert np.array_equal(le, le_r)
        assert np.array_equal(lv, lv_r)


def test_radii_match_oracle(vc):
    rng = np.random.default_rng(4)
    for it in range(5):
        nv, edges, faces, fs, sites = _complex(rng)
        vts = (rng.random((nv, 3)) * 40 - 5).astype(np.float32)
        say("radii", it, nv, len(faces))
        got = V.voro_radii(vc, vts, faces, fs, sites)
        assert np.array_equal(got, R.radii(vts, faces, fs, sites))
    # known answer: the last face holding a vertex decides; no face -> 0
    got = V.voro_radii(vc, [(3, 4, 0), (3, 4, 12), (0, 0, 0), (9, 9, 9)],
                       [[0, 1, 2], [0, 2]], [(0, 1), (2, 0)],
                       [(0, 0, 0), (1, 1, 1), (0, 0, 4)])
    assert got[0] == np.float32(np.sqrt(np.float32(41))) and got[1] == 13.0
    assert got[2] == 4.0 and got[3] == 0.0
    assert len(V.voro_radii(vc, [], [], [], [(0, 0, 0)])) == 0


def test_voronoi_from_ft_matches_oracle(vc, oracle):
    """vc_voro_from_ft against the oracle's discrete Voronoi diagram of the
    oracle feature transform: vertices, faces, face sites and sites
    bit-exact, on random volumes (odd shapes, sparse and dense), a torus and
    a full block (no background: no face)."""
    from voxel_ma_amd import synth
    rng = np.random.default_rng(31)
    cases = []
    for i in range(8):
        shape = tuple(int(s) for s in rng.integers(1, 24, size=3))
        cases.append((rng.random(shape) < rng.choice([0.3, 0.6, 0.9])).astype(np.uint8))
    cases.append(synth.torus_c1())
    cases.append(np.ones((5, 6, 7), np.uint8))
    cases.append(np.zeros((4, 4, 4), np.uint8))
    for i, occ in enumerate(cases):
        say("voronoi from ft", i, occ.shape)
        vol = vc.upload(occ)
        vol.edt()
        got = V.voro_from_ft(vol)
        _, ft = oracle.edt_ft(occ)
        want = R.voronoi_from_ft(occ, ft)
        for name, g, w in zip(("vts", "faces", "face_sites", "sites"), got, want):
            assert g.shape == w.shape and np.array_equal(g, w), (i, name)
        vol.release()


def test_voronoi_from_ft_errors_and_measures(vc):
    """Before vc_edt the call fails with VC_ERR_STATE; short capacities give
    VC_ERR_CAPACITY with the counts; the two sites of a face are different
    voxel centres, so its lambda |s1 - s2| is at least 1."""
    import ctypes
    rng = np.random.default_rng(32)
    occ = (rng.random((20, 18, 22)) < 0.7).astype(np.uint8)
    vol = vc.upload(occ)
    with pytest.raises(V.VcError):
        V.voro_from_ft(vol)
    vol.edt()
    vts, faces, fs, sites = V.voro_from_ft(vol)
    assert len(faces) > 0
    L = V.library()
    nv, nf, ns = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    bv = np.zeros((len(vts), 3), np.float32)
    bf = np.zeros((len(faces), 4), np.int32)
    bs = np.zeros((len(faces), 2), np.int32)
    bst = np.zeros((len(sites), 3), np.float32)
    rc = L.vc_voro_from_ft(vc.handle, vol.handle, bv.ctypes.data, len(vts), ctypes.byref(nv),
                           bf.ctypes.data, bs.ctypes.data, len(faces) - 1, ctypes.byref(nf),
                           bst.ctypes.data, len(sites), ctypes.byref(ns))
    assert rc == V.VC_ERR_CAPACITY
    assert (nv.value, nf.value, ns.value) == (len(vts), len(faces), len(sites))
    lam_f = np.linalg.norm(sites[fs[:, 0]] - sites[fs[:, 1]], axis=1)
    assert np.all(lam_f >= 1.0)
    vol.release()


def test_extract_inside_device_matches_host(vc):
    """vc_voro_extract_inside_device against the host extraction (pinned to
    the oracle in tests/test_voro.py): the hand-derived square (a stored
    edge (3,0) does not hide the side (0,3)), random polygon complexes with
    stored edges in both orientations and random validity, and the discrete
    Voronoi diagram of a ball."""
    edges = [(0, 1), (1, 2), (2, 3), (3, 0)]
    got = V.voro_extract_inside_device(vc, 4, edges, [[0, 1, 2, 3]], [1, 1, 1, 1],
                                       [1.0, 2.0, 3.0, 4.0], [7.0])
    assert [tuple(e) for e in got[1]] == [(0, 1), (1, 2), (2, 3), (3, 0), (0, 3), (1, 3)]
    assert list(got[2]) == [1.0, 2.0, 3.0, 4.0, 7.0, 7.0]
    assert [tuple(t) for t in got[3]] == [(3, 0, 1), (3, 1, 2)]
    rng = np.random.default_rng(41)
    cases = []
    for it in range(6):
        nv, edges, faces, fs, sites = _complex(rng)
        edges = [(b, a) if rng.random() < 0.3 else (a, b) for a, b in edges]
        edges = [e for e in edges if rng.random() < 0.6]
        valid = (rng.random(nv) < rng.choice([0.5, 0.9, 1.0])).astype(np.uint8)
        cases.append((nv, edges, faces, valid, rng.random(len(edges)).astype(np.float32),
                      rng.random(len(faces)).astype(np.float32)))
    n = 24
    g = np.indices((n, n, n)) - (n - 1) / 2.0
    occ = ((g ** 2).sum(0) <= 9.5 ** 2).astype(np.uint8)
    vol = vc.upload(occ)
    vol.edt()
    vts, faces, fs, sites = V.voro_from_ft(vol)
    valid = V.voro_tag_points(vol, vts)
    vol.release()
    lam = np.linalg.norm(sites[fs[:, 0]] - sites[fs[:, 1]], axis=1).astype(np.float32)
    cases.append((len(vts), [], faces, valid, np.zeros(0, np.float32), lam))
    for i, (nv, edges, faces, valid, le, lf) in enumerate(cases):
        say("extract inside device", i, nv, len(edges), len(faces))
        want = V.voro_extract_inside(nv, edges, faces, valid, le, lf)
        got = V.voro_extract_inside_device(vc, nv, edges, faces, valid, le, lf)
        for k, (a, b) in enumerate(zip(got, want)):
            assert np.array_equal(np.asarray(a), np.asarray(b)), (i, k)


def test_merge_close_device_matches_host(vc):
    """vc_voro_merge_close_device against the host merge: random complexes
    whose adjacent vertices are apart (the device path, with polygons that
    repeat a vertex: deduplicated, or turned into an edge when two distinct
    ids remain), with and without only_inside; the discrete Voronoi diagram
    of a ball; and complexes with coincident neighbours (the host path)."""
    rng = np.random.default_rng(52)
    cases = []
    for it in range(6):
        nv, edges, faces, fs, sites = _complex(rng)
        vts = (rng.random((nv, 3)) * 100).astype(np.float32)
        faces = faces + [[0, 1, 0, 2], [3, 4, 3], [5, 6, 5, 6, 7]]
        valid = (rng.random(nv) < 0.8).astype(np.uint8)
        cases.append((vts, edges, faces, valid, 1e-3, it % 2 == 0, True))
    n = 20
    g = np.indices((n, n, n)) - (n - 1) / 2.0
    occ = ((g ** 2).sum(0) <= 8.0 ** 2).astype(np.uint8)
    vol = vc.upload(occ)
    vol.edt()
    vts, faces, fs, sites = V.voro_from_ft(vol)
    valid = V.voro_tag_points(vol, vts)
    vol.release()
    cases.append((vts, [], [list(f) for f in faces], valid, 1e-6, True, True))
    # coincident neighbours: vertex 1 on top of vertex 0, 5 near 4
    vts = (rng.random((12, 3)) * 10).astype(np.float32)
    vts[1] = vts[0]
    vts[5] = vts[4] + np.float32(1e-5)
    edges = [(0, 1), (1, 2), (2, 3), (4, 5), (5, 6), (0, 7), (8, 9), (10, 11)]
    faces = [[0, 1, 2], [4, 5, 6, 3], [7, 8, 9, 10]]
    cases.append((vts, edges, faces, np.ones(12, np.uint8), 1e-3, False, False))
    for i, (vts, edges, faces, valid, eps, oi, dev_expected) in enumerate(cases):
        say("merge close device", i, len(vts), len(edges), len(faces), oi)
        want = V.voro_merge_close(vts, edges, faces, valid, eps, oi)
        got = V.voro_merge_close_device(vc, vts, edges, faces, valid, eps, oi)
        assert got[6] == dev_expected, i
        for k in range(6):
            assert got[k] == want[k], (i, k)

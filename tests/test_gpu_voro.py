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

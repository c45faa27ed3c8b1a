"""GPU: the FIFO simple-pair collapse on the device (vc_cc_prune_device)
against the host collapse vc_cc_prune, which tests/test_ccthin.py pins to
the oracle (src/ccthin.cpp:201-771): removed flags and final edge counts
equal on hand-derived cases, random polygon complexes, large grids and the
complexes of the voxel core."""
import numpy as np
import pytest

import voxel_ma_amd as V
from conftest import say
from test_ccthin import _grid, _random_complex

pytestmark = pytest.mark.gpu


def _both(vc, nv, e, foff, fe, em, fm, f_t, l_t):
    rh = V.cc_prune(nv, e, foff, fe, em, fm, f_t, l_t)
    rd = V.cc_prune_device(vc, nv, e, foff, fe, em, fm, f_t, l_t)
    assert rh[0] == 0 and rd[0] == 0, (rh[0], rd[0])
    for a, b in zip(rh[1:], rd[1:5]):
        assert np.array_equal(a, b)
    return rd


def test_device_known_answers(vc):
    """The hand-derived FIFO traces of the triangle and of the thresholded
    square (see tests/test_ccthin.py) through the device collapse."""
    e, foff, fe = V.cc_finalize(3, [], [[0, 1, 2]])
    rc, rv, re, rf, ref, gens = V.cc_prune_device(vc, 3, e, foff, fe, np.zeros(3),
                                                  np.zeros(1), 1.0, 1.0)
    assert rc == 0 and list(rv) == [1, 1, 0] and list(re) == [1, 1, 1] and list(rf) == [1]
    # queue generations: [FE(f0,e0), FE(f0,e1), FE(f0,e2)]; the first pop
    # pushes EV(e2,v0) and EV(e1,v1); popping EV(e2,v0) leaves vertex 2 with
    # one edge and pushes EV(e1,v2), which is skipped (e1 already gone)
    assert gens == 3
    e, foff, fe = V.cc_finalize(4, [], [[0, 1, 2], [0, 2, 3]])
    rc, rv, re, rf, ref, gens = V.cc_prune_device(vc, 4, e, foff, fe, np.zeros(5),
                                                  [0.5, 2.0], 1.0, 1.0)
    assert rc == 0
    assert list(rv) == [0, 1, 0, 0] and list(re) == [1, 1, 0, 0, 0]
    assert list(rf) == [1, 0] and list(ref) == [0, 0, 1, 1, 1]


@pytest.mark.parametrize("seed", range(3))
def test_device_matches_host_on_random_complexes(vc, seed):
    rng = np.random.default_rng(500 + seed)
    for i in range(60):
        nv, extra, faces = _random_complex(rng)
        e, foff, fe = V.cc_finalize(nv, extra, faces)
        em = rng.random(len(e)).astype(np.float32)
        fm = rng.random(len(faces)).astype(np.float32)
        f_t, l_t = float(rng.uniform(0, 1.2)), float(rng.uniform(0, 1.2))
        _both(vc, nv, e, foff, fe, em, fm, f_t, l_t)


def test_device_matches_host_on_large_grids(vc):
    """Triangulated grids up to 120 x 90 (21,600 faces) with random measures
    and quantised measures (many equal values, long collapse chains)."""
    rng = np.random.default_rng(7)
    for m, n, q in ((40, 30, 0), (120, 90, 0), (120, 90, 4), (64, 64, 2)):
        nv, faces = _grid(m, n)
        e, foff, fe = V.cc_finalize(nv, [], faces)
        em = rng.random(len(e)).astype(np.float32)
        fm = rng.random(len(faces)).astype(np.float32)
        if q:
            em = np.floor(em * q) / q
            fm = np.floor(fm * q) / q
        for f_t, l_t in ((1.1, 1.1), (0.5, 0.7), (0.3, 0.0)):
            say("device collapse grid", m, n, q, f_t, l_t)
            _both(vc, nv, e, foff, fe, em, fm, f_t, l_t)


def test_device_matches_host_on_voxel_core(vc):
    """The inside part of the discrete Voronoi diagram of a ball and of a
    torus (triangulated, with the lambda measures of its faces and edges),
    pruned at several thresholds."""
    from voxel_ma_amd import synth
    n = 28
    g = np.indices((n, n, n)) - (n - 1) / 2.0
    ball = ((g ** 2).sum(0) <= 11.0 ** 2).astype(np.uint8)
    for name, occ in (("ball", ball), ("torus", synth.torus_c1())):
        vol = vc.upload(occ)
        vol.edt()
        vts, faces, fs, sites = V.voro_from_ft(vol)
        valid = V.voro_tag_points(vol, vts)
        vol.release()
        lam_f = np.linalg.norm(sites[fs[:, 0]] - sites[fs[:, 1]], axis=1).astype(np.float32)
        vi, eo, em, tr, ff, fm = V.voro_extract_inside(len(vts), [], faces, valid,
                                                       np.zeros(0, np.float32), lam_f)
        tris = np.asarray(tr, np.int64).reshape(-1, 3)
        e, foff, fe = V.cc_finalize(len(vi), np.asarray(eo, np.int64).reshape(-1, 2),
                                    [list(t) for t in tris])
        em_all = np.zeros(len(e), np.float32)
        em_all[:len(em)] = em
        size = float(max(np.ptp(np.argwhere(occ), axis=0)))
        for t in (0.0, 0.04, 0.2, 1.0):
            say("device collapse voxel core", name, len(tris), t)
            thr = np.float32(size * t)
            rd = _both(vc, len(vi), e, foff, fe, em_all, np.asarray(fm, np.float32), thr, thr)
            # a face's lambda is at least 1, so below a threshold of 1 nothing
            # is queued; above it the collapse runs
            assert (rd[5] >= 1) == (thr > 1.0)


def test_device_rejects_degenerate_complexes(vc):
    """An edge with equal ends or a face listing a side twice is refused
    (VC_ERR_INVALID); vc_cc_prune handles those sequentially."""
    e = np.array([[0, 0], [0, 1]], np.int32)
    foff = np.array([0], np.int64)
    rc = V.cc_prune_device(vc, 2, e, foff, np.zeros(1, np.int32), np.zeros(2), np.zeros(0),
                           1.0, 1.0)[0]
    assert rc == V.VC_ERR_INVALID
    e = np.array([[0, 1], [1, 2], [0, 2]], np.int32)
    foff = np.array([0, 4], np.int64)
    fe = np.array([0, 1, 2, 1], np.int32)
    rc = V.cc_prune_device(vc, 3, e, foff, fe, np.zeros(3), np.zeros(1), 1.0, 1.0)[0]
    assert rc == V.VC_ERR_INVALID


def test_finalize_device_matches_host(vc):
    """vc_cc_finalize_device against the host finalize: the reference's
    hand-built complex (tests/test_voro_merge.cpp:10-47: duplicate given
    edges collapse, face sides reuse given ids), random complexes with given
    edges in both orientations (an edge stored (b, a) does not absorb the
    normalised side (a, b)), and a voxel-core complex."""
    edges = [(0, 1), (1, 2), (2, 3), (0, 3), (2, 4), (1, 4), (1, 5), (4, 5), (5, 7), (6, 7),
             (4, 6), (0, 3), (3, 8), (0, 8), (0, 8), (8, 9), (0, 9)]
    faces = [[0, 1, 2, 3], [1, 2, 4], [1, 4, 5], [4, 5, 7, 6], [0, 3, 8], [0, 8, 9]]
    cases = [(10, edges, faces)]
    rng = np.random.default_rng(61)
    for _ in range(20):
        nv, extra, fs = _random_complex(rng)
        extra = extra + [(b, a) for a, b in extra[:2]]
        cases.append((nv, extra, fs))
    nv, faces = _grid(150, 120)
    cases.append((nv, [(1, 0), (0, 151), (3, 2)], faces))
    for i, (nv, e_in, fs) in enumerate(cases):
        say("finalize device", i, nv, len(e_in), len(fs))
        h = V.cc_finalize(nv, e_in, fs)
        d = V.cc_finalize(nv, e_in, fs, ctx=vc)
        for a, b in zip(h, d):
            assert np.array_equal(a, b), i

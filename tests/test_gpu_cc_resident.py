"""GPU: the cell complex resident on the device (include/vc_cc.h) -- finalize,
repeated prune calls on one collapse state, the remaining complex, the
kept-id lists and the topology checks -- against the oracle
(or_cc_finalize / or_cc_prune / or_simple_triangulate, pinned in
tests/test_ccthin.py) and against scipy.sparse.csgraph for the component
counts.  The remaining complex is composed from the oracle's removal flags
as src/ccthin.cpp:467-512 composes it (isolated kept edges, zig-zag
triangles of kept faces, util::compactify)."""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

import voxel_ma_amd as V
from conftest import say
from test_ccthin import _grid, _random_complex

pytestmark = pytest.mark.gpu

OCTA = [[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]]


def _ncomp(n, pairs):
    """connected components of n nodes joined by the given pairs (scipy)"""
    if n == 0:
        return 0
    p = np.asarray(pairs, np.int64).reshape(-1, 2)
    m = coo_matrix((np.ones(len(p)), (p[:, 0], p[:, 1])), shape=(n, n))
    return int(connected_components(m, directed=False)[0])


def _oracle_remaining(oracle, e, foff, fv, rv, re, rf, ref):
    keep_v = np.flatnonzero(rv == 0)
    new = {int(v): i for i, v in enumerate(keep_v)}
    keep_e = np.flatnonzero((re == 0) & (ref == 0))
    edges = [(new.get(int(a), 0), new.get(int(b), 0)) for a, b in e[keep_e]]
    tris, ff = [], []
    for f in np.flatnonzero(rf == 0):
        poly = fv[foff[f]:foff[f + 1]]
        if len(poly) < 3:
            continue
        for t in oracle.simple_triangulate(poly):
            tris.append([new.get(int(x), 0) for x in t])
            ff.append(int(f))
    return (keep_v, np.array(edges, np.int32).reshape(-1, 2),
            np.array(tris, np.int32).reshape(-1, 3), np.array(ff, np.int32),
            keep_e, np.flatnonzero(rf == 0))


def _euler_of(nv, edges, tris):
    """util::computeEulerChar(vts, edges, tris): pairs as given plus the
    normalised triangle sides, components by scipy"""
    pairs = {(int(a), int(b)) for a, b in edges}
    for t in tris:
        for k in range(3):
            a, b = int(t[k]), int(t[(k + 1) % 3])
            pairs.add((min(a, b), max(a, b)))
    return {"V": nv, "E": len(pairs), "F": len(tris), "C": _ncomp(nv, list(pairs)),
            "euler": nv - len(pairs) + len(tris)}


def _faces_csr(faces):
    foff = np.zeros(len(faces) + 1, np.int64)
    foff[1:] = np.cumsum([len(f) for f in faces])
    fv = np.concatenate([np.asarray(f, np.int32) for f in faces]) if faces else \
        np.zeros(0, np.int32)
    return foff, fv


def _cases(seed, n):
    rng = np.random.default_rng(seed)
    out = [(6, [], OCTA), (4, [], [[0, 1, 2], [0, 2, 3]])]
    for _ in range(n):
        out.append(_random_complex(rng))
    nv, faces = _grid(60, 45)
    out.append((nv, [(1, 0), (0, 61), (3, 2)], faces))
    return rng, out


def test_resident_finalize_and_topology(vc, oracle):
    rng, cases = _cases(801, 40)
    for i, (nv, extra, faces) in enumerate(cases):
        cc = V.CellComplex(vc, nv, extra, faces)
        e, fe = cc.finalized()
        eo, fo, feo = oracle.cc_finalize(nv, extra, faces)
        assert np.array_equal(e, eo) and np.array_equal(fe, feo), i
        t = cc.topology()
        assert (t["V"], t["E"], t["F"]) == (nv, len(eo), len(faces))
        assert t["euler"] == nv - len(eo) + len(faces)
        assert t["C"] == _ncomp(nv, eo), i
        # faces joined through shared edge ids
        pairs = []
        first = {}
        for f in range(len(faces)):
            for x in feo[fo[f]:fo[f + 1]]:
                if int(x) in first:
                    pairs.append((first[int(x)], f))
                else:
                    first[int(x)] = f
        assert t["face_components"] == _ncomp(len(faces), pairs), i
        # compute_conn_cmpnts(subset, 1): vertices of the chosen faces
        # through the chosen faces' sides
        sub = rng.random(len(faces)) < 0.5
        foff, fv = _faces_csr(faces)
        vs = sorted({int(v) for f in np.flatnonzero(sub) for v in fv[foff[f]:foff[f + 1]]})
        idx = {v: k for k, v in enumerate(vs)}
        sp = [(idx[int(fv[foff[f] + j]) ], idx[int(fv[foff[f] + (j + 1) % (foff[f + 1] - foff[f])])])
              for f in np.flatnonzero(sub) for j in range(foff[f + 1] - foff[f])]
        assert cc.subset_components(sub) == _ncomp(len(vs), sp), i
        cc.release()


def test_resident_prune_many_thresholds(vc, oracle):
    """One collapse state, several prune calls (each from the full complex),
    every result against the oracle's prune of the same threshold; the
    remaining complex, the kept-id lists and the Euler characteristic of
    what remains after each."""
    rng, cases = _cases(802, 40)
    for i, (nv, extra, faces) in enumerate(cases):
        eo, fo, feo = oracle.cc_finalize(nv, extra, faces)
        foff, fv = _faces_csr(faces)
        em = rng.random(len(eo)).astype(np.float32)
        fm = rng.random(len(faces)).astype(np.float32)
        if i % 3 == 0:
            em = np.floor(em * 3) / 3
            fm = np.floor(fm * 3) / 3
        cc = V.CellComplex(vc, nv, extra, faces)
        th = cc.thinning()
        th.assign(em, fm)
        for f_t, l_t in ((1.1, 1.1), (0.4, 0.7), (0.0, 0.0), (0.7, 0.2)):
            rc, gens = th.prune(f_t, l_t)
            ro = oracle.cc_prune(nv, eo, fo, feo, em, fm, f_t, l_t)
            assert rc == ro[0] == 0, (i, rc)
            got = th.removed()
            for a, b in zip(ro[1:], got):
                assert np.array_equal(a, b), (i, f_t, l_t)
            vi, ed, tr, ff = th.remaining()
            ov, oe, ot, off, okeep_e, okeep_f = _oracle_remaining(oracle, eo, foff, fv, *ro[1:])
            assert np.array_equal(vi, ov) and np.array_equal(ed, oe), (i, f_t)
            assert np.array_equal(tr, ot) and np.array_equal(ff, off), (i, f_t)
            rv_ids, re_ids, rf_ids = th.remained()
            assert np.array_equal(rv_ids, ov) and np.array_equal(re_ids, okeep_e)
            assert np.array_equal(rf_ids, okeep_f)
            want = _euler_of(len(ov), oe, ot)
            got_e = th.remaining_euler()
            assert {k: got_e[k] for k in want} == want, (i, f_t)
            got_a = V.euler_char(vc, len(ov), oe, ot)
            assert {k: got_a[k] for k in want} == want, (i, f_t)
        th.release()
        cc.release()


def test_closed_components(vc, oracle):
    """nClosedComponents: known answers (a closed octahedron is one closed
    pocket; two are two; a disk and an annulus collapse away; an octahedron
    beside a disk is one), then random complexes against the oracle's
    prune at measures 0 / threshold 1 and scipy on the remaining triangles
    joined through shared sides."""
    def octa(o):
        return [[a + o, b + o, c + o] for a, b, c in OCTA]
    nv_d, disk = _grid(5, 4)
    known = [(6, OCTA, 1), (12, OCTA + octa(6), 2), (nv_d, disk, 0),
             (6 + nv_d, OCTA + [[a + 6, b + 6, c + 6] for a, b, c in disk], 1)]
    for nv, faces, want in known:
        cc = V.CellComplex(vc, nv, [], faces)
        assert cc.closed_components() == want, (nv, want)
        cc.release()
    rng, cases = _cases(803, 30)
    for i, (nv, extra, faces) in enumerate(cases):
        eo, fo, feo = oracle.cc_finalize(nv, extra, faces)
        foff, fv = _faces_csr(faces)
        ro = oracle.cc_prune(nv, eo, fo, feo, np.zeros(len(eo)), np.zeros(len(faces)), 1.0, 1.0)
        _, _, ot, _, _, _ = _oracle_remaining(oracle, eo, foff, fv, *ro[1:])
        first, pairs = {}, []
        for t, tri in enumerate(ot):
            for k in range(3):
                a, b = int(tri[k]), int(tri[(k + 1) % 3])
                key = (min(a, b), max(a, b))
                if key in first:
                    pairs.append((first[key], t))
                else:
                    first[key] = t
        want = _ncomp(len(ot), pairs) if len(ot) else 0
        cc = V.CellComplex(vc, nv, extra, faces)
        assert cc.closed_components() == want, i
        cc.release()


def test_resident_voxel_core(vc, oracle):
    """The inside part of the discrete Voronoi diagram of the torus, pruned
    at the -tt thresholds on one collapse state, against the oracle."""
    from voxel_ma_amd import synth
    occ = synth.torus_c1()
    vol = vc.upload(occ)
    vol.edt()
    vts, faces, fs, sites = V.voro_from_ft(vol)
    valid = V.voro_tag_points(vol, vts)
    vol.release()
    lam_f = np.linalg.norm(sites[fs[:, 0]] - sites[fs[:, 1]], axis=1).astype(np.float32)
    vi, eo_in, em_in, tr, ff, fm = V.voro_extract_inside(len(vts), [], faces, valid,
                                                         np.zeros(0, np.float32), lam_f)
    tris = [list(t) for t in np.asarray(tr, np.int64).reshape(-1, 3)]
    extra = np.asarray(eo_in, np.int64).reshape(-1, 2)
    nv = len(vi)
    eo, fo, feo = oracle.cc_finalize(nv, extra, tris)
    em = np.zeros(len(eo), np.float32)
    em[:len(em_in)] = em_in
    fm = np.asarray(fm, np.float32)
    foff, fv = _faces_csr(tris)
    cc = V.CellComplex(vc, nv, extra, tris)
    th = cc.thinning()
    th.assign(em, fm)
    size = float(max(np.ptp(np.argwhere(occ), axis=0)))
    for t in (0.01, 0.04, 0.2):
        thr = np.float32(size * t)
        rc, gens = th.prune(thr, thr)
        ro = oracle.cc_prune(nv, eo, fo, feo, em, fm, thr, thr)
        assert rc == ro[0] == 0
        for a, b in zip(ro[1:], th.removed()):
            assert np.array_equal(a, b), t
        got = th.remaining()
        want = _oracle_remaining(oracle, eo, foff, fv, *ro[1:])
        for a, b in zip(got, want[:4]):
            assert np.array_equal(a, b), t
        say("voxel core", len(tris), "triangles, t", t, "generations", gens,
            "remaining", len(got[2]))
    th.release()
    cc.release()


def test_resident_degenerate_complex(vc):
    """An edge with equal ends: the complex is created and its topology
    computed; the device collapse refuses it (VC_ERR_INVALID), as
    vc_cc_prune_device does."""
    cc = V.CellComplex(vc, 3, [(0, 0), (0, 1)], [[0, 1, 2]])
    t = cc.topology()
    assert t["V"] == 3 and t["F"] == 1
    th = cc.thinning()
    rc, _ = th.prune(1.0, 1.0)
    assert rc == V.VC_ERR_INVALID
    th.release()
    cc.release()

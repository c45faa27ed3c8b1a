"""GPU: the discrete Voronoi diagram kept resident on the device
(include/vc_vdiag.h) against the host-array entry points of the same steps,
each pinned to the oracle in test_gpu_voro.py / test_voro.py / test_ccthin.py:
vc_voro_from_ft + vc_cc_finalize (computeVD's geometry), vc_voro_radii,
vc_voro_tag_points, vc_voro_merge_close (host BFS version), vc_voro_measures
and vc_voro_extract_inside (host version), on the same volumes.  The chain
follows vol2ma (src/highlevelalgo.cpp:487-586, :662-837): radii, tag, merge
with the inside-size eps, extraction with measures, and the finalized
complex of the output."""
import numpy as np
import pytest

import voxel_ma_amd as V
from conftest import say

pytestmark = pytest.mark.gpu


def _ball(n, r):
    g = np.indices((n, n, n)) - (n - 1) / 2.0
    return ((g ** 2).sum(0) <= r * r).astype(np.uint8)


def _blobs(n, seed):
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    f = gaussian_filter(rng.standard_normal((n, n, n)), 2.5)
    return (f > np.quantile(f, 0.6)).astype(np.uint8)


def _volumes():
    from voxel_ma_amd import synth
    return [("torus", synth.torus_c1()), ("ball", _ball(24, 9.0)), ("blobs", _blobs(40, 7)),
            ("slab", np.pad(np.ones((6, 20, 9), np.uint8), 2))]


def _csr(nv, ne, edges, foff, fe):
    """edges of each vertex in edge order, faces of each edge in face order"""
    rows = edges.ravel().astype(np.int64)
    cols = np.repeat(np.arange(ne, dtype=np.int32), 2)
    o = np.lexsort((cols, rows))
    ve_off = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=nv))]).astype(np.int64)
    ve = cols[o].astype(np.int32)
    nf = len(foff) - 1
    fid = np.repeat(np.arange(nf, dtype=np.int32), np.diff(foff))
    o2 = np.lexsort((fid, fe))
    ef_off = np.concatenate([[0], np.cumsum(np.bincount(fe, minlength=ne))]).astype(np.int64)
    ef = fid[o2].astype(np.int32)
    return ve_off, ve, ef_off, ef


def _faces_of(foff, fv):
    return [list(fv[foff[f]:foff[f + 1]]) for f in range(len(foff) - 1)]


def _host_extract(vc, nv, edges, faces, fs, sites, valid):
    eo, foff, fe = V.cc_finalize(nv, edges, faces)
    fv = np.concatenate([np.asarray(f, np.int32) for f in faces]) if faces else \
        np.zeros(0, np.int32)
    fvalid = np.array([all(valid[v] for v in f) for f in faces], np.uint8)
    ve_off, ve, ef_off, ef = _csr(nv, len(eo), eo, foff, fe)
    lv, le, lf = V.voro_measures(vc, nv, eo, fs, sites, fvalid, ve_off, ve, ef_off, ef)
    out = V.voro_extract_inside(nv, eo, faces, valid, le, lf)
    return lv, out


def _same_extract(got, lv, host):
    vi, vm, eo, em, tr, ff, fm = got[:7]
    hvi, heo, hem, htr, hff, hfm = host
    assert np.array_equal(vi, np.asarray(hvi, np.int32))
    assert np.array_equal(vm, lv[vi])
    assert np.array_equal(eo, np.asarray(heo, np.int32).reshape(-1, 2))
    assert np.array_equal(em, np.asarray(hem, np.float32))
    assert np.array_equal(tr, np.asarray(htr, np.int32).reshape(-1, 3))
    assert np.array_equal(ff, np.asarray(hff, np.int32))
    assert np.array_equal(fm, np.asarray(hfm, np.float32))


def test_resident_diagram_matches_host_steps(vc):
    for name, occ in _volumes():
        vol = vc.upload(occ)
        vol.edt()
        vts, faces, fs, sites = V.voro_from_ft(vol)
        d = V.VoroDiagram(vol)
        g = d.get()
        say("vdiag", name, len(vts), "vertices", len(faces), "faces")
        # computeVD's geometry: the diagram finalized as a cellcomplex
        assert np.array_equal(g["vts"], vts) and np.array_equal(g["sites"], sites)
        assert np.array_equal(g["face_off"], 4 * np.arange(len(faces) + 1))
        assert np.array_equal(g["face_vts"], faces.reshape(-1))
        assert np.array_equal(g["face_sites"], fs)
        eo, _, _ = V.cc_finalize(len(vts), [], [list(f) for f in faces])
        assert np.array_equal(g["edges"], eo)
        assert not g["valid"].any() and g["radii"] is None
        # radii, tagging and the valid vertices' box
        d.radii()
        r = V.voro_radii(vc, vts, [list(f) for f in faces], fs, sites)
        assert np.array_equal(d.get()["radii"], r)
        d.tag(vol)
        valid = V.voro_tag_points(vol, vts).astype(np.uint8)
        info = d.info()
        assert np.array_equal(d.get()["valid"], valid)
        assert info["n_valid"] == int(valid.sum())
        if valid.any():
            assert np.array_equal(info["bb_min"], vts[valid == 1].min(0))
            assert np.array_equal(info["bb_max"], vts[valid == 1].max(0))
        # extraction before merging: validity from the tag
        lv, host = _host_extract(vc, len(vts), [], [list(f) for f in faces], fs, sites, valid)
        _same_extract(d.extract_inside(all_valid=False), lv, host)
        # a merge that joins vertices stays off the device and changes nothing
        assert not d.merge_close(2.0, True)
        assert np.array_equal(d.get()["edges"], eo)
        # vol2ma's merge: eps = inside size / longest axis * 1e-6
        size = float(np.max(np.asarray(info["bb_max"]) - np.asarray(info["bb_min"])))
        eps = np.float32(size * (1.0 / max(occ.shape)) * 0.000001)
        o2n, nvts, nfaces, rem, nedges, keep = V.voro_merge_close(vts, eo, faces, valid, eps,
                                                                 True)
        assert d.merge_close(eps, True)
        g2 = d.get()
        nvts = np.asarray(nvts, np.int64)
        kept = np.asarray(rem) == 0
        assert np.array_equal(g2["vts"], vts[nvts])
        assert np.array_equal(g2["radii"], r[nvts])
        assert g2["valid"].all() and d.info()["n_valid"] == len(nvts)
        assert np.array_equal(g2["face_sites"], fs[kept])
        e2, fo2, _ = V.cc_finalize(len(nvts), nedges, nfaces)
        assert np.array_equal(g2["edges"], e2)
        assert _faces_of(g2["face_off"], g2["face_vts"]) == [list(f) for f in nfaces]
        # extraction of the merged diagram and its finalized complex
        ones = np.ones(len(nvts), np.uint8)
        lv2, host2 = _host_extract(vc, len(nvts), nedges, nfaces, fs[kept], sites, ones)
        got = d.extract_inside(all_valid=False, keep_complex=True)
        _same_extract(got, lv2, host2)
        cc = got[7]
        ce, cfe = cc.finalized()
        he, _, hfe = V.cc_finalize(len(got[0]), got[2], [list(t) for t in got[4]])
        assert np.array_equal(ce, he) and np.array_equal(cfe, hfe)
        # every vertex valid (onlyHasInside): the same as all flags set
        _same_extract(d.extract_inside(all_valid=True), lv2, host2)
        cc.release()
        d.release()
        vol.release()


def test_resident_diagram_empty_and_errors(vc):
    """No inside voxel: an empty diagram through every step; a diagram
    before vc_edt is refused."""
    vol = vc.upload(np.zeros((8, 9, 10), np.uint8))
    with pytest.raises(V.VcError):
        V.VoroDiagram(vol)
    vol.edt()
    d = V.VoroDiagram(vol)
    i = d.info()
    assert (i["nv"], i["ne"], i["nf"]) == (0, 0, 0)
    d.tag(vol)
    assert d.merge_close(0.0, True)
    out = d.extract_inside()
    assert all(len(a) == 0 for a in out)
    d.release()
    vol.release()

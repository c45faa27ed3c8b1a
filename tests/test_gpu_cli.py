"""GPU: the C++ drop-in layer (Surfacer, VoxelThinning through voroUtility)
against the oracle."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import say

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import export_ref  # noqa: E402

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "voxel_ma_amd", "voroUtility")


def _read_off(path):
    with open(path) as f:
        assert f.readline().strip() == "OFF"
        nv, nf, _ = map(int, f.readline().split())
        v = np.array([list(map(float, f.readline().split())) for _ in range(nv)], np.float32)
        t = np.array([list(map(int, f.readline().split()))[1:] for _ in range(nf)], np.int64)
    return v, t


def _read_ply_points(path):
    with open(path) as f:
        lines = f.read().splitlines()
    start = lines.index("end_header") + 1
    return np.array([list(map(float, l.split())) for l in lines[start:] if l.strip()],
                    np.float32).reshape(-1, 3)


def test_cli_vol2mesh_matches_oracle(tmp_path, oracle):
    from voxel_ma_amd import synth
    occ = synth.torus_c1()
    raw = str(tmp_path / "t.raw")
    occ.tofile(raw)
    base = str(tmp_path / "t")
    say("cli vol2mesh torus")
    r = subprocess.run([CLI, "-md=vol2mesh", raw, base, "-dims=64,64,64", "-needEuler",
                        "-onlyBndryVts=false", "-write2node"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    v, t = _read_off(base + ".off")
    s, q, e = oracle.boundary(occ)
    tris = np.concatenate([q[:, [0, 1, 2]], q[:, [2, 3, 0]]], axis=1).reshape(-1, 3)
    assert np.array_equal(v, s) and np.array_equal(t, tris)
    assert "euler = V - E + F = %d" % e["euler"] in r.stdout
    with open(base + ".off", "rb") as f:
        assert f.read() == export_ref.off(s, tris)
    with open(base + ".node", "rb") as f:
        assert f.read() == export_ref.tetnodes(s)
    # default -onlyBndryVts: the boundary points alone
    say("cli vol2mesh torus points")
    r = subprocess.run([CLI, "-md=vol2mesh", raw, base + "_pts", "-dims=64,64,64"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    with open(base + "_pts.off", "rb") as f:
        assert f.read() == export_ref.off(s)


def test_cli_vol2skel_matches_oracle(tmp_path, oracle):
    from voxel_ma_amd import synth
    occ = synth.torus_c1()
    raw = str(tmp_path / "t.raw")
    occ.tofile(raw)
    prefix = str(tmp_path / "sk")
    tt = (0.01, 0.03)
    say("cli vol2skel torus")
    r = subprocess.run([CLI, "-md=vol2skel", raw, prefix, "-dims=64,64,64",
                        "-tt=%g,%g" % tt, "-outToDotma"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    levels = synth.levels_from_tt(synth.inside_size(occ), tt)
    d, f = oracle.edt_ft(occ)
    lab, _ = oracle.thin(occ, oracle.levels(occ, oracle.lambda2(occ, f), levels), len(levels))
    for k, t in enumerate(tt):
        pts = _read_ply_points("%s_thinned%g.ply" % (prefix, t))
        exp = np.argwhere((lab == 1) | (lab > 2 + k)).astype(np.float32)
        assert np.array_equal(pts, exp)
        idx = exp.astype(np.int64)
        rad = np.sqrt(d[idx[:, 0], idx[:, 1], idx[:, 2]].astype(np.float32))
        with open("%s_thinned%g.ply" % (prefix, t), "rb") as fh:
            assert fh.read() == export_ref.ply(exp)
        with open("%s_thinned%g.r" % (prefix, t), "rb") as fh:
            assert fh.read() == export_ref.radii(exp, rad)
        with open("%s_thinned%g.ma" % (prefix, t), "rb") as fh:
            assert fh.read() == export_ref.dotma(exp, rad)


def _read_ply(path):
    with open(path) as f:
        lines = f.read().splitlines()
    counts = {}
    for l in lines:
        if l.startswith("element "):
            _, name, n = l.split()
            counts[name] = int(n)
        if l == "end_header":
            break
    body = lines[lines.index("end_header") + 1:]
    nv, ne, nf = counts["vertex"], counts["edge"], counts["face"]
    v = np.array([list(map(float, l.split())) for l in body[:nv]], np.float32).reshape(-1, 3)
    e = np.array([list(map(int, l.split())) for l in body[nv:nv + ne]], np.int64).reshape(-1, 2)
    t = np.array([list(map(int, l.split()))[1:] for l in body[nv + ne:nv + ne + nf]],
                 np.int64).reshape(-1, 3)
    return v, e, t


def _chi(v, e, t):
    sides = {tuple(sorted(p)) for p in e.tolist()}
    for a, b, c in t.tolist():
        sides |= {tuple(sorted((a, b))), tuple(sorted((b, c))), tuple(sorted((c, a)))}
    return len(v) - len(sides) + len(t)


def test_cli_vol2ma_voxel_core(tmp_path, vc):
    """-md=vol2ma on a ball: the discrete Voronoi diagram of the feature
    transform, tagged, merged, its inside part written in full and
    FIFO-pruned per threshold.  The full complex holds exactly the valid
    vertices and faces (all four corners inside) of vc_voro_from_ft, each
    quad as two triangles; a collapse is a homotopy equivalence, so every
    pruned complex has the full complex's Euler characteristic."""
    import voxel_ma_amd as V
    n = 24
    g = np.indices((n, n, n)) - (n - 1) / 2.0
    occ = ((g ** 2).sum(0) <= 9.0 ** 2).astype(np.uint8)
    raw = str(tmp_path / "ball.raw")
    occ.tofile(raw)
    out = str(tmp_path / "ball.ply")
    say("cli vol2ma ball")
    r = subprocess.run([CLI, "-md=vol2ma", raw, out, "-dims=%d,%d,%d" % (n, n, n),
                        "-tt=0.04,0.2", "-fullOrPruned=2", "-outToDotma"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    vol = vc.upload(occ)
    vol.edt()
    vts, faces, fs, sites = V.voro_from_ft(vol)
    valid = V.voro_tag_points(vol, vts).astype(bool)
    vol.release()
    vf = faces[valid[faces].all(1)]
    v, e, t = _read_ply(out)
    assert len(v) == int(valid.sum()) and len(t) == 2 * len(vf)
    assert np.array_equal(np.sort(v, axis=0), np.sort(vts[valid], axis=0))
    chi_full = _chi(v, e, t)
    for tt in ("0.04", "0.2"):
        vt, et, tt_ = _read_ply(str(tmp_path / ("ball_thinned%s.ply" % tt)))
        assert 0 < len(vt) <= len(v)
        assert _chi(vt, et, tt_) == chi_full, tt
        assert os.path.exists(str(tmp_path / ("ball_thinned%s.r" % tt)))
        assert os.path.exists(str(tmp_path / ("ball_thinned%s.ma" % tt)))


def test_cli_vol2ma_resident_equals_host_chain(tmp_path):
    """-md=vol2ma with the diagram resident on the device (computeVD keeps
    it there through tagging, merging, extraction and the collapse) writes
    the same bytes as the chain over host arrays (-hostVoro): the full
    inside part, each pruned complex, the radii and the .ma files."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(11)
    f = gaussian_filter(rng.standard_normal((48, 44, 40)), 3.0)
    occ = (f > np.quantile(f, 0.55)).astype(np.uint8)
    raw = str(tmp_path / "blobs.raw")
    occ.tofile(raw)
    outs = {}
    for mode in ("device", "host"):
        d = tmp_path / mode
        d.mkdir()
        cmd = [CLI, "-md=vol2ma", raw, str(d / "core.ply"), "-dims=48,44,40",
               "-tt=0.04,0.2", "-fullOrPruned=2", "-outToDotma"]
        if mode == "host":
            cmd.append("-hostVoro")
        say("cli vol2ma", mode)
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        outs[mode] = {p.name: p.read_bytes() for p in sorted(d.iterdir())}
    assert sorted(outs["device"]) == sorted(outs["host"])
    assert len(outs["device"]) == 8
    for name in outs["device"]:
        assert outs["device"][name] == outs["host"][name], name


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

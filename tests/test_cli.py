"""GPU: the voroUtility front end's topo mode (closed pockets of a mesh via
the cell-complex collapse, src/highlevelalgo.cpp:993-1020), which finalizes,
collapses and counts components on the device."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "voxel_ma_amd", "voroUtility")


def _write_off(path, vts, faces):
    with open(path, "w") as f:
        f.write("OFF\n%d %d 0\n" % (len(vts), len(faces)))
        for v in vts:
            f.write("%g %g %g\n" % tuple(v))
        for fc in faces:
            f.write("%d %s\n" % (len(fc), " ".join(str(i) for i in fc)))


def _topo(tmp_path, vts, faces):
    p = str(tmp_path / "m.off")
    _write_off(p, vts, faces)
    r = subprocess.run([CLI, "-md=topo", "-meshToCheck=" + p], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith("# closed components")][0]
    return int(line.split(":")[1])


def test_topo_closed_octahedron(tmp_path):
    vts = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    faces = [[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5],
             [0, 3, 5]]
    assert _topo(tmp_path, vts, faces) == 1


def test_topo_open_disk_and_two_spheres(tmp_path):
    assert _topo(tmp_path, [(0, 0, 0), (1, 0, 0), (0, 1, 0)], [[0, 1, 2]]) == 0
    vts = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    oct_f = [[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5],
             [0, 3, 5]]
    vts2 = vts + [(x + 5, y, z) for x, y, z in vts]
    faces2 = oct_f + [[a + 6, b + 6, c + 6] for a, b, c in oct_f]
    assert _topo(tmp_path, vts2, faces2) == 2

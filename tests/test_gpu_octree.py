"""Octree input on the device: the dense (R+1)^3 occupancy built by
vc_volume_from_octree equals OctreeVolume::getDataAt > 0 (oracle) at every
voxel, and the octree volume runs through EDT and thinning like any other."""
import os
import sys

import numpy as np
import pytest

import voxel_ma_amd as V
from conftest import ROOT, say

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import octree_ref as O  # noqa: E402

pytestmark = pytest.mark.gpu


def _dense_oracle(R, root):
    n = R + 1
    out = np.zeros((n, n, n), np.uint8)
    for x in range(n):
        for y in range(n):
            for z in range(n):
                out[x, y, z] = O.get_data_at(R, root, x, y, z) > 0
    return out


@pytest.mark.parametrize("R", [2, 8, 16])
def test_dense_occupancy_matches_oracle(vc, R):
    rng = np.random.default_rng(100 + R)
    for it in range(3):
        root = O.random_tree(R, rng)
        for sog in (False, True):
            data = O.write_sog(R, root) if sog else O.write_sof(R, root)
            say("octree", R, it, "sog" if sog else "sof")
            t = V.Octree(data, sog=sog)
            vol = t.volume(vc)
            got = vol.get(V.BUF_OCC)
            assert np.array_equal(got, _dense_oracle(R, root))
            vol.release()
            t.free()


def test_octree_volume_through_the_path(vc, oracle):
    rng = np.random.default_rng(7)
    R = 32
    root = O.random_tree(R, rng, p_split=0.7, p_empty=0.2)
    t = V.Octree(O.write_sof(R, root))
    vol = t.volume(vc)
    occ = vol.get(V.BUF_OCC)
    vol.edt().measure()
    levels = [1.0, 2.0, 3.5]
    vol.thin_levels(levels)
    lab = vol.get(V.BUF_LABEL)
    d, f = oracle.edt_ft(occ)
    assert np.array_equal(vol.get(V.BUF_EDT), d)
    lab_o, _ = oracle.thin(occ, oracle.levels(occ, oracle.lambda2(occ, f), levels), 3)
    assert np.array_equal(lab, lab_o)
    vol.release()
    t.free()

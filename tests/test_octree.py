"""Octree input (.sof / .sog, include/vc_octree.h), host side: parsing,
OctreeVolume::getDataAt on every coordinate and the Surfacer leaf centres
against oracle/octree_ref.py, which is pinned by hand-derived answers."""
import os
import sys

import numpy as np
import pytest

import voxel_ma_amd as V

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import octree_ref as O  # noqa: E402


def _hand_tree():
    """R = 2: one internal root; child (0,0,0) a leaf with byte 0b00000010
    (only corner 1 = (0,0,1) outside), child (1,1,1) an empty node with sign
    1 (outside), every other child an empty node with sign 0 (inside)."""
    kids = [O.Node(O.EMPTY, 0) for _ in range(8)]
    kids[0] = O.Node(O.LEAF, 0b10)
    kids[7] = O.Node(O.EMPTY, 1)
    return 2, O.Node(O.INTERNAL, 0, kids)


def test_known_answers():
    R, root = _hand_tree()
    for sog in (False, True):
        data = O.write_sog(R, root, (1.0, 2.0, 3.0), 4.0) if sog else O.write_sof(R, root)
        t = V.Octree(data, sog=sog)
        assert t.info.resolution == 2 and t.info.n_leaf == 1 and t.info.n_empty == 7
        assert t.info.n_internal == 1
        if sog:
            assert list(t.info.lower_left) == [1.0, 2.0, 3.0] and t.info.bbox_len == 4.0
        # cell (0,0,0): bit 0 of 0b10 is 0 -> inside
        assert t.value(0, 0, 0) == 1.0 == O.get_data_at(R, O.parse(data, sog)[1], 0, 0, 0)
        # cell (1,1,1) is the outside empty node; (2,2,2) reads its corner 7
        assert t.value(1, 1, 1) == -1.0 and t.value(2, 2, 2) == -1.0
        # (0,0,2) reads corner bit 1 (z) of cell (0,0,1): an inside empty node
        assert t.value(0, 0, 2) == 1.0
        # the leaf is the only non-homogeneous one: centre (0.5, 0.5, 0.5)
        assert t.leaf_centers().tolist() == [[0.5, 0.5, 0.5]]
        t.free()


def test_z_boundary_corner_bit():
    """A single leaf R = 1: coordinate 1 on z reads bit 1 of the byte."""
    for byte, expect in ((0b10, -1.0), (0b01, 1.0)):
        root = O.Node(O.LEAF, byte)
        t = V.Octree(O.write_sof(1, root))
        assert t.value(0, 0, 1) == expect == O.get_data_at(1, root, 0, 0, 1)
        assert t.value(0, 0, 0) == O.get_data_at(1, root, 0, 0, 0)
        t.free()


@pytest.mark.parametrize("R", [4, 8, 16])
def test_random_trees_match_oracle(R):
    rng = np.random.default_rng(R)
    for it in range(4):
        root = O.random_tree(R, rng)
        for sog in (False, True):
            data = O.write_sog(R, root) if sog else O.write_sof(R, root)
            t = V.Octree(data, sog=sog)
            for x in range(R + 1):
                for y in range(R + 1):
                    for z in range(R + 1):
                        assert t.value(x, y, z) == O.get_data_at(R, root, x, y, z)
            got = [tuple(float(c) for c in p) for p in t.leaf_centers()]
            assert got == O.leaf_centers(R, root)
            t.free()


def test_malformed_streams_rejected():
    R, root = _hand_tree()
    data = O.write_sof(R, root)
    for bad in (data[:-1], data[:4] + b"\x07", b"\x03\x00\x00\x00" + data[4:],
                b"SOG" + data):
        with pytest.raises(V.VcError):
            V.Octree(bad)

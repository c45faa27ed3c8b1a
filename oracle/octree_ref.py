"""CPU restatement of the octree input (test infrastructure: the checker
for include/vc_octree.h).

  parse       OctreeVolume::read_sof_file / read_sog_file
              (src/octree.cpp:259-497): pre-order node stream, children
              (i, j, k) with k fastest, .sog leaves carry 3 extra floats
  get_data_at OctreeVolume::getDataAt (src/octree.cpp:200-255) via
              get_node_values (:676-774): +1 inside (bit 0 of the cell's
              sign byte, or the corner bit for a coordinate equal to R)
  leaf_centers  Surfacer::extractBoundaryVts octree branch
              (src/surfacing.cpp:286-314, :666-686)
  write_sof / write_sog  the same streams, for building test inputs
Pinned by hand-derived known answers in tests/test_octree.py.
"""
import struct

import numpy as np

INTERNAL, EMPTY, LEAF = 0, 1, 2


class Node:
    def __init__(self, type_, value=0, children=None):
        self.type, self.value, self.children = type_, value, children or []


def _read_node(buf, at, sog):
    t = buf[at]
    at += 1
    if t in (LEAF, EMPTY):
        v = buf[at]
        at += 1
        if t == LEAF and sog:
            at += 12
        return Node(t, v), at
    if t != INTERNAL:
        raise ValueError("bad node type %d" % t)
    kids = []
    for _ in range(8):
        c, at = _read_node(buf, at, sog)
        kids.append(c)
    return Node(INTERNAL, 0, kids), at


def parse(data, sog=False):
    at = 0
    hdr = None
    if sog:
        hdr = struct.unpack_from("<4f", data, 15)
        at = 128
    R = struct.unpack_from("<i", data, at)[0]
    root, _ = _read_node(data, at + 4, sog)
    return R, root, hdr


def _node_values(node, i, j, k, off, ln):
    if not (off[0] <= i < off[0] + ln and off[1] <= j < off[1] + ln and
            off[2] <= k < off[2] + ln):
        return False, 0
    if node.type == LEAF:
        return True, node.value
    if node.type == EMPTY:
        return True, 0xFF if node.value == 1 else 0
    h = ln // 2
    q = 0
    for a in range(2):
        for b in range(2):
            for c in range(2):
                found, v = _node_values(node.children[q], i, j, k,
                                        (off[0] + a * h, off[1] + b * h, off[2] + c * h), h)
                q += 1
                if found:
                    return True, v
    return False, 0


def get_data_at(R, root, x, y, z):
    off = 0
    if x == R:
        x, off = R - 1, off | 4
    if y == R:
        y, off = R - 1, off | 2
    if z == R:
        z, off = R - 1, off | 1
    _, v = _node_values(root, x, y, z, (0, 0, 0), R)
    return 1 if ((v >> off) & 1) == 0 else -1


def leaf_centers(R, root):
    out = []

    def walk(n, off, ln):
        if n.type == LEAF:
            if n.value not in (0, 0xFF):
                out.append((off[0] + 0.5, off[1] + 0.5, off[2] + 0.5))
            return
        if n.type == INTERNAL:
            h = ln // 2
            q = 0
            for a in range(2):
                for b in range(2):
                    for c in range(2):
                        walk(n.children[q], (off[0] + a * h, off[1] + b * h, off[2] + c * h), h)
                        q += 1

    walk(root, (0, 0, 0), R)
    return out


def _write_node(n, out, sog):
    out.append(n.type)
    if n.type in (LEAF, EMPTY):
        out.append(n.value)
        if n.type == LEAF and sog:
            out.extend(struct.pack("<3f", 0.0, 0.0, 0.0))
        return
    for c in n.children:
        _write_node(c, out, sog)


def write_sof(R, root):
    out = bytearray(struct.pack("<i", R))
    _write_node(root, out, False)
    return bytes(out)


def write_sog(R, root, lower_left=(0.0, 0.0, 0.0), bbox_len=1.0):
    hdr = bytearray(128)
    title = b"SOG.Format 1.0\x00"
    hdr[:len(title)] = title
    struct.pack_into("<4f", hdr, len(title), *lower_left, bbox_len)
    out = bytearray(hdr) + struct.pack("<i", R)
    _write_node(root, out, True)
    return bytes(out)


def random_tree(R, rng, p_split=0.6, p_empty=0.3):
    """Leaves at unit cells, empty nodes anywhere (as the reference files)."""
    def make(ln):
        if ln == 1:
            return Node(LEAF, int(rng.integers(0, 256)))
        r = rng.random()
        if r < p_empty:
            return Node(EMPTY, int(rng.integers(0, 2)))
        if r < p_empty + p_split or ln > R // 2:
            return Node(INTERNAL, 0, [make(ln // 2) for _ in range(8)])
        return Node(EMPTY, int(rng.integers(0, 2)))
    return make(R)

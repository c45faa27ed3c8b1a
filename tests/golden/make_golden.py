"""Writes the golden fixtures under tests/golden/ (run in the build container).

  edt_scipy.npz       inputs + squared EDT from scipy.ndimage.distance_transform_edt
                      (a library routine independent of the oracle), squared and
                      rounded to int32.
  cube_components.json  the per-mask surface-component count of the reference's
                      marching-cubes cycle table (reference cycle8.txt, read as
                      data; only the derived 256 counts and the file's SHA-256 are
                      stored), with the four hole masks forced to 0 as
                      src/surfacing.cpp:545-550 does.

Usage: python tests/golden/make_golden.py [--reference /root/reference]
"""
import argparse
import hashlib
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def edt_fixtures():
    from scipy import ndimage
    from voxel_ma_amd import synth

    rng = np.random.default_rng(1234)
    vols = {
        "rand_24x20x16": (rng.random((24, 20, 16)) < 0.7).astype(np.uint8),
        "rand_17x9x33": (rng.random((17, 9, 33)) < 0.9).astype(np.uint8),
        "torus_c1": synth.torus_c1(),
    }
    g = np.indices((32, 32, 32))
    vols["ball_32"] = (((g - 15.5) ** 2).sum(0) <= 12.0 ** 2).astype(np.uint8)
    out = {}
    for name, occ in vols.items():
        d = ndimage.distance_transform_edt(occ)
        out[name + "__occ"] = occ
        out[name + "__d2"] = np.rint(d * d).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "edt_scipy.npz"), **out)


def cube_components(ref):
    path = os.path.join(ref, "cycle8.txt")
    raw = open(path, "rb").read()
    lines = raw.decode().strip().split("\n")
    assert len(lines) == 256
    counts = []
    for line in lines:
        parent = {}

        def find(a):
            while parent[a] != a:
                parent[a] = parent[parent[a]]
                a = parent[a]
            return a

        for cyc in re.findall(r"\{(\{\d+, \d+\}(?:, \{\d+, \d+\})*)\}", line):
            edges = [tuple(map(int, e)) for e in re.findall(r"\{(\d+), (\d+)\}", cyc)]
            for e in edges:
                parent.setdefault(e, e)
            for a, b in zip(edges, edges[1:]):
                ra, rb = find(a), find(b)
                if ra != rb:
                    parent[ra] = rb
        counts.append(len({find(e) for e in parent}))
    for m in (126, 189, 219, 231):
        counts[m] = 0
    doc = {"source": "reference cycle8.txt (data table), components of cycle-linked edges",
           "sha256": hashlib.sha256(raw).hexdigest(), "counts": counts}
    with open(os.path.join(HERE, "cube_components.json"), "w") as f:
        json.dump(doc, f)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    a = ap.parse_args()
    edt_fixtures()
    cube_components(a.reference)
    print("fixtures written to", HERE)

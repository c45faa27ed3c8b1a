"""GPU parity at the full benchmark configurations, checked against the CPU
oracle (its OpenMP loops on the host's cores):

* the 512^3 blurred-noise volume of the headline step (sigma 12, 70th
  percentile, 1-voxel shell, seed 3; SHA-256 prefix 506bc3377c18ad94):
  EDT, feature transform, lambda^2, levels and labels bit for bit, the
  removed count per level and the number of sweeps, Betti numbers of the
  skeleton at every level equal to the object's; boundary sites, quads and
  Euler terms against the Surfacer restatement;
* the batch of 256 solids (edges 32-191, seed 4): every member under the
  autotuned schedule;
* the learned measure on the 512^3 volume: tiled inference equals the
  whole-volume pass, and the exact thinning under the gated measure equals
  the oracle's thinning under the same measure, with Betti numbers kept.

The thresholds follow exportInsideVoroMesh's -tt loop: level = insideSize * t
for t in 0.01, 0.02, 0.04 (reference src/highlevelalgo.cpp:825-837).
"""
import hashlib
import time

import numpy as np
import pytest

import voxel_ma_amd as V
from conftest import say

pytestmark = pytest.mark.gpu

C3_SHA16 = "506bc3377c18ad94"
C4_SHA16 = "1c82b5e17d7d46b8"


@pytest.fixture(scope="module")
def c3(vc, oracle):
    from voxel_ma_amd import synth
    n = 512
    say("C3 generate %d^3 blurred noise" % n)
    occ_dev, _ = vc.synth_blurred_noise((n, n, n), sigma=12.0, fill=0.30, seed=3)
    occ = occ_dev.download(np.uint8, (n, n, n))
    vol = vc.from_device(occ_dev.ptr, (n, n, n))
    vc.synchronize()
    occ_dev.free()
    levels = synth.levels_from_tt(synth.inside_size(occ))
    oracle.set_threads(0)
    yield occ, vol, levels
    vol.release()


def _skeleton(lab, k):
    return ((lab == 1) | (lab > 2 + k)).astype(np.uint8)


def test_c3_headline_volume_bit_exact(vc, oracle, c3):
    occ, vol, levels = c3
    assert hashlib.sha256(occ.tobytes()).hexdigest()[:16] == C3_SHA16
    say("C3 GPU step, levels", levels)
    vol.edt().measure()
    st = vol.thin_levels(levels)
    t0 = time.time()
    d, f = oracle.edt_ft(occ)
    assert np.array_equal(vol.get(V.BUF_EDT), d), "edt"
    assert np.array_equal(vol.get(V.BUF_FT), f), "ft"
    del d
    lam = oracle.lambda2(occ, f)
    del f
    assert np.array_equal(vol.get(V.BUF_LAM2).view(np.int32), lam), "lam2"
    lvl = oracle.levels(occ, lam, levels)
    del lam
    assert np.array_equal(vol.get(V.BUF_LEVEL), lvl), "levels"
    lab, sto = oracle.thin(occ, lvl, len(levels))
    say("C3 oracle EDT..thinning %.1f s" % (time.time() - t0))
    assert np.array_equal(vol.get(V.BUF_LABEL), lab), "labels"
    for k in range(len(levels)):
        assert st.removed[k] == sto.removed[k], k
    assert st.passes == 8 * sum(sto.sweeps[k] for k in range(len(levels)))
    t0 = time.time()
    b = oracle.betti(occ)
    for k in range(len(levels)):
        assert oracle.betti(_skeleton(lab, k)) == b, k
    say("C3 Betti", b, "kept at %d levels (%.1f s)" % (len(levels), time.time() - t0))


def test_c3_boundary_matches_surfacer_restatement(vc, oracle, c3):
    occ, vol, _ = c3
    say("C3 boundary")
    s, q, e = vol.boundary()
    t0 = time.time()
    so, qo, eo = oracle.boundary(occ)
    say("C3 boundary oracle %.1f s: %d sites, %d quads" % (time.time() - t0, len(so), len(qo)))
    assert np.array_equal(s, so)
    assert np.array_equal(q, qo)
    for key, v in eo.items():
        assert e[key] == v, key
    assert e["chi"] == oracle.euler_voxels(occ)


def test_c4_full_batch_tuned_schedule(vc, oracle):
    from voxel_ma_amd import synth
    specs = synth.batch_c4_specs(count=256, seed=4)
    members, levels = [], []
    h = hashlib.sha256()
    say("C4 generate 256 solids")
    for L, prims in specs:
        v = vc.solids_volume((L, L, L), prims)
        occ = v.get(V.BUF_OCC)
        h.update(occ.tobytes())
        levels.append(synth.levels_from_tt(synth.inside_size(occ)))
        members.append(v)
    assert h.hexdigest()[:16] == C4_SHA16
    levels = np.array(levels, np.float32)
    b = V.Batch(vc, members)
    b.edt().measure()
    sc, _ = b.tune(levels, reps=2)
    say("C4 tuned schedule", sc.as_dict())
    st = b.run(levels)
    oracle.set_threads(0)
    t0 = time.time()
    removed = np.zeros(levels.shape[1], np.int64)
    for i, (m, (L, prims)) in enumerate(zip(members, specs)):
        occ = m.get(V.BUF_OCC)
        lv = [float(x) for x in levels[i]]
        d, f = oracle.edt_ft(occ)
        lam = oracle.lambda2(occ, f)
        lvl = oracle.levels(occ, lam, lv)
        lab, sto = oracle.thin(occ, lvl, len(lv))
        assert np.array_equal(m.get(V.BUF_EDT), d), (i, "edt")
        assert np.array_equal(m.get(V.BUF_FT), f), (i, "ft")
        assert np.array_equal(m.get(V.BUF_LAM2).view(np.int32), lam), (i, "lam2")
        assert np.array_equal(m.get(V.BUF_LEVEL), lvl), (i, "level")
        assert np.array_equal(m.get(V.BUF_LABEL), lab), (i, "label")
        removed += np.array(sto.removed[:len(lv)])
    say("C4 oracle over 256 volumes %.1f s" % (time.time() - t0))
    assert list(st.removed[:levels.shape[1]]) == list(removed)
    b.release()
    for m in members:
        m.release()


def test_c5_surrogate_on_headline_volume(vc, oracle, c3):
    from voxel_ma_amd import surrogate
    occ, vol, levels = c3
    tau = 0.01
    net = V.Surrogate(vc, seed=1, path=surrogate.WEIGHTS)
    vol.edt()
    say("C5 whole-volume inference 512^3")
    lam_w, rel_w = net.infer(vol, 0)
    say("C5 tiled inference, core 192, halo", V.SURROGATE_HALO)
    lam_t, rel_t = net.infer(vol, 192)
    assert np.array_equal(lam_w, lam_t) and np.array_equal(rel_w, rel_t)
    del lam_t, rel_t
    st = net.measure(vol, max_rel_err=tau)
    inside = occ > 0
    assert st.inside == int(inside.sum())
    assert st.accepted == int((rel_w[inside] <= tau).sum())
    vol.thin_levels(levels)
    lab = vol.get(V.BUF_LABEL)
    lam2 = vol.get(V.BUF_LAM2) if st.fallback else np.zeros(occ.shape, np.uint32)
    meas = np.where(rel_w <= np.float32(tau), np.maximum(lam_w, np.float32(0)),
                    np.sqrt(lam2.astype(np.float32)))
    del lam_w, rel_w, lam2
    lvl = np.full(occ.shape, 255, np.uint8)
    for k in reversed(range(len(levels))):
        lvl[inside & (meas < np.float32(levels[k]))] = k
    del meas
    assert np.array_equal(vol.get(V.BUF_LEVEL), lvl)
    lab_o, _ = oracle.thin(occ, lvl, len(levels))
    assert np.array_equal(lab, lab_o)
    b = oracle.betti(occ)
    for k in range(len(levels)):
        assert oracle.betti(_skeleton(lab, k)) == b, k
    say("C5 gate", st.as_dict(), "Betti", b)
    net.release()

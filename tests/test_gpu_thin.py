"""GPU parity: measure-ordered subfield thinning through the C ABI against the
CPU oracle's dense schedule (bit-exact labels) plus topology invariants."""
import itertools

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def _gpu_thin(vc, occ, levels):
    vol = vc.upload(occ)
    vol.edt().measure()
    st = vol.thin_levels(levels)
    lab, lvl = vol.get(4), vol.get(5)
    vol.release()
    return lab, lvl, st


def _oracle_thin(oracle, occ, levels):
    d, f = oracle.edt_ft(occ)
    lam = oracle.lambda2(occ, f)
    lvl = oracle.levels(occ, lam, levels)
    lab, st = oracle.thin(occ, lvl, len(levels))
    return lab, lvl, st


def _compare(vc, oracle, occ, levels, name):
    say("thin", name, occ.shape, levels)
    lab, lvl, st = _gpu_thin(vc, occ, levels)
    labo, lvlo, sto = _oracle_thin(oracle, occ, levels)
    assert np.array_equal(lvl, lvlo), name
    assert np.array_equal(lab, labo), name
    for k in range(len(levels)):
        assert st.removed[k] == sto.removed[k]
    # the same number of sweeps as the dense schedule (8 passes per sweep)
    assert st.passes == 8 * sum(sto.sweeps[k] for k in range(len(levels)))
    return lab


def test_thin_random_small(vc, oracle):
    rng = np.random.default_rng(9)
    for i in range(25):
        shape = tuple(int(s) for s in rng.integers(2, 30, size=3))
        occ = (rng.random(shape) < rng.choice([0.5, 0.8, 0.95])).astype(np.uint8)
        levels = sorted(float(x) for x in rng.uniform(0.5, 8.0, size=int(rng.integers(1, 5))))
        _compare(vc, oracle, occ, levels, "rand%d" % i)


def test_thin_config1_torus(vc, oracle):
    from voxel_ma_amd import synth
    occ = synth.torus_c1()
    for levels in ([1.5, 3.0, 6.0], synth.levels_from_tt(synth.inside_size(occ), (0.04,))):
        lab = _compare(vc, oracle, occ, levels, "c1")
        b = oracle.betti(occ)
        for k in range(len(levels)):
            assert oracle.betti(((lab == 1) | (lab > 2 + k)).astype(np.uint8)) == b


def test_thin_config2_spheres(vc, oracle):
    from voxel_ma_amd import synth
    occ = synth.spheres_c2()
    levels = synth.levels_from_tt(synth.inside_size(occ))
    lab = _compare(vc, oracle, occ, levels, "c2")
    b = oracle.betti(occ)
    sk = ((lab == 1) | (lab > 2 + len(levels) - 1)).astype(np.uint8)
    assert oracle.betti(sk) == b


def test_thin_float_measure_path(vc, oracle):
    """vc_set_measure_f32: the thinning uses the caller's float measure."""
    import voxel_ma_amd
    from voxel_ma_amd import synth
    occ = synth.torus_c1()
    rng = np.random.default_rng(3)
    meas = rng.uniform(0, 5, size=occ.shape).astype(np.float32)
    levels = [1.0, 2.5, 4.0]
    say("thin float measure")
    vol = vc.upload(occ)
    dm = voxel_ma_amd.DeviceBuffer.from_array(meas)
    vol.set_measure_f32(dm.ptr)
    vol.thin_levels(levels)
    lab, lvl = vol.get(4), vol.get(5)
    vol.release()
    dm.free()
    lvlo = np.full(occ.shape, 255, np.uint8)
    for k in reversed(range(len(levels))):
        lvlo[(occ > 0) & (meas < np.float32(levels[k]))] = k
    assert np.array_equal(lvl, lvlo)
    labo, _ = oracle.thin(occ, lvlo, len(levels))
    assert np.array_equal(lab, labo)


def test_thin_empty_and_full(vc, oracle):
    for name, occ in {"empty": np.zeros((9, 9, 9), np.uint8),
                      "full": np.ones((9, 9, 9), np.uint8),
                      "one": np.pad(np.ones((1, 1, 1), np.uint8), 2)}.items():
        _compare(vc, oracle, occ, [2.0, 5.0], name)

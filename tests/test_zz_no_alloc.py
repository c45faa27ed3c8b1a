"""GPU: vc_skeleton_noalloc (the entry point that allocates nothing) against
the allocating path it mirrors (vc_volume_from_device + vc_edt + vc_measure +
vc_thin_levels + vc_get_device(VC_BUF_LABEL)), through the C ABI: identical
labels and statistics on two calls in a row, the second on a smaller volume
(the scratch array keeps the first call's contents), then the limits."""
import ctypes

import numpy as np
import pytest

import voxel_ma_amd as V
from conftest import say

pytestmark = pytest.mark.gpu


def _reference_path(vc, d_occ, shape, levels):
    vol = vc.from_device(d_occ.ptr, shape)
    vol.edt().measure()
    st = vol.thin_levels(levels)
    out = V.DeviceBuffer(int(np.prod(shape)))
    vol.get_device(V.BUF_LABEL, out.ptr, out.nbytes)
    vc.synchronize()
    lab = out.download(np.uint8, shape)
    out.free()
    vol.release()
    return lab, st


def _noalloc(vc, d_occ, shape, levels, d_label):
    lv = np.ascontiguousarray(levels, np.float32)
    st = V.ThinStats()
    vc.check(vc._lib.vc_skeleton_noalloc(vc.handle, ctypes.c_void_p(d_occ.ptr), *shape,
                                         lv.ctypes.data_as(ctypes.c_void_p), len(lv),
                                         ctypes.c_void_p(d_label.ptr), ctypes.byref(st)))
    vc.synchronize()
    return d_label.download(np.uint8, shape), st


def _volume(shape, seed):
    from voxel_ma_amd import synth
    rng = np.random.default_rng(seed)
    prims = synth.solid_specs(min(shape), rng)
    occ = synth.solids_host(min(shape), prims)
    out = np.zeros(shape, np.uint8)
    s = tuple(min(a, b) for a, b in zip(shape, occ.shape))
    out[:s[0], :s[1], :s[2]] = occ[:s[0], :s[1], :s[2]]
    return out


def test_noalloc_matches_allocating_path_twice(vc, oracle):
    from voxel_ma_amd import synth
    # the second call is smaller: the scratch array still holds the first
    # call's larger state
    cases = [((128, 120, 136), 5), ((40, 36, 70), 6)]
    for shape, seed in cases:
        occ = _volume(shape, seed)
        levels = synth.levels_from_tt(synth.inside_size(occ))
        say("noalloc", shape, levels)
        d_occ = V.DeviceBuffer.from_array(occ)
        d_label = V.DeviceBuffer(occ.size)
        lab, st = _noalloc(vc, d_occ, shape, levels, d_label)
        lab_ref, st_ref = _reference_path(vc, d_occ, shape, levels)
        d_occ.free()
        d_label.free()
        assert np.array_equal(lab, lab_ref), shape
        for k in range(len(levels)):
            assert st.removed[k] == st_ref.removed[k]
        assert st.passes == st_ref.passes and st.tests == st_ref.tests
        assert st.voxel_passes == st_ref.voxel_passes
        # and the oracle's labels
        d, f = oracle.edt_ft(occ)
        lo, _ = oracle.thin(occ, oracle.levels(occ, oracle.lambda2(occ, f), levels),
                            len(levels))
        assert np.array_equal(lab, lo), shape


def test_noalloc_rejects_beyond_limits(vc):
    d = V.DeviceBuffer(16)
    lv = np.array([1.0], np.float32)
    # too many voxels; few enough voxels but one plane word per voxel (nz = 1)
    # overflows the scratch layout; an empty axis
    for shape in [(256, 256, 65), (2048, 2048, 1), (0, 4, 4)]:
        rc = vc._lib.vc_skeleton_noalloc(vc.handle, ctypes.c_void_p(d.ptr), *shape,
                                         lv.ctypes.data_as(ctypes.c_void_p), 1,
                                         ctypes.c_void_p(d.ptr), None)
        assert rc == V.VC_ERR_INVALID, shape
        say("rejected", shape, vc.last_error())
    d.free()

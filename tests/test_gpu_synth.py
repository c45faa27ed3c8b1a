"""GPU: the device generator of the blurred-noise volume (benchmark input)."""
import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def test_blurred_noise_generator(vc):
    from voxel_ma_amd import synth
    n = 48
    say("synth blurred noise", n)
    buf, thr = vc.synth_blurred_noise((n, n, n), sigma=3.0, fill=0.3, seed=7)
    occ = buf.download(np.uint8, (n, n, n))
    buf2, thr2 = vc.synth_blurred_noise((n, n, n), sigma=3.0, fill=0.3, seed=7)
    occ2 = buf2.download(np.uint8, (n, n, n))
    buf.free()
    buf2.free()
    assert np.array_equal(occ, occ2) and thr == thr2          # deterministic
    assert occ[0].sum() == occ[-1].sum() == occ[:, 0].sum() == occ[:, :, -1].sum() == 0
    host = synth.blur_c3_host(n, sigma=3.0, fill=0.3, seed=7)
    assert (occ != host).mean() < 1e-3                          # same spec on the host
    # exactly ceil(0.3*N) voxels above the threshold before the shell is cleared
    assert abs(occ.sum() - host.sum()) <= 0.001 * occ.size

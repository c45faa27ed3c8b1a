"""GPU parity of the batched path (include/vc_batch.h): every member volume of
a batch gets exactly the oracle's EDT, FT, levels and labels (the same as
the single-volume path), under every schedule, with captured graph segments
replayed repeatedly; the device solids generator equals its host
restatement byte for byte."""
import numpy as np
import pytest

import voxel_ma_amd as V
from conftest import say

pytestmark = pytest.mark.gpu


def _oracle_all(oracle, occ, levels):
    d, f = oracle.edt_ft(occ)
    lam = oracle.lambda2(occ, f)
    lvl = oracle.levels(occ, lam, levels)
    lab, st = oracle.thin(occ, lvl, len(levels))
    return d, f, lvl, lab, st


def _ragged_volumes(seed=11):
    """Uneven non-cubic volumes (odd and even axes, nz across word and
    512-chunk boundaries) plus a few batched-solids cubes."""
    from voxel_ma_amd import synth
    rng = np.random.default_rng(seed)
    vols = []
    for shape in [(7, 9, 65), (33, 40, 70), (20, 3, 130), (1, 17, 5), (12, 12, 520),
                  (41, 2, 1), (16, 31, 64)]:
        p = rng.choice([0.55, 0.8, 0.95])
        vols.append((rng.random(shape) < p).astype(np.uint8))
    for L, prims in synth.batch_c4_specs(count=5, seed=21, lmin=24, ratio=2.5):
        vols.append(synth.solids_host(L, prims))
    return vols


def _levels_for(occ):
    from voxel_ma_amd import synth
    size = synth.inside_size(occ)
    lv = synth.levels_from_tt(size)
    if lv[-1] <= 0.0:
        lv = [0.7, 1.5, 3.0]
    return lv


@pytest.fixture(scope="module")
def ragged(oracle):
    vols = _ragged_volumes()
    levels = np.array([_levels_for(o) for o in vols], np.float32)
    ref = [_oracle_all(oracle, o, list(map(float, lv))) for o, lv in zip(vols, levels)]
    return vols, levels, ref


def _check_members(members, vols, levels, ref, name):
    for i, (m, occ, r) in enumerate(zip(members, vols, ref)):
        d, f, lvl, lab, _ = r
        assert np.array_equal(m.get(V.BUF_EDT), d), (name, i, "edt")
        assert np.array_equal(m.get(V.BUF_FT), f), (name, i, "ft")
        assert np.array_equal(m.get(V.BUF_LEVEL), lvl), (name, i, "level")
        assert np.array_equal(m.get(V.BUF_LABEL), lab), (name, i, "label")


def test_batch_matches_oracle_per_volume(vc, ragged):
    vols, levels, ref = ragged
    say("batch default schedule", len(vols), "volumes")
    members = [vc.upload(o) for o in vols]
    b = V.Batch(vc, members)
    st = b.run(levels)
    _check_members(members, vols, levels, ref, "default")
    # stats sum over the volumes; passes follow the slowest volume
    for k in range(levels.shape[1]):
        assert st.removed[k] == sum(r[4].removed[k] for r in ref)
    assert st.passes == 8 * sum(max(r[4].sweeps[k] for r in ref)
                                for k in range(levels.shape[1]))
    b.release()
    for m in members:
        m.release()


@pytest.mark.parametrize("sched", [
    dict(edt_lds_kb=16, edt_grid_per_cu=0, thin_grid_per_cu=0, sweeps_per_sync=1, use_graph=0,
         tail_sweep=-1, tile_sweep=-1),
    dict(edt_lds_kb=64, edt_grid_per_cu=1, thin_grid_per_cu=1, sweeps_per_sync=32, use_graph=1,
         tail_sweep=0, tile_sweep=-1),
    dict(edt_lds_kb=32, edt_grid_per_cu=8, thin_grid_per_cu=8, sweeps_per_sync=0, use_graph=1,
         tail_sweep=2, tile_sweep=-1),
    dict(edt_lds_kb=8, edt_grid_per_cu=2, thin_grid_per_cu=0, sweeps_per_sync=3, use_graph=1,
         tail_sweep=1, tile_sweep=-1),
    dict(edt_lds_kb=32, edt_grid_per_cu=0, thin_grid_per_cu=0, sweeps_per_sync=0, use_graph=0,
         tail_sweep=3, tile_sweep=-1),
    dict(edt_lds_kb=32, edt_grid_per_cu=0, thin_grid_per_cu=0, sweeps_per_sync=0, use_graph=0,
         tail_sweep=-1, tile_sweep=0),
    dict(edt_lds_kb=32, edt_grid_per_cu=0, thin_grid_per_cu=2, sweeps_per_sync=1, use_graph=0,
         tail_sweep=1, tile_sweep=2),
    dict(edt_lds_kb=16, edt_grid_per_cu=0, thin_grid_per_cu=0, sweeps_per_sync=0, use_graph=1,
         tail_sweep=-1, tile_sweep=4),
    dict(edt_lds_kb=32, edt_grid_per_cu=1, thin_grid_per_cu=1, sweeps_per_sync=2, use_graph=1,
         tail_sweep=-1, tile_sweep=0),
    dict(edt_lds_kb=16, edt_grid_per_cu=0, thin_grid_per_cu=0, sweeps_per_sync=-1, use_graph=1,
         tail_sweep=-1, tile_sweep=0),
    dict(edt_lds_kb=16, edt_grid_per_cu=0, thin_grid_per_cu=0, sweeps_per_sync=-1, use_graph=0,
         tail_sweep=2, tile_sweep=-1),
])
def test_batch_every_schedule_same_result(vc, ragged, sched):
    vols, levels, ref = ragged
    say("batch schedule", sched)
    members = [vc.upload(o) for o in vols]
    b = V.Batch(vc, members)
    b.set_schedule(**sched)
    assert b.schedule().as_dict() == sched
    b.run(levels)
    _check_members(members, vols, levels, ref, str(sched))
    b.release()
    for m in members:
        m.release()


def test_batch_graph_replays(vc, ragged):
    """A captured segment is replayed many times, with eager work (another
    volume's thinning, host syncs) between calls; every replay must equal the
    eager result."""
    vols, levels, ref = ragged
    members = [vc.upload(o) for o in vols]
    b = V.Batch(vc, members)
    b.set_schedule(use_graph=1, sweeps_per_sync=2)
    other = vc.upload(vols[1])
    other.edt().measure()
    for rep in range(4):
        say("graph replay", rep)
        b.run(levels)
        _check_members(members, vols, levels, ref, "replay%d" % rep)
        other.thin_levels(list(map(float, levels[1])))
        vc.synchronize()
    other.release()
    b.release()
    for m in members:
        m.release()


@pytest.mark.parametrize("tail", [0, 2])
def test_sparse_pass_variant_single_volumes(vc, oracle, tail):
    """The sparse-pass variant (one wavefront per 256 words, flagged words
    compacted in LDS) on single volumes of every shape class: labels equal the
    oracle."""
    rng = np.random.default_rng(31 + tail)
    shapes = [tuple(int(s) for s in rng.integers(2, 40, size=3)) for _ in range(12)]
    shapes += [(64, 64, 64), (3, 70, 129)]
    for i, shape in enumerate(shapes):
        occ = (rng.random(shape) < rng.choice([0.6, 0.85, 0.97])).astype(np.uint8)
        if shape == (64, 64, 64):
            from voxel_ma_amd import synth
            occ = synth.torus_c1()
        levels = sorted(float(x) for x in rng.uniform(0.8, 6.0, size=3))
        say("sparse variant", tail, i, shape)
        m = vc.upload(occ)
        b = V.Batch(vc, [m])
        b.set_schedule(tail_sweep=tail, tile_sweep=-1)
        st = b.run(levels)
        d, f, lvl, lab, sto = _oracle_all(oracle, occ, levels)
        assert np.array_equal(m.get(V.BUF_LABEL), lab), (tail, i, shape)
        assert st.passes == 8 * sum(sto.sweeps[k] for k in range(3))
        b.release()
        m.release()


def test_batch_of_one_equals_single_path(vc, oracle):
    from voxel_ma_amd import synth
    occ = synth.torus_c1()
    levels = [1.5, 3.0, 6.0]
    say("batch of one")
    a = vc.upload(occ)
    a.edt().measure()
    sa = a.thin_levels(levels)
    m = vc.upload(occ)
    b = V.Batch(vc, [m])
    b.set_schedule(use_graph=1, thin_grid_per_cu=4)
    sb = b.run(levels)
    assert np.array_equal(a.get(V.BUF_LABEL), m.get(V.BUF_LABEL))
    assert list(sa.removed[:3]) == list(sb.removed[:3])
    assert sa.passes == sb.passes and sa.tests == sb.tests
    b.release()
    a.release()
    m.release()


def test_batch_float_measure_members(vc, oracle):
    """Members keep their own measure: a float measure set on one volume is
    used by the batch thinning of that volume only."""
    from voxel_ma_amd import synth
    rng = np.random.default_rng(5)
    occ_a = synth.torus_c1()
    occ_b = (rng.random((20, 24, 28)) < 0.85).astype(np.uint8)
    meas = rng.uniform(0, 5, size=occ_a.shape).astype(np.float32)
    levels = np.array([[1.0, 2.5, 4.0], [1.2, 2.0, 3.5]], np.float32)
    say("batch float measure member")
    a, bv = vc.upload(occ_a), vc.upload(occ_b)
    b = V.Batch(vc, [a, bv])
    b.edt().measure()
    buf = V.DeviceBuffer.from_array(meas)
    a.set_measure_f32(buf.ptr)
    b.thin(levels)
    lvl_a = np.full(occ_a.shape, 255, np.uint8)
    for k in reversed(range(3)):
        lvl_a[(occ_a > 0) & (meas < levels[0, k])] = k
    lab_a, _ = oracle.thin(occ_a, lvl_a, 3)
    assert np.array_equal(a.get(V.BUF_LABEL), lab_a)
    _, _, _, lab_b, _ = _oracle_all(oracle, occ_b, list(map(float, levels[1])))
    assert np.array_equal(bv.get(V.BUF_LABEL), lab_b)
    buf.free()
    b.release()
    a.release()
    bv.release()


def test_batch_tune_installs_valid_schedule(vc, ragged):
    vols, levels, ref = ragged
    say("batch tune")
    members = [vc.upload(o) for o in vols]
    b = V.Batch(vc, members)
    b.edt().measure()
    sc, (t_front, t_thin) = b.tune(levels, reps=1)
    assert sc.edt_lds_kb in (8, 16, 32, 64) and t_front > 0 and t_thin > 0
    assert b.schedule().as_dict() == sc.as_dict()
    _check_members(members, vols, levels, ref, "tuned")
    # same shape key: the cached schedule is reused
    b2 = V.Batch(vc, members)
    assert b2.schedule().as_dict() == sc.as_dict()
    b2.release()
    b.release()
    for m in members:
        m.release()


def test_batch_errors(vc):
    occ = np.ones((6, 6, 6), np.uint8)
    a = vc.upload(occ)
    with pytest.raises(V.VcError) as e:
        V.Batch(vc, [a, a])
    assert e.value.code == V.VC_ERR_INVALID
    b = V.Batch(vc, [a])
    with pytest.raises(V.VcError) as e:
        b.thin([1.0, 2.0])
    assert e.value.code == V.VC_ERR_STATE
    b.edt().measure()
    with pytest.raises(V.VcError) as e:
        b.thin([2.0, 1.0])
    assert e.value.code == V.VC_ERR_INVALID
    with pytest.raises(V.VcError) as e:
        b.set_schedule(sweeps_per_sync=33)
    assert e.value.code == V.VC_ERR_INVALID
    with pytest.raises(V.VcError) as e:
        b.set_schedule(edt_lds_kb=48)
    assert e.value.code == V.VC_ERR_INVALID
    b.release()
    a.release()


def test_solids_generator_matches_host(vc):
    from voxel_ma_amd import synth
    for i, (L, prims) in enumerate(synth.batch_c4_specs(count=4, seed=8, lmin=32, ratio=2.0)):
        say("solids", i, L, len(prims))
        buf = vc.synth_solids((L, L, L), prims)
        got = buf.download(np.uint8, (L, L, L))
        buf.free()
        assert np.array_equal(got, synth.solids_host(L, prims)), i


@pytest.mark.parametrize("tile", [0, 1, 4])
def test_tile_sweep_variant_single_volumes(vc, oracle, tile):
    """Tile sweeps (32 x 32 x 64 tiles, all 8 passes in LDS with an 8-row
    halo) on single volumes whose axes cross tile and word borders, from the
    first sweep or after dense sweeps: labels equal the oracle, the sweep
    count too."""
    rng = np.random.default_rng(71 + tile)
    shapes = [tuple(int(s) for s in rng.integers(2, 40, size=3)) for _ in range(6)]
    shapes += [(64, 64, 64), (70, 33, 129), (33, 97, 64), (96, 40, 200)]
    for i, shape in enumerate(shapes):
        occ = (rng.random(shape) < rng.choice([0.6, 0.85, 0.97])).astype(np.uint8)
        if shape == (64, 64, 64):
            from voxel_ma_amd import synth
            occ = synth.torus_c1()
        levels = sorted(float(x) for x in rng.uniform(0.8, 6.0, size=3))
        say("tile sweeps", tile, i, shape)
        m = vc.upload(occ)
        b = V.Batch(vc, [m])
        b.set_schedule(tile_sweep=tile, sweeps_per_sync=1 + i % 3)
        st = b.run(levels)
        d, f, lvl, lab, sto = _oracle_all(oracle, occ, levels)
        assert np.array_equal(m.get(V.BUF_LABEL), lab), (tile, i, shape)
        assert list(st.removed[:3]) == [int(x) for x in sto.removed[:3]], (tile, i, shape)
        assert st.passes == 8 * sum(sto.sweeps[k] for k in range(3))
        b.release()
        m.release()


def test_tile_sweeps_blobs_match_dense(vc):
    """A 160 x 150 x 140 volume of blurred noise (many tiles, long removal
    chains): tile sweeps from sweep 0 and from sweep 3 give the labels of the
    dense passes."""
    occ, _ = vc.synth_blurred_noise((160, 150, 140), sigma=4.0, fill=0.35, seed=5)
    host = occ.download(np.uint8, (160, 150, 140))
    occ.free()
    from voxel_ma_amd import synth
    levels = synth.levels_from_tt(synth.inside_size(host))
    out = []
    for tile in (-1, 0, 3):
        say("blobs tile", tile)
        m = vc.upload(host)
        b = V.Batch(vc, [m])
        b.set_schedule(tile_sweep=tile)
        st = b.run(levels)
        out.append((m.get(V.BUF_LABEL), list(st.removed[:3]), st.passes))
        b.release()
        m.release()
    for lab, rem, passes in out[1:]:
        assert np.array_equal(lab, out[0][0])
        assert rem == out[0][1] and passes == out[0][2]


def test_sweep_timing_and_visited_voxels(vc):
    """The tile-sweep timing hook brackets each tile-sweep launch with HIP
    events without changing any result, and the visited-voxel counter stays
    within the dense bound (tiles of 32 x 32 x 64 voxels, 8 passes each; word
    passes 256 voxels per visited word)."""
    shape = (128, 96, 128)   # 4 x 3 x 2 whole tiles
    occ, _ = vc.synth_blurred_noise(shape, sigma=4.0, fill=0.35, seed=6)
    host = occ.download(np.uint8, shape)
    occ.free()
    from voxel_ma_amd import synth
    levels = synth.levels_from_tt(synth.inside_size(host))
    N = host.size
    runs = {}
    for name, sched, timed in (("tile", dict(tile_sweep=0), False),
                               ("tile-timed", dict(tile_sweep=0), True),
                               ("tile-graph-timed", dict(tile_sweep=0, use_graph=1,
                                                         sweeps_per_sync=4), True),
                               ("dense", dict(tile_sweep=-1, tail_sweep=-1), True)):
        say("sweep timing", name)
        m = vc.upload(host)
        b = V.Batch(vc, [m])
        b.set_schedule(**sched)
        vc.thin_timing(timed)
        st = b.run(levels)
        launches, ms = vc.thin_timing_read()
        vc.thin_timing(False)
        runs[name] = (m.get(V.BUF_LABEL), list(st.removed[:3]), st.passes,
                      st.voxel_passes)
        sweeps = st.passes // 8
        assert 0 < st.voxel_passes <= 8 * N * sweeps, name
        if name.startswith("tile"):
            assert st.voxel_passes % (8 * 32 * 32 * 64) == 0, name
            if timed:
                assert sweeps <= launches <= st.launches and ms > 0.0, name
        else:
            assert st.voxel_passes % 256 == 0
            assert launches == 0 and ms == 0.0   # no tile sweeps ran
        if not timed:
            assert launches == 0 and ms == 0.0
        b.release()
        m.release()
    ref = runs["tile"]
    for name, r in runs.items():
        assert np.array_equal(r[0], ref[0]), name
        assert r[1] == ref[1] and r[2] == ref[2], name
    # the tile schedules visit the same tiles whether timed or not
    assert runs["tile-timed"][3] == ref[3]
    assert runs["tile-graph-timed"][3] == ref[3]

"""Benchmark: Mvoxels/s of the plain GPU path (exact EDT + measure + thinning)
on the 512^3 blurred-noise volume, one MI355X.

A step = vc_edt + vc_measure + vc_thin_levels over the whole volume, inputs
resident in HBM.  Prints one JSON line (rank 0).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 512]
                    [--no-cpu-baseline] [--no-batch] [--batch-only]
                    [--profile-steps]

The JSON line also carries "batch": the batched-solids workload (256 cubic
volumes of edge 32-192, one launch per pass over all of them, autotuned
schedule), timed the same way, and "surrogate": the same 512^3 step with the
learned measure (gate at 1% predicted relative error, exact fallback), its
error against the exact measure and the Euler characteristics of both
skeletons.
"""
import argparse
import ctypes
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
EDT_BYTES_PER_VOXEL = 21  # passes z, y, x: (1+2) + (2+4) + (4+4+4)
# dense subfield pass over bit planes: alive state (1/8 B) read with its
# neighbour rows amortised (x 9/8) + the eligibility plane (1/8 B) / 8
THIN_BYTES_PER_VOXEL_PASS = (1.0 / 8) * (9.0 / 8) + (1.0 / 8) / 8
# a tile sweep does the 8 subfield passes of one sweep in one launch
THIN_BYTES_PER_VOXEL_SWEEP = 8 * THIN_BYTES_PER_VOXEL_PASS


class HipEvents:
    """hipEvent timing on the stream the library launches on (the HIP runtime
    libvoxelcore.so links)."""

    def __init__(self):
        import voxel_ma_amd
        self.h = voxel_ma_amd.hip()
        self.h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                               ctypes.c_void_p]
        self.h.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def create(self):
        e = ctypes.c_void_p()
        assert self.h.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        out = ctypes.c_float()
        assert self.h.hipEventElapsedTime(ctypes.byref(out), a, b) == 0
        return out.value


# `traffic` is not counted in the bench run itself: hardware counters need a
# profiler pass of their own, so it is read from the committed summary of
# such passes over this same workload
TRAFFIC_SOURCE = ("profiles/traffic.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, "
                  "one run each over this workload, scaled by a copy kernel's reading")


def measured_traffic(kernel, n):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (512^3 workload only), else None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if n != 512 or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f).get("traffic_bytes_per_dispatch", {})
    v = t.get(kernel)
    return None if v is None else round(v)


def _host_cpu():
    """Host CPU description from lscpu (model, sockets, cores per socket,
    threads per core), /proc/cpuinfo as a fallback."""
    import subprocess
    info = {"nproc_visible": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                             text=True, timeout=10).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets",
                "Core(s) per socket": "cores_per_socket", "Thread(s) per core": "threads_per_core",
                "CPU(s)": "cpus", "CPU max MHz": "max_mhz"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    if "model" not in info:
        try:
            with open("/proc/cpuinfo") as fh:
                for line in fh:
                    if line.startswith("model name"):
                        info["model"] = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    return info


def _native_oracle():
    """Builds the oracle with -O3 -march=native on this machine (the
    Makefile's `native` target, into a fresh temporary directory, so no
    build made for another CPU is picked up); returns (path, flags), or the
    portable build and its flags if the native build fails."""
    import subprocess
    import tempfile
    odir = os.path.join(ROOT, "oracle")
    out = tempfile.mkdtemp(prefix="vc_oracle_native_")
    r = subprocess.run(["make", "-B", "-s", "-C", odir, "native", "OUT=" + out],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    flags = subprocess.run(["make", "-s", "-C", odir, "native-flags"], stdout=subprocess.PIPE,
                           stderr=subprocess.DEVNULL, text=True).stdout.strip()
    path = os.path.join(out, "native", "liboracle.so")
    if r.returncode == 0 and os.path.exists(path):
        return path, flags
    print("[bench] native oracle build failed, using the portable one:\n" + r.stdout,
          file=sys.stderr, flush=True)
    return os.path.join(odir, "_build", "liboracle.so"), "-O2 -march=x86-64-v2 -fopenmp"


def cpu_baseline(occ_host, levels, omp_runs=5):
    """The oracle (the C restatement, built -O3 -march=native on this host)
    on the full volume: EDT + measure + levels + thinning.  The OpenMP leg on
    all host threads is the median of `omp_runs` runs after one warm-up; the
    sequential leg (the reference's single-threaded design) runs once."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_lib as O

    path, flags = _native_oracle()
    O.LIB_PATH = path
    O._lib = None
    O.lib()
    if "vc_oracle_native_" in path:
        import shutil
        shutil.rmtree(os.path.dirname(os.path.dirname(path)), ignore_errors=True)
    occ = np.ascontiguousarray(occ_host)
    n_vox = occ.size

    def once():
        t0 = time.perf_counter()
        d, f = O.edt_ft(occ)
        del d
        lam = O.lambda2(occ, f)
        del f
        lvl = O.levels(occ, lam, levels)
        del lam
        O.thin(occ, lvl, len(levels))
        return time.perf_counter() - t0

    threads = O.set_threads(0)
    print("[bench] cpu baseline: oracle on %d threads, %d runs + warm-up"
          % (threads, omp_runs), file=sys.stderr, flush=True)
    once()  # warm-up; also fills the oracle's simple-point memo
    ts = sorted(once() for _ in range(omp_runs))
    med = ts[len(ts) // 2]
    O.set_threads(1)
    print("[bench] cpu baseline: oracle on 1 thread", file=sys.stderr, flush=True)
    t1 = once()
    O.set_threads(threads)
    shape = "x".join(str(s) for s in occ.shape)
    # the boundary row (sites, quads, Euler terms) on the central 128^3 crop
    n = occ.shape[0]
    side = min(n, 128)
    c0 = (n - side) // 2
    crop = np.ascontiguousarray(occ[c0:c0 + side, c0:c0 + side, c0:c0 + side])
    t0 = time.perf_counter()
    breps = 0
    while True:
        O.boundary(crop)
        breps += 1
        bel = time.perf_counter() - t0
        if bel > 3.0 or breps >= 20:
            break
    # the discrete Voronoi diagram of the feature transform (numpy oracle) on
    # a 64^3 crop
    import voro_ref as VR
    vside = min(n, 64)
    v0 = (n - vside) // 2
    vcrop = np.ascontiguousarray(occ[v0:v0 + vside, v0:v0 + vside, v0:v0 + vside])
    _, vft = O.edt_ft(vcrop)
    t0 = time.perf_counter()
    vreps = 0
    while True:
        VR.voronoi_from_ft(vcrop, vft)
        vreps += 1
        vel = time.perf_counter() - t0
        if vel > 3.0 or vreps >= 10:
            break
    return {"value": round(n_vox / med / 1e6, 3), "unit": "Mvoxels/s", "cores": threads,
            "kind": "port",
            "sample": "oracle (C restatement, %s) EDT+measure+levels+thinning on the full %s "
                      "volume with the same levels, OpenMP on %d threads: median of %d runs "
                      "(%s s) after one warm-up" % (flags, shape, threads, omp_runs,
                                                    ", ".join("%.2f" % t for t in ts), ),
            "single_thread": {"value": round(n_vox / t1 / 1e6, 3), "unit": "Mvoxels/s",
                              "cores": 1, "seconds": round(t1, 2),
                              "sample": "the same oracle on 1 thread (the reference's "
                                        "single-threaded design), the full %s volume, "
                                        "one run" % shape},
            "build": flags,
            "host": _host_cpu(),
            "rows": {"R3_R4_boundary": {"value": round(crop.size * breps / bel / 1e6, 3),
                                        "unit": "Mvoxels/s", "sample": "oracle boundary + "
                                        "Euler (sequential) on the central %d^3 crop, %d reps"
                                        % (side, breps)},
                     "voronoi_from_ft": {"value": round(vcrop.size * vreps / vel / 1e6, 3),
                                         "unit": "Mvoxels/s",
                                         "sample": "oracle (numpy) discrete Voronoi diagram "
                                                   "on the central %d^3 crop, %d reps"
                                                   % (vside, vreps)}}}


def batch_leg(ctx, ev, steps, warmup, count=256, seed=4):
    """The batched-solids workload: `count` cubic volumes, edge
    round(32 * 6^u), unions of 4-16 capsules/ellipsoids, generated on the
    device; per-volume levels from each volume's inside size.  Autotunes the
    schedule, then times whole-batch EDT + measure + thinning steps."""
    import numpy as np
    import voxel_ma_amd
    from voxel_ma_amd import synth

    specs = synth.batch_c4_specs(count=count, seed=seed)
    vols, levels = [], []
    h = hashlib.sha256()
    for L, prims in specs:
        v = ctx.solids_volume((L, L, L), prims)
        occ = v.get(voxel_ma_amd.BUF_OCC)
        h.update(occ.tobytes())
        levels.append(synth.levels_from_tt(synth.inside_size(occ)))
        vols.append(v)
    levels = np.array(levels, np.float32)
    b = voxel_ma_amd.Batch(ctx, vols)
    n, voxels = b.size()
    print("[bench] batch of %d volumes, %d voxels: autotune" % (n, voxels), file=sys.stderr,
          flush=True)
    b.edt().measure()
    t0 = time.perf_counter()
    sc, tt = b.tune(levels, reps=2)
    tune_s = time.perf_counter() - t0
    stream = ctx.stream()
    for _ in range(warmup):
        b.run(levels, stats=False)
    ctx.synchronize()
    evs = [[ev.create() for _ in range(4)] for _ in range(steps)]
    t0 = time.perf_counter()
    st = None
    for k in range(steps):
        ev.record(evs[k][0], stream)
        b.edt()
        ev.record(evs[k][1], stream)
        b.measure()
        ev.record(evs[k][2], stream)
        st = b.thin(levels, stats=(k == steps - 1))
        ev.record(evs[k][3], stream)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    edt_ms = sum(ev.ms(e[0], e[1]) for e in evs) / steps
    meas_ms = sum(ev.ms(e[1], e[2]) for e in evs) / steps
    thin_ms = sum(ev.ms(e[2], e[3]) for e in evs) / steps
    ms_step = wall * 1e3 / steps
    out = {
        "workload": "%d cubic volumes, edge round(32*6^u) in [%d, %d], unions of 4-16 "
                    "capsules/ellipsoids (seed %d), per-volume -tt levels 0.01/0.02/0.04"
                    % (n, min(L for L, _ in specs), max(L for L, _ in specs), seed),
        "volumes": n, "voxels": int(voxels), "volumes_sha256_16": h.hexdigest()[:16],
        "value": round(voxels / (ms_step * 1e-3) / 1e6, 2), "unit": "Mvoxels/s",
        "ms_per_step": round(ms_step, 3),
        "stages_ms": {"edt": round(edt_ms, 3), "measure": round(meas_ms, 3),
                      "thin": round(thin_ms, 3)},
        "edt_achieved_gbs": round(EDT_BYTES_PER_VOXEL * voxels / (edt_ms * 1e-3) / 1e9, 1),
        "thin_stats": {"removed": [int(x) for x in st.removed[:levels.shape[1]]],
                       "passes_with_work": int(st.passes), "tests": int(st.tests),
                       "voxel_passes": int(st.voxel_passes),
                       "launches": int(st.launches)},
        "schedule": sc.as_dict(), "tune_s": round(tune_s, 2),
        "tuned_ms": {"edt_measure": round(tt[0], 3), "thin": round(tt[1], 3)},
    }
    b.release()
    for v in vols:
        v.release()
    return out


def _grid_complex(m):
    """An m x m grid of unit quads in the z = 0 plane (vertices (i, j, 0)),
    each quad a face whose two sites sit at (i+.5, j+.5, +-h) with h varying
    smoothly, so lambda = 2h varies over the complex."""
    import numpy as np
    i, j = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    v = lambda a, b: (a * (m + 1) + b).ravel()
    faces = np.stack([v(i, j), v(i + 1, j), v(i + 1, j + 1), v(i, j + 1)], 1).astype(np.int32)
    gi, gj = np.meshgrid(np.arange(m + 1), np.arange(m + 1), indexing="ij")
    vts = np.stack([gi.ravel(), gj.ravel(), np.zeros(gi.size)], 1).astype(np.float32)
    h = (1.0 + 0.5 * np.sin(i / 7.0) * np.cos(j / 5.0)).ravel().astype(np.float32)
    c = np.stack([i.ravel() + 0.5, j.ravel() + 0.5], 1).astype(np.float32)
    sites = np.concatenate([np.column_stack([c, h]), np.column_stack([c, -h])]).astype(np.float32)
    nf = m * m
    face_sites = np.stack([np.arange(nf), np.arange(nf) + nf], 1).astype(np.int32)
    return vts, faces, sites, face_sites


def _csr(n_rows, rows, cols):
    """CSR of (row, col) pairs, cols in ascending order within a row."""
    import numpy as np
    order = np.lexsort((cols, rows))
    off = np.zeros(n_rows + 1, np.int64)
    np.add.at(off, rows + 1, 1)
    return np.cumsum(off), np.ascontiguousarray(cols[order], np.int32)


def _random_octree(R, seed):
    """A .sof stream of resolution R: internal nodes split with probability
    falling with depth, leaves carry random corner signs, empty nodes a
    random sign (pre-order, children (i, j, k) with k fastest)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = bytearray(np.int32(R).tobytes())

    def node(ln, depth):
        if ln > 1 and rng.random() < (0.95 if depth < 3 else 0.45):
            out.append(0)
            for _ in range(8):
                node(ln // 2, depth + 1)
        elif ln == 1 or rng.random() < 0.5:
            out.append(2)
            out.append(int(rng.integers(1, 255)))
        else:
            out.append(1)
            out.append(int(rng.integers(0, 2)))
    node(R, 0)
    return bytes(out)


def boundary_row(ctx, occ_host, reps=3):
    """R3/R4: boundary sites, quads and cube-complex Euler terms of the
    headline volume (computed once per volume: a fresh upload per
    repetition, the size query computes all), best of `reps`."""
    def best_of(fn, k=reps):
        ts = []
        for _ in range(k):
            ctx.synchronize()
            t0 = time.perf_counter()
            r = fn()
            ctx.synchronize()
            ts.append(time.perf_counter() - t0)
        return min(ts), r

    n = occ_host.shape[0]
    vols = [ctx.upload(occ_host) for _ in range(reps)]
    it = iter(vols)

    def bnd():
        v = next(it)
        L = ctx._lib
        ns, nq = ctypes.c_int64(), ctypes.c_int64()
        L.vc_boundary_extract(ctx.handle, v.handle, None, 0, ctypes.byref(ns), None, 0,
                              ctypes.byref(nq), None)
        return ns.value, nq.value
    t_b, (ns, nq) = best_of(bnd)
    eu = vols[0].boundary(with_euler=True)[2]
    for v in vols:
        v.release()
    N = occ_host.size
    algo_b = N * 1 + ns * 12 + nq * 16
    return {
        "workload": "boundary sites + quads + Euler terms of the %d^3 headline volume" % n,
        "ms": round(t_b * 1e3, 3), "sites": ns, "quads": nq, "euler": eu["euler"],
        "chi": eu["chi"], "value": round(N / t_b / 1e6, 1), "unit": "Mvoxels/s",
        "roofline": {"bound": "hbm", "achieved": round(algo_b / t_b / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(algo_b / t_b / 1e9 / HBM_PEAK_GBS, 4),
                     "bytes": "1 B/voxel + 12 B/site + 16 B/quad"}}


def edt_ball_row(ctx, n, reps=3):
    """EDT + FT of an n^3 grid holding a ball of radius n * 250 / 512 about
    the grid centre, timed with HIP events on the library's stream (median of
    `reps` after a warm-up); the same 21 bytes per voxel as the headline."""
    import numpy as np
    g = (np.arange(n) - (n - 1) / 2.0) ** 2
    rad = n * 250.0 / 512.0
    occ = ((g[:, None, None] + g[None, :, None] + g[None, None, :]) <= rad * rad)
    vol = ctx.upload(occ.astype(np.uint8))
    del occ
    ev = HipEvents()
    stream = ctx.stream()
    vol.edt()
    ctx.synchronize()
    ts = []
    for _ in range(reps):
        a, b = ev.create(), ev.create()
        ev.record(a, stream)
        vol.edt()
        ev.record(b, stream)
        ts.append(ev.ms(a, b))
    vol.release()
    ms = sorted(ts)[len(ts) // 2]
    gbs = EDT_BYTES_PER_VOXEL * n ** 3 / (ms * 1e-3) / 1e9
    return {"workload": "ball of radius %.0f centred in a %d^3 grid" % (rad, n),
            "ms": round(ms, 3), "value": round(n ** 3 / ms / 1e3, 1), "unit": "Mvoxels/s",
            "achieved_gbs": round(gbs, 1), "frac": round(gbs / 8000.0, 4),
            "note": "column passes search outward until the squared distance passes the "
                    "best cost, so their work grows with the distance to the surface"}


def rows_leg(ctx, occ_host, label_host, reps=3):
    """The other rows of the path, each on its own input, timed on the
    device stream (wall clock around the C-ABI call; the Voronoi-side calls
    take host arrays, so their figures include the PCIe copies)."""
    import numpy as np
    import voxel_ma_amd as V
    from voxel_ma_amd import synth
    out = {}

    def best_of(fn, k=reps):
        ts = []
        for _ in range(k):
            ctx.synchronize()
            t0 = time.perf_counter()
            r = fn()
            ctx.synchronize()
            ts.append(time.perf_counter() - t0)
        return min(ts), r

    out["R3_R4_boundary"] = boundary_row(ctx, occ_host, reps)
    n = occ_host.shape[0]

    # R5 on a thick solid: a ball of radius 250 in the 512^3 grid (half the
    # volume inside, mean depth ~62 voxels), the worst case of the column
    # passes' outward searches, whose length follows the distance
    out["R5_edt_thick_ball"] = edt_ball_row(ctx, n, reps)

    # R2: dense occupancy from a .sof octree (R = 256)
    stream = _random_octree(256, 9)
    tree = V.Octree(stream)
    t_o, vo = best_of(lambda: tree.volume(ctx))
    vo.release()
    n3 = 257 ** 3
    out["R2_octree"] = {"workload": "random .sof octree, R = 256 (%d nodes), to a 257^3 volume"
                        % (tree.info.n_internal + tree.info.n_empty + tree.info.n_leaf),
                        "ms": round(t_o * 1e3, 3), "value": round(n3 / t_o / 1e6, 1),
                        "unit": "Mvoxels/s"}
    tree.free()

    # the voxel-core Voronoi construction (discrete diagram of the feature
    # transform) on the central 256^3 crop; host outputs, so PCIe included
    side = min(n, 256)
    c0 = (n - side) // 2
    crop = np.ascontiguousarray(occ_host[c0:c0 + side, c0:c0 + side, c0:c0 + side])
    vcrop = ctx.upload(crop)
    vcrop.edt()
    t_v, (vv, vf, vfs, vs) = best_of(lambda: V.voro_from_ft(vcrop))
    vcrop.release()
    out["voronoi_from_ft"] = {
        "workload": "discrete Voronoi diagram of the feature transform, central %d^3 crop "
                    "of the headline volume" % side,
        "ms": round(t_v * 1e3, 2), "faces": int(len(vf)), "vertices": int(len(vv)),
        "sites": int(len(vs)), "value": round(crop.size / t_v / 1e6, 1),
        "unit": "Mvoxels/s", "note": "size query + fill (two passes over the volume), "
                                     "outputs copied to host memory"}
    del vv, vf, vfs, vs

    # R6/R7: tagging and lambda measures; R8/R9 extraction and merging;
    # R10-R12 finalize + prune, on a 500 x 500 grid complex
    m = 500
    vts, faces, sites, fs = _grid_complex(m)
    vol = ctx.upload(occ_host)
    pts = (np.random.default_rng(5).random((8 << 20, 3)) * n).astype(np.float32)
    t_t, _ = best_of(lambda: V.voro_tag_points(vol, pts))
    vol.release()
    out["R6_tag"] = {"workload": "%d random points in the headline volume" % len(pts),
                     "ms": round(t_t * 1e3, 3), "value": round(len(pts) / t_t / 1e6, 1),
                     "unit": "Mpoints/s", "note": "host arrays, PCIe copies included"}
    edges, foff, fe = V.cc_finalize(len(vts), [], faces)
    ne, nf = len(edges), len(faces)
    fe_rows = np.repeat(np.arange(nf), np.diff(foff))
    ef_off, ef = _csr(ne, fe.astype(np.int64), fe_rows.astype(np.int32))
    ve_rows = edges.ravel().astype(np.int64)
    ve_cols = np.repeat(np.arange(ne, dtype=np.int32), 2)
    ve_off, ve = _csr(len(vts), ve_rows, ve_cols)
    fvalid = np.ones(nf, np.uint8)
    t_m, (lv, le, lf) = best_of(lambda: V.voro_measures(ctx, len(vts), edges, fs, sites, fvalid,
                                                         ve_off, ve, ef_off, ef))
    out["R7_measures"] = {"workload": "lambda of %d faces, %d edges, %d vertices" % (nf, ne, len(vts)),
                          "ms": round(t_m * 1e3, 3), "value": round(nf / t_m / 1e6, 2),
                          "unit": "Mfaces/s", "note": "host arrays, PCIe copies included"}
    valid = np.ones(len(vts), np.uint8)
    fl = faces
    t_x, xh = best_of(lambda: V.voro_extract_inside(len(vts), edges, fl, valid, le, lf), 1)
    t_xd, xd = best_of(lambda: V.voro_extract_inside_device(ctx, len(vts), edges, fl, valid,
                                                            le, lf))
    # the merge's C-ABI calls alone (arrays prepared once)
    call_h, out_h = V.merge_close_call(vts, edges, fl, valid, 0.25, True, None)
    call_d, out_d = V.merge_close_call(vts, edges, fl, valid, 0.25, True, ctx)
    t_g, _ = best_of(call_h, 1)
    t_gd, _ = best_of(call_d)
    gh, gd = out_h(), out_d()
    out["R8_R9"] = {"workload": "extract inside + merge close vertices on %d faces" % nf,
                    "extract_host_ms": round(t_x * 1e3, 1),
                    "extract_device_ms": round(t_xd * 1e3, 1),
                    "merge_host_ms": round(t_g * 1e3, 1), "merge_device_ms": round(t_gd * 1e3, 1),
                    "merge_on_device": bool(gd[6]),
                    "device_equals_host": all(np.array_equal(np.asarray(a), np.asarray(b))
                                              for a, b in zip(xh, xd)) and
                                          all(gd[k] == gh[k] for k in range(6)),
                    "note": "host arrays in and out; extraction through its wrapper, the merge "
                            "as its C-ABI call alone"}
    t_f, fh = best_of(lambda: V.cc_finalize(len(vts), [], fl), 1)
    t_fd, fd = best_of(lambda: V.cc_finalize(len(vts), [], fl, ctx=ctx))
    t_p, rh = best_of(lambda: V.cc_prune(len(vts), edges, foff, fe, le, lf, 1.5, 1.5), 1)
    t_pd, rd = best_of(lambda: V.cc_prune_device(ctx, len(vts), edges, foff, fe, le, lf,
                                                 1.5, 1.5))
    same = all(np.array_equal(a, b) for a, b in zip(rh[1:], rd[1:5]))
    out["R10_R12_cellcomplex"] = {"workload": "finalize + FIFO prune of %d quads" % nf,
                                  "finalize_ms": round(t_f * 1e3, 1),
                                  "finalize_device_ms": round(t_fd * 1e3, 1),
                                  "finalize_device_equals_host":
                                      all(np.array_equal(a, b) for a, b in zip(fh, fd)),
                                  "prune_ms": round(t_p * 1e3, 1),
                                  "prune_device_ms": round(t_pd * 1e3, 1),
                                  "generations": int(rd[5]), "device_equals_host": same}

    # the voxel core of a 128^3 crop: discrete Voronoi diagram, inside part
    # triangulated, FIFO collapse at t = 0.04 on the host and on the device
    side = min(n, 128)
    c0 = (n - side) // 2
    crop = np.ascontiguousarray(occ_host[c0:c0 + side, c0:c0 + side, c0:c0 + side])
    vc_ = ctx.upload(crop)
    vc_.edt()
    cv, cf, cfs, cs = V.voro_from_ft(vc_)
    cvalid = V.voro_tag_points(vc_, cv)
    vc_.release()
    clam = np.linalg.norm(cs[cfs[:, 0]] - cs[cfs[:, 1]], axis=1).astype(np.float32)
    vi, eo, em_, tr, _, fm_ = V.voro_extract_inside(len(cv), [], cf, cvalid,
                                                    np.zeros(0, np.float32), clam)
    tris = np.asarray(tr, np.int32).reshape(-1, 3)
    ce, cfo, cfe = V.cc_finalize(len(vi), np.asarray(eo, np.int32).reshape(-1, 2), tris)
    cem = np.zeros(len(ce), np.float32)
    cem[:len(em_)] = em_
    cfm = np.asarray(fm_, np.float32)
    thr = float(synth.inside_size(crop)) * 0.04
    t_ch, rch = best_of(lambda: V.cc_prune(len(vi), ce, cfo, cfe, cem, cfm, thr, thr), 1)
    t_cd, rcd = best_of(lambda: V.cc_prune_device(ctx, len(vi), ce, cfo, cfe, cem, cfm,
                                                  thr, thr))
    out["voxel_core_collapse"] = {
        "workload": "inside part of the discrete Voronoi diagram of the central %d^3 crop "
                    "(%d triangles, %d edges), FIFO collapse at t = 0.04" % (side, len(tris),
                                                                            len(ce)),
        "host_ms": round(t_ch * 1e3, 1), "device_ms": round(t_cd * 1e3, 1),
        "generations": int(rcd[5]), "faces_removed": int(rcd[3].sum()),
        "device_equals_host": all(np.array_equal(a, b) for a, b in zip(rch[1:], rcd[1:5])),
        "note": "host arrays in and out; device time includes the adjacency build"}

    out["vol2ma"] = vol2ma_row(crop, side)

    # output files: the skeleton of the headline volume as PLY and .r
    import tempfile
    sk = np.argwhere(label_host == 1).astype(np.float32)
    r = np.ones(len(sk), np.float32)
    with tempfile.TemporaryDirectory() as td:
        t_w, _ = best_of(lambda: V.write_ply(os.path.join(td, "s.ply"), sk), 1)
        sz = os.path.getsize(os.path.join(td, "s.ply"))
        t_r, _ = best_of(lambda: V.write_radii(os.path.join(td, "s.r"), sk, r), 1)
    out["writers"] = {"workload": "PLY + .r of the %d skeleton voxels" % len(sk),
                      "ply_ms": round(t_w * 1e3, 1), "ply_MB": round(sz / 1e6, 1),
                      "radii_ms": round(t_r * 1e3, 1)}
    return out


def vol2ma_row(crop, side, tt="0.04,0.2"):
    """`voroUtility -md=vol2ma` (src/voroUtility.cpp:454-525) on the crop as a
    user runs it: a child process reading the raw volume and writing the full
    inside part (PLY + .r) and each pruned complex (PLY + .r + .ma).  Wall time
    of the process (start-up included), the program's own stage times, and a
    digest of every output file (so runs of different builds can be compared
    byte for byte)."""
    import subprocess
    import tempfile
    cli = os.path.join(ROOT, "voxel_ma_amd", "voroUtility")
    with tempfile.TemporaryDirectory() as td:
        raw = os.path.join(td, "crop.raw")
        crop.tofile(raw)
        outp = os.path.join(td, "core.ply")
        cmd = [cli, "-md=vol2ma", raw, outp, "-dims=%d,%d,%d" % crop.shape, "-tt=" + tt,
               "-fullOrPruned=2", "-outToDotma"]
        t0 = time.perf_counter()
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=600)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            return {"error": r.stdout[-2000:]}
        stages = {}
        for line in r.stdout.splitlines():
            if line.startswith("time -> "):
                k, v = line[len("time -> "):].rsplit(":", 1)
                stages[k.strip()] = float(v.split()[0])
        h = hashlib.sha256()
        files = sorted(f for f in os.listdir(td) if f != "crop.raw")
        sizes = 0
        for f in files:
            with open(os.path.join(td, f), "rb") as fh:
                b = fh.read()
            sizes += len(b)
            h.update(f.encode() + b"\0" + hashlib.sha256(b).digest())
    return {"workload": "voroUtility -md=vol2ma on the central %d^3 crop, -tt=%s, full + pruned "
                        "outputs with .ma" % (side, tt),
            "wall_ms": round(wall * 1e3, 1), "stages_ms": stages, "files": files,
            "bytes": sizes, "outputs_sha256": h.hexdigest()[:16]}


def _betti(occ, chi):
    """(b0, b1, b2) of a voxel set as 26-connected object / 6-connected
    background, with its Euler characteristic chi"""
    import numpy as np
    from scipy import ndimage
    b0 = int(ndimage.label(occ, structure=np.ones((3, 3, 3), np.uint8))[1])
    b2 = int(ndimage.label(np.pad(occ == 0, 1, constant_values=True))[1]) - 1
    return [b0, b0 + b2 - int(chi), b2]


def surrogate_leg(ctx, ev, vol, levels, n, steps, plain_ms, tau=0.01):
    """Surrogate-guided pruning on the same volume: EDT + surrogate measure
    (inference, gate at relative error tau, exact fallback) + thinning,
    against the plain step; quality against the exact measure."""
    import numpy as np
    import voxel_ma_amd
    from voxel_ma_amd import surrogate

    net = voxel_ma_amd.Surrogate(ctx, seed=1, path=surrogate.WEIGHTS)
    stream = ctx.stream()
    N = n ** 3

    def step():
        vol.edt()
        st = net.measure(vol, max_rel_err=tau)
        vol.thin_levels(levels, stats=False)
        return st

    step()
    ctx.synchronize()
    e0, e1 = ev.create(), ev.create()
    ev.record(e0, stream)
    t0 = time.perf_counter()
    infer_each = []
    for _ in range(steps):
        st = step()
        infer_each.append(st.ms_infer)
    ev.record(e1, stream)
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    # quality: surrogate outputs vs the exact measure (float64), decisions,
    # and the skeletons of the exact and the pure-surrogate measure
    lam, rel = net.infer(vol)
    vol.measure()
    lam2 = vol.get(voxel_ma_amd.BUF_LAM2)
    occ = vol.get(voxel_ma_amd.BUF_OCC) > 0
    exact = np.sqrt(lam2.astype(np.float64))
    err = (np.abs(lam.astype(np.float64) - exact) / np.maximum(exact, 1.0))[occ]
    T = np.array(levels, np.float32)
    agree = float((np.searchsorted(T, np.maximum(lam, 0), side="right")[occ] ==
                   np.searchsorted(T, np.sqrt(lam2.astype(np.float32)), side="right")[occ]).mean())
    del lam, rel, exact, lam2
    vol.thin_levels(levels, stats=False)
    sk_e = (vol.get(voxel_ma_amd.BUF_LABEL) == 1).astype(np.uint8)
    net.measure(vol, max_rel_err=3.0e38)
    vol.thin_levels(levels, stats=False)
    sk_s = (vol.get(voxel_ma_amd.BUF_LABEL) == 1).astype(np.uint8)
    iou = float((sk_e & sk_s).sum() / max(int((sk_e | sk_s).sum()), 1))
    eul = []
    for sk in (sk_e, sk_s):
        v2 = ctx.upload(sk)
        eul.append(int(v2.boundary(with_euler=True)[2]["chi"]))
        v2.release()
    eul.append(int(vol.boundary(with_euler=True)[2]["chi"]))
    # Betti numbers of the (26, 6) voxel sets: b0 = 26-components, b2 =
    # 6-components of the background (padded: the outside is one) - 1, b1 =
    # b0 + b2 - chi (chi of the union of closed voxel cubes, computed above)
    betti = [_betti(sk, chi) for sk, chi in ((sk_e, eul[0]), (sk_s, eul[1]))]
    betti.append(_betti(occ, eul[2]))
    infer_ms = float(np.median(infer_each))  # the timed steps' inference times
    tflops = 116e3 * N / (infer_ms * 1e-3) / 1e12
    net.release()
    return {
        "tau_rel_err": tau, "weights": "voxel_ma_amd/data/unet.bin",
        "ms_per_step": round(ms, 3), "plain_ms_per_step": round(plain_ms, 3),
        "speedup_vs_plain": round(plain_ms / ms, 4),
        "accepted": int(st.accepted), "fallback": int(st.fallback),
        "accepted_frac": round(st.accepted / max(st.inside, 1), 6),
        "ms_infer": round(infer_ms, 3), "ms_exact_fallback": round(st.ms_exact, 3),
        "rel_err_vs_exact_fp64": {"median": round(float(np.median(err)), 4),
                                  "p90": round(float(np.quantile(err, 0.9)), 4),
                                  "frac_within_1pct": round(float((err <= 0.01).mean()), 4)},
        "level_decision_agreement": round(agree, 4),
        "skeleton_iou_pure_surrogate": round(iou, 4),
        "euler_chi": {"exact_skeleton": eul[0], "surrogate_skeleton": eul[1], "object": eul[2]},
        "betti_b0_b1_b2": {"exact_skeleton": betti[0], "surrogate_skeleton": betti[1],
                           "object": betti[2]},
        "roofline_infer": {"bound": "mfma", "achieved": round(tflops, 1), "peak": 2500.0,
                           "unit": "TFLOP/s", "frac": round(tflops / 2500.0, 4),
                           "flop_per_voxel": 116000},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched-solids leg (256 volumes)")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batched-solids leg (for profiling)")
    ap.add_argument("--vol2ma-only", action="store_true",
                    help="only the vol2ma row on the central 128^3 crop of the headline volume")
    ap.add_argument("--boundary-only", action="store_true",
                    help="only the boundary row on the headline volume (for profiling)")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the surrogate-guided pruning leg")
    ap.add_argument("--no-rows", action="store_true",
                    help="skip the per-row measurements (boundary, octree, Voronoi side, "
                         "cell complex, writers)")
    ap.add_argument("--no-sweep-events", action="store_true",
                    help="time the tile sweeps by the thinning stage instead of HIP "
                         "events around each launch")
    ap.add_argument("--profile-steps", action="store_true",
                    help="no CPU baseline, no hashing (for rocprofv3 runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 3 device-to-device copies of 1 GiB (HBM counter "
                         "calibration for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)")
    a = ap.parse_args()
    if a.copy_calibration:
        import voxel_ma_amd
        nbytes = 1 << 30
        src, dst = voxel_ma_amd.DeviceBuffer(nbytes), voxel_ma_amd.DeviceBuffer(nbytes)
        h = voxel_ma_amd.hip()
        for _ in range(3):
            assert h.hipMemcpy(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), nbytes, 3) == 0
        h.hipDeviceSynchronize()
        print(json.dumps({"copy_bytes": nbytes, "copies": 3}))
        return
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (--gpus %d requested); "
              "sharding over several GPUs is a design, not a measured path" % a.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import voxel_ma_amd
    from voxel_ma_amd import synth

    n = a.size
    ctx = voxel_ma_amd.Context(0)
    if a.batch_only:
        print(json.dumps({"batch": batch_leg(ctx, HipEvents(), a.steps, a.warmup)}), flush=True)
        ctx.close()
        return
    print("[bench] generating blurred-noise volume %d^3 on the device" % n, file=sys.stderr,
          flush=True)
    occ_dev, _thr = ctx.synth_blurred_noise((n, n, n), sigma=12.0, fill=0.30, seed=3)
    occ_host = occ_dev.download(np.uint8, (n, n, n))
    fill = float(occ_host.mean())
    size = synth.inside_size(occ_host)
    levels = synth.levels_from_tt(size)
    vol_hash = hashlib.sha256(occ_host.tobytes()).hexdigest()[:16]
    if a.boundary_only:
        occ_dev.free()
        print(json.dumps({"R3_R4_boundary": boundary_row(ctx, occ_host)}), flush=True)
        ctx.close()
        return
    if a.vol2ma_only:
        occ_dev.free()
        side = min(n, 128)
        c0 = (n - side) // 2
        crop = np.ascontiguousarray(occ_host[c0:c0 + side, c0:c0 + side, c0:c0 + side])
        ctx.close()
        print(json.dumps({"vol2ma": vol2ma_row(crop, side)}), flush=True)
        return

    vol = ctx.from_device(occ_dev.ptr, (n, n, n))
    ctx.synchronize()
    occ_dev.free()
    stream = ctx.stream()
    ev = HipEvents()
    N = n * n * n

    def step(events=None):
        if events:
            ev.record(events[0], stream)
        vol.edt()
        if events:
            ev.record(events[1], stream)
        vol.measure()
        if events:
            ev.record(events[2], stream)
        st = vol.thin_levels(levels, stats=True)
        if events:
            ev.record(events[3], stream)
        return st

    print("[bench] warmup %d" % a.warmup, file=sys.stderr, flush=True)
    for _ in range(a.warmup):
        step()
    ctx.synchronize()

    evs = [[ev.create() for _ in range(4)] for _ in range(a.steps)]
    ctx.synchronize()
    t0 = time.perf_counter()
    stats = None
    for k in range(a.steps):
        stats = step(evs[k])
    ctx.synchronize()
    wall = time.perf_counter() - t0
    print("[bench] timed %d steps in %.3f s" % (a.steps, wall), file=sys.stderr, flush=True)
    # the thinning kernel's own launch durations: HIP events on the context
    # stream around every tile-sweep launch, over a second timed region of
    # the same steps (the event records add host gaps between launches, so
    # they stay out of the region that gives `value`)
    sweep_launches, sweep_kernel_ms = 0, 0.0
    if not a.no_sweep_events:
        ctx.thin_timing(True)
        for k in range(a.steps):
            step()
        ctx.synchronize()
        sweep_launches, sweep_kernel_ms = ctx.thin_timing_read()
        ctx.thin_timing(False)

    edt_ms = sum(ev.ms(e[0], e[1]) for e in evs) / a.steps
    meas_ms = sum(ev.ms(e[1], e[2]) for e in evs) / a.steps
    thin_ms = sum(ev.ms(e[2], e[3]) for e in evs) / a.steps
    ms_step = wall * 1e3 / a.steps
    value = N / (ms_step * 1e-3) / 1e6

    edt_gbs = EDT_BYTES_PER_VOXEL * N / (edt_ms * 1e-3) / 1e9
    # thinning: one tile-sweep launch runs all 8 subfield passes of a sweep
    # over the tiles whose 3x3x3 tile neighbourhood changed in the sweep
    # before.  Algorithmic bytes of a step = voxels touched, summed over the
    # subfield passes (the library counts them: 8 x the voxels of each visited
    # tile) x the per-pass model; launches past a level's convergence visit
    # nothing.  Achieved = those bytes / the summed launch durations (HIP
    # events), i.e. algorithmic bytes per launch / average launch duration.
    sweeps = max(1, stats.passes // 8)
    thin_bytes_step = THIN_BYTES_PER_VOXEL_PASS * stats.voxel_passes
    launches_per_step = sweep_launches / a.steps if sweep_launches else 0
    if sweep_launches:
        kernel_ms_step = sweep_kernel_ms / a.steps
        avg_launch_us = sweep_kernel_ms * 1e3 / sweep_launches
    else:
        # timing switched off: the thinning stage's time per sweep (level
        # set-ups and host waits included) stands in for the kernel time
        kernel_ms_step = thin_ms
        avg_launch_us = thin_ms * 1e3 / sweeps
    thin_gbs = thin_bytes_step / (kernel_ms_step * 1e-3) / 1e9
    # the same launches credited with a dense sweep over all N voxels each
    dense_gbs = THIN_BYTES_PER_VOXEL_SWEEP * N * sweeps / (kernel_ms_step * 1e-3) / 1e9
    out = {
        "metric": "Mvoxels/s EDT+thinning at 512^3",
        "value": round(value, 2),
        "unit": "Mvoxels/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "blurred-noise %d^3 (sigma=12, 70th-percentile threshold, "
                               "1-voxel shell), exact EDT+FT, lambda measure, subfield "
                               "thinning at %d levels" % (n, len(levels)),
                   "volume_sha256_16": vol_hash, "fill": round(fill, 4),
                   "levels": [round(l, 4) for l in levels], "inside_size": float(size)},
        "stages_ms": {"edt": round(edt_ms, 3), "measure": round(meas_ms, 3),
                      "thin": round(thin_ms, 3)},
        "thin_stats": {"removed": [int(x) for x in stats.removed[:len(levels)]],
                       "passes_with_work": int(stats.passes), "tests": int(stats.tests),
                       "voxel_passes": int(stats.voxel_passes),
                       "launches": int(stats.launches)},
        "roofline": {"bound": "hbm", "kernel": "thin_tile_sweep (8 subfield passes)",
                     "achieved": round(thin_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(thin_gbs / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic("thin_tile_sweep", n),
                     "traffic_source": TRAFFIC_SOURCE,
                     "bytes_per_voxel_pass": round(THIN_BYTES_PER_VOXEL_PASS, 4),
                     "voxel_passes_per_step": int(stats.voxel_passes),
                     "touched_frac": round(stats.voxel_passes / (8.0 * N * sweeps), 4),
                     "dense_equivalent_gbs": round(dense_gbs, 1),
                     "sweeps_per_step": sweeps,
                     "launches_per_step": round(launches_per_step, 2),
                     "avg_launch_us": round(avg_launch_us, 2),
                     "kernel_ms_per_step": round(kernel_ms_step, 3),
                     "timing": "hip events per launch" if sweep_launches
                               else "thinning stage / sweeps"},
        "roofline_edt": {"bound": "hbm", "kernel": "edt (3 passes)", "achieved": round(edt_gbs, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(edt_gbs / HBM_PEAK_GBS, 4), "bytes_per_voxel": 21,
                         "traffic": sum(filter(None, (measured_traffic(k, n) for k in
                                                      ("edt_pass_z", "edt_pass_y",
                                                       "edt_pass_x")))) or None,
                         "traffic_source": TRAFFIC_SOURCE},
    }
    if not (a.no_surrogate or a.profile_steps):
        print("[bench] surrogate-guided pruning", file=sys.stderr, flush=True)
        out["surrogate"] = surrogate_leg(ctx, ev, vol, levels, n, max(1, a.steps // 2),
                                         ms_step)
    label_host = vol.get(voxel_ma_amd.BUF_LABEL)
    vol.release()
    if not (a.no_rows or a.profile_steps):
        print("[bench] other rows", file=sys.stderr, flush=True)
        out["rows"] = rows_leg(ctx, occ_host, label_host)
    if not (a.no_batch or a.profile_steps):
        out["batch"] = batch_leg(ctx, ev, a.steps, a.warmup)
    if not (a.no_cpu_baseline or a.profile_steps):
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(occ_host, levels)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

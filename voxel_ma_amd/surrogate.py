"""Training and evaluation of the measure surrogate (include/vc_surrogate.h).

Training volumes are generated on the device (torus / ball unions in the
style of the torus and spheres workloads, synth.training_specs; optionally
blurred-noise volumes at 128^3), labelled by the plain path (vc_edt +
vc_measure: the target is the exact lambda = sqrt(lam2)), and the network is
trained by the library's own forward / backward kernels with Adam.  The
committed weights (4,000 volumes of 128^3, about one minute on one MI355X)
come from

    python -m voxel_ma_amd.surrogate --steps 2000 --batch 2 --size 128 \
        --lr 1e-3 --seed 5 --blur-frac 0 --out voxel_ma_amd/data/unet.bin
"""
import argparse
import os
import sys
import time

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(PKG, "data", "unet.bin")


def make_volume(ctx, L, rng, blur_frac=0.0):
    """One labelled training volume of edge L: with probability blur_frac a
    blurred-noise volume (sigma 12 voxels, fill 0.30, random seed >= 1000 --
    the benchmark volume uses seed 3), else a torus / ball union
    (synth.training_specs); vc_edt + vc_measure run on it."""
    from voxel_ma_amd import synth
    # no draw when blur_frac is 0: the generator stream (and so the volumes
    # of a seed) is the one the committed weights were trained on
    if blur_frac > 0 and rng.random() < blur_frac:
        buf, _ = ctx.synth_blurred_noise((L, L, L), sigma=12.0, fill=0.30,
                                         seed=int(rng.integers(1000, 1 << 40)))
        vol = ctx.from_device(buf.ptr, (L, L, L))
        ctx.synchronize()
        buf.free()
    else:
        vol = ctx.solids_volume((L, L, L), synth.training_specs(L, rng))
    vol.edt().measure()
    return vol


def train(ctx, net, steps, batch=2, L=128, lr=1e-3, seed=5, fresh_every=1, log=None,
          blur_frac=0.0):
    """`steps` Adam steps, each on `batch` new volumes of edge L (a fresh
    volume pool every `fresh_every` steps; blur_frac of them blurred noise).
    Returns the loss history."""
    rng = np.random.default_rng(seed)
    hist = []
    pool = []
    t0 = time.perf_counter()
    for k in range(steps):
        if k % fresh_every == 0:
            for v in pool:
                v.release()
            pool = [make_volume(ctx, L, rng, blur_frac) for _ in range(batch)]
        lr_k = lr * 0.5 * (1.0 + np.cos(np.pi * k / max(steps, 1)))  # cosine decay
        loss = net.train_step(pool, max(lr_k, lr * 0.02))
        hist.append(loss)
        if log and (k % 20 == 0 or k == steps - 1):
            log("step %d loss %.5f (%.1f s)" % (k, loss, time.perf_counter() - t0))
    for v in pool:
        v.release()
    return hist


def levels_of(occ):
    from voxel_ma_amd import synth
    return synth.levels_from_tt(synth.inside_size(occ))


def evaluate(ctx, net, vols, max_rel_err=0.01):
    """Quality of the surrogate against the exact measure on volumes whose
    vc_edt and vc_measure have run:
      relative error |lambda_s - lambda| / max(lambda, 1), lambda =
      sqrt(lam2) in float64, on inside voxels (median, 90th percentile,
      fraction within 1%);
      accepted / fallback voxels of the gate at max_rel_err;
      decision agreement: fraction of inside voxels whose first eligible
      level (the -tt levels of the volume) is the same under the surrogate
      measure as under the exact one;
      skeleton IoU at the last level between thinning with the surrogate
      measure everywhere and thinning with the exact measure (both exact
      thinnings, so both keep the topology of the object)."""
    import voxel_ma_amd as V
    errs, acc, fb, agree, n_in, inter, union = [], 0, 0, 0, 0, 0, 0
    for v in vols:
        occ = v.get(V.BUF_OCC) > 0
        lam2 = v.get(V.BUF_LAM2)
        levels = levels_of(occ)
        lam, rel = net.infer(v)
        exact = np.sqrt(lam2.astype(np.float64))
        e = np.abs(lam.astype(np.float64) - exact) / np.maximum(exact, 1.0)
        errs.append(e[occ])
        ok = rel[occ] <= max_rel_err
        acc += int(ok.sum())
        fb += int((~ok).sum())
        T = np.array([np.float32(l) for l in levels], np.float32)
        lvl_s = np.searchsorted(T, np.maximum(lam, 0.0), side="right")
        lvl_e = np.searchsorted(T, np.sqrt(lam2.astype(np.float32)), side="right")
        agree += int((lvl_s[occ] == lvl_e[occ]).sum())
        n_in += int(occ.sum())
        # skeletons: exact measure, then surrogate measure everywhere
        v.measure()
        v.thin_levels(levels, stats=False)
        sk_e = v.get(V.BUF_LABEL) == 1
        net.measure(v, max_rel_err=3.0e38)
        v.thin_levels(levels, stats=False)
        sk_s = v.get(V.BUF_LABEL) == 1
        inter += int((sk_e & sk_s).sum())
        union += int((sk_e | sk_s).sum())
        v.measure()  # leave the exact measure installed
    e = np.concatenate(errs) if errs else np.zeros(0)
    return {"median_rel_err": float(np.median(e)) if e.size else 0.0,
            "p90_rel_err": float(np.quantile(e, 0.9)) if e.size else 0.0,
            "frac_within_1pct": float((e <= 0.01).mean()) if e.size else 0.0,
            "accepted": acc, "fallback": fb,
            "decision_agreement": agree / max(n_in, 1),
            "skeleton_iou": inter / max(union, 1)}


def main():
    import voxel_ma_amd as V
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=2)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--blur-frac", type=float, default=0.0,
                    help="fraction of blurred-noise training volumes")
    ap.add_argument("--out", default=WEIGHTS)
    a = ap.parse_args()
    ctx = V.Context(0)
    net = V.Surrogate(ctx, seed=a.seed)
    hist = train(ctx, net, a.steps, a.batch, a.size, a.lr, a.seed,
                 log=lambda m: print("[train]", m, file=sys.stderr, flush=True),
                 blur_frac=a.blur_frac)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    net.save(a.out)
    rng = np.random.default_rng(a.seed + 1000)
    held_solids = [make_volume(ctx, a.size, rng, 0.0) for _ in range(3)]
    held_blur = [make_volume(ctx, a.size, rng, 1.0) for _ in range(3)]
    print({"loss_first": hist[0], "loss_last": hist[-1],
           "held_out_solids": evaluate(ctx, net, held_solids),
           "held_out_blurred_noise": evaluate(ctx, net, held_blur)}, flush=True)


if __name__ == "__main__":
    main()

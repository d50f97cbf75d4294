"""Train the chemistry surrogate on the GPU and write its weights file.

    python -m openstc_amd.train_surrogate --out openstc_amd/data/surrogate_h2air9.npz
    python -m openstc_amd.train_surrogate --workload jet --out openstc_amd/data/surrogate_h2jet9.npz

1. Run a finite-rate workload on the GPU -- the mixing layer (192x96x96,
   FP32 bulk, DPLUR; 400 iterations, every 20th sampled) or the four-block H2
   jet in vitiated air (4 x 256x192x128, SST; its first 60 iterations, every
   3rd sampled) -- and sample interior states (T, p, Y, local dt) of every
   block (training seed) and, separately, from a run with another initial
   perturbation (held-out seed).
2. Label every cell of each snapshot with the plain point-implicit step on
   the GPU (FP64) and draw training states evenly over decades of max|dY|.
3. Train the MLP with the library's kernels (Adam, warm-up + cosine).
4. Calibrate the trusted output range dy_max on half of the held-out run
   (largest value meeting the error target) and report the other half.
5. Save weights, descriptor, training history and a held-out sample.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import _lib as L
from . import cases, surrogate as S
from .runtime import Context


def log(*a):
    print(*a, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--workload", choices=["mixing", "jet"], default="mixing")
    ap.add_argument("--steps", type=int, default=None, help="solver iterations per harvest run")
    ap.add_argument("--every", type=int, default=None)
    ap.add_argument("--per-snapshot", type=int, default=None, help="training states per block and snapshot")
    ap.add_argument("--val-per-snapshot", type=int, default=None)
    ap.add_argument("--cfl", type=float, default=None)
    ap.add_argument("--grid", type=int, nargs=3, default=None)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--keep-val", type=int, default=8192, help="held-out states stored with the weights")
    ap.add_argument("--uniform", action="store_true", help="sample training states uniformly, not by |dY| decade")
    ap.add_argument("--mag-weight", type=float, default=16.0)
    ap.add_argument("--target", type=float, default=8e-4, help="rel_l2 the output gate is calibrated to")
    ap.add_argument("--activation", choices=sorted(S.ACTIVATIONS), default="hswish")
    a = ap.parse_args(argv)
    jet = a.workload == "jet"
    defaults = dict(steps=60, every=3, per_snapshot=25000, val_per_snapshot=5000, cfl=0.2, grid=[256, 192, 128]) \
        if jet else dict(steps=400, every=20, per_snapshot=100000, val_per_snapshot=20000, cfl=0.2, grid=[192, 96, 96])
    for k, v in defaults.items():
        if getattr(a, k) is None:
            setattr(a, k, v)
    mx, my, mz = a.grid
    t0 = time.time()

    def build(seed):
        if jet:
            return cases.jet_crossflow(mx=mx, my=my, mz=mz, cfl=a.cfl, seed=seed)
        return cases.mixing_layer(mx=mx, my=my, mz=mz, cfl=a.cfl, seed=seed)

    train_seed, val_seed = (6, 7) if jet else (3, 4)
    log(f"[train_surrogate] harvest {a.workload} {mx}x{my}x{mz}, {a.steps} iterations, CFL {a.cfl}")
    tr = S.harvest(build(train_seed), a.steps, a.every, a.per_snapshot, seed=a.seed, log=log,
                   stratify=not a.uniform)
    # held-out: another initial perturbation, sampled uniformly (the solver's own distribution)
    va = S.harvest(build(val_seed), a.steps, a.every, a.val_per_snapshot, seed=a.seed + 1, log=log,
                   stratify=False)
    mech, info = cases.load_mechanism(True)
    param = cases.mixture_param(mech, False, True, 4, a.cfl, 1,
                                cases.InflowState(1.0, 1.0, 0, 0, 1e5, 1000.0, 1.0, 0.0, [0.0] * 9))
    ctx = Context(param)
    t, p, y, dt, dy, it = tr
    vt, vp, vy, vdt, vdy, vit = va
    log(f"[train_surrogate] {len(t)} training / {len(vt)} held-out states labelled ({time.time() - t0:.1f}s)")
    desc = S.make_desc(t, p, y, dt, dy, param, info, mag_weight=a.mag_weight,
                       activation=S.ACTIVATIONS[a.activation])
    target = S.standardised_targets(dy, desc)
    sur = S.Surrogate(ctx, desc, max_batch=a.batch, seed=a.seed)
    vdev = [S.DeviceArray(np.ascontiguousarray(v)) for v in (vt, vp, vy, vdt)]
    val = (len(vt), *vdev, vdy)
    hist = S.train(sur, t, p, y, dt, target, a.epochs, a.batch, a.lr, seed=a.seed, val=val, log=log)
    ungated = S.evaluate(sur, *val)
    log("[train_surrogate] held-out, no output gate:", json.dumps(ungated))
    # trusted output range: calibrated on half of the held-out run, reported on the other half
    half = np.random.default_rng(a.seed + 7).permutation(len(vt))
    cal, tst = half[: len(vt) // 2], half[len(vt) // 2:]
    dy_max, table = S.calibrate_dy_max(sur, vt[cal], vp[cal], vy[cal], vdt[cal], vdy[cal], target=a.target)
    log(f"[train_surrogate] dy_max = {dy_max:.3e}")
    for thr, rel, fb in table:
        log(f"[calibrate] dy_max {thr:.2e} rel_l2 {rel:.3e} fallback {fb:.4f}")
    w = sur.weights()
    sur.close()
    desc.dy_max = dy_max
    sur = S.Surrogate(ctx, desc, w)
    tdev = [S.DeviceArray(np.ascontiguousarray(v[tst])) for v in (vt, vp, vy, vdt)]
    final = S.evaluate(sur, len(tst), *tdev, vdy[tst])
    log("[train_surrogate] held-out test half, gated:", json.dumps(final))
    for d in tdev:
        d.free()
    keep = tst[np.random.default_rng(a.seed).choice(len(tst), size=min(a.keep_val, len(tst)), replace=False)]
    sur.save(a.out, history=np.array(json.dumps(hist)), heldout_t=vt[keep], heldout_p=vp[keep],
             heldout_y=vy[keep], heldout_dt=vdt[keep], heldout_dy=vdy[keep],
             info=np.array(json.dumps({"workload": a.workload, "grid": a.grid, "cfl": a.cfl, "steps": a.steps,
                                       "every": a.every,
                                       "n_train": int(len(t)), "n_heldout": int(len(vt)), "epochs": a.epochs,
                                       "batch": a.batch, "lr": a.lr, "mag_weight": a.mag_weight,
                                       "activation": a.activation,
                                       "heldout_ungated": ungated, "dy_max": dy_max, "target": a.target,
                                       "heldout_test_gated": final})))
    log(f"[train_surrogate] wrote {a.out} ({time.time() - t0:.1f}s)")
    for d in vdev:
        d.free()
    sur.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

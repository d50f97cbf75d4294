"""Train the chemistry surrogate on the GPU and write its weights file.

    python -m openstc_amd.train_surrogate --out openstc_amd/data/surrogate_h2air9.npz
    python -m openstc_amd.train_surrogate --workload mixing --activation hswish --augment 0 --out ...

1. Run the finite-rate workloads on the GPU -- the mixing layer (192x96x96,
   FP32 bulk, DPLUR, CFL 0.2; 400 iterations, every 20th sampled) and the
   four-block H2 jet in vitiated air (4 x 256x192x128, SST, CFL 0.2; 120
   iterations through its ignition, every 10th sampled) -- and sample
   interior states (T, p, Y, local dt) of every block (training seeds) and,
   separately, from runs with other initial perturbations (held-out seeds).
2. Widen the states to the full box: compositions drawn from the harvested
   states at T uniform in [300, 2800] K, p log-uniform in [0.5, 3] atm and dt
   log-uniform over the harvested local time steps (training and held-out
   draws apart).
3. Label every state with the plain point-implicit step on the GPU (FP64);
   harvested training states are drawn evenly over decades of max|dY|.
4. Train the MLP (GELU by default) with the library's kernels (Adam,
   warm-up + cosine).
5. Calibrate the trusted output range dy_max on half of the held-out states
   (largest value meeting the error target) and report the other half, per
   source and without the gate.
6. Save weights, descriptor, training history and a held-out sample.
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


def augment(ctx, y_pool, n, t_range, p_range_atm, dt_range, rng):
    """n states over the full box: compositions from the pool, T uniform,
    p and dt log-uniform; labelled by the plain step on the GPU."""
    if n <= 0:
        return None
    y = y_pool[rng.integers(0, len(y_pool), size=n)].copy()
    t = rng.uniform(t_range[0], t_range[1], size=n)
    p = S.ATM * np.exp(rng.uniform(np.log(p_range_atm[0]), np.log(p_range_atm[1]), size=n))
    dt = np.exp(rng.uniform(np.log(dt_range[0]), np.log(dt_range[1]), size=n))
    dy = S.plain_step(ctx, t, p, y, dt)
    return t, p, y, dt, dy, np.full(n, -1)


def cat(*parts):
    parts = [q for q in parts if q is not None]
    return tuple(np.concatenate([q[i] for q in parts]) for i in range(6))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--workload", choices=["mixing", "jet", "both"], default="both")
    ap.add_argument("--steps", type=int, default=400, help="mixing-layer iterations per harvest run")
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--per-snapshot", type=int, default=60000, help="training states per block and snapshot")
    ap.add_argument("--val-per-snapshot", type=int, default=12000)
    ap.add_argument("--grid", type=int, nargs=3, default=[192, 96, 96])
    ap.add_argument("--jet-steps", type=int, default=120)
    ap.add_argument("--jet-every", type=int, default=10)
    ap.add_argument("--jet-per-snapshot", type=int, default=25000)
    ap.add_argument("--jet-val-per-snapshot", type=int, default=5000)
    ap.add_argument("--jet-grid", type=int, nargs=3, default=[256, 192, 128])
    ap.add_argument("--cfl", type=float, default=0.2)
    ap.add_argument("--augment", type=int, default=600000, help="synthetic training states over the full box")
    ap.add_argument("--t-range", type=float, nargs=2, default=[300.0, 2800.0])
    ap.add_argument("--p-range-atm", type=float, nargs=2, default=[0.5, 3.0])
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--keep-val", type=int, default=8192, help="held-out states stored with the weights")
    ap.add_argument("--uniform", action="store_true", help="sample training states uniformly, not by |dY| decade")
    ap.add_argument("--mag-weight", type=float, default=16.0)
    ap.add_argument("--target", type=float, default=8e-4, help="rel_l2 the output gate is calibrated to")
    ap.add_argument("--activation", choices=sorted(S.ACTIVATIONS), default="gelu")
    a = ap.parse_args(argv)
    t0 = time.time()
    rng = np.random.default_rng(a.seed + 11)
    tr_parts, va_parts, sources = [], [], []
    if a.workload in ("mixing", "both"):
        mx, my, mz = a.grid
        log(f"[train_surrogate] harvest mixing layer {mx}x{my}x{mz}, {a.steps} iterations, CFL {a.cfl}")
        tr_parts.append(S.harvest(cases.mixing_layer(mx=mx, my=my, mz=mz, cfl=a.cfl, seed=3), a.steps, a.every,
                                  a.per_snapshot, seed=a.seed, log=log, stratify=not a.uniform))
        # held-out: another initial perturbation, sampled uniformly (the solver's own distribution)
        va_parts.append(S.harvest(cases.mixing_layer(mx=mx, my=my, mz=mz, cfl=a.cfl, seed=4), a.steps, a.every,
                                  a.val_per_snapshot, seed=a.seed + 1, log=log, stratify=False))
        sources.append("mixing")
    if a.workload in ("jet", "both"):
        mx, my, mz = a.jet_grid
        log(f"[train_surrogate] harvest jet 4 x {mx}x{my}x{mz}, {a.jet_steps} iterations, CFL {a.cfl}")
        tr_parts.append(S.harvest(cases.jet_crossflow(mx=mx, my=my, mz=mz, cfl=a.cfl, seed=6), a.jet_steps,
                                  a.jet_every, a.jet_per_snapshot, seed=a.seed + 2, log=log,
                                  stratify=not a.uniform))
        va_parts.append(S.harvest(cases.jet_crossflow(mx=mx, my=my, mz=mz, cfl=a.cfl, seed=7), a.jet_steps,
                                  a.jet_every, a.jet_val_per_snapshot, seed=a.seed + 3, log=log, stratify=False))
        sources.append("jet")
    mech, info = cases.load_mechanism(True)
    param = cases.mixture_param(mech, False, True, 4, a.cfl, 1,
                                cases.InflowState(1.0, 1.0, 0, 0, 1e5, 1000.0, 1.0, 0.0, [0.0] * 9))
    ctx = Context(param)
    harv = cat(*tr_parts)
    dt_range = (float(harv[3].min()), float(harv[3].max()))
    aug_tr = augment(ctx, harv[2], a.augment, a.t_range, a.p_range_atm, dt_range, rng)
    aug_va = augment(ctx, harv[2], a.augment // 5, a.t_range, a.p_range_atm, dt_range, rng)
    if aug_va is not None:
        va_parts.append(aug_va)
        sources.append("box")
    t, p, y, dt, dy, it = cat(harv, aug_tr)
    src_id = np.concatenate([np.full(len(v[0]), k) for k, v in enumerate(va_parts)])
    vt, vp, vy, vdt, vdy, vit = cat(*va_parts)
    log(f"[train_surrogate] {len(t)} training / {len(vt)} held-out states labelled "
        f"(T {t.min():.0f}..{t.max():.0f} K, dt {dt.min():.2e}..{dt.max():.2e} s; {time.time() - t0:.1f}s)")
    desc = S.make_desc(t, p, y, dt, dy, param, info, mag_weight=a.mag_weight,
                       activation=S.ACTIVATIONS[a.activation])
    target = S.standardised_targets(dy, desc)
    sur = S.Surrogate(ctx, desc, max_batch=a.batch, seed=a.seed)
    vdev = [S.DeviceArray(np.ascontiguousarray(v)) for v in (vt, vp, vy, vdt)]
    val = (len(vt), *vdev, vdy)
    hist = S.train(sur, t, p, y, dt, target, a.epochs, a.batch, a.lr, seed=a.seed, val=val, log=log)

    def per_source(s_, idx):
        out = {}
        for k, name in enumerate(sources):
            sel = idx[src_id[idx] == k]
            if len(sel) == 0:
                continue
            dev = [S.DeviceArray(np.ascontiguousarray(v[sel])) for v in (vt, vp, vy, vdt)]
            out[name] = S.evaluate(s_, len(sel), *dev, vdy[sel])
            for d in dev:
                d.free()
        return out

    everything = np.arange(len(vt))
    ungated = S.evaluate(sur, *val)
    ungated_src = per_source(sur, everything)
    log("[train_surrogate] held-out, no output gate:", json.dumps(ungated), json.dumps(ungated_src))
    # trusted output range: calibrated on half of the held-out states, reported on the other half
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
    final_src = per_source(sur, tst)
    log("[train_surrogate] held-out test half, gated:", json.dumps(final), json.dumps(final_src))
    for d in tdev:
        d.free()
    keep = tst[np.random.default_rng(a.seed).choice(len(tst), size=min(a.keep_val, len(tst)), replace=False)]
    sur.save(a.out, history=np.array(json.dumps(hist)), heldout_t=vt[keep], heldout_p=vp[keep],
             heldout_y=vy[keep], heldout_dt=vdt[keep], heldout_dy=vdy[keep],
             info=np.array(json.dumps({"workload": a.workload, "sources": sources, "grid": a.grid,
                                       "jet_grid": a.jet_grid, "cfl": a.cfl, "steps": a.steps, "every": a.every,
                                       "jet_steps": a.jet_steps, "jet_every": a.jet_every,
                                       "augment": a.augment, "t_range": a.t_range, "p_range_atm": a.p_range_atm,
                                       "dt_range": dt_range,
                                       "n_train": int(len(t)), "n_heldout": int(len(vt)), "epochs": a.epochs,
                                       "batch": a.batch, "lr": a.lr, "mag_weight": a.mag_weight,
                                       "activation": a.activation,
                                       "heldout_ungated": ungated, "heldout_ungated_by_source": ungated_src,
                                       "dy_max": dy_max, "target": a.target,
                                       "heldout_test_gated": final, "heldout_test_gated_by_source": final_src})))
    log(f"[train_surrogate] wrote {a.out} ({time.time() - t0:.1f}s)")


if __name__ == "__main__":
    sys.exit(main())

#!/usr/bin/env python3
"""Throughput of the per-iteration residual-and-update path on one MI355X.

Workload (default): the viscous RANS configuration -- single-species air,
k-omega SST, one 256x128x64 flat-plate block with tanh wall clustering,
Mach 2.5, Re=1e7/m, T=220 K, MUSCL+minmod+AUSM+ with the 2nd-order central
viscous flux, DPLUR with 4 sweeps, CFL 5; FP32 bulk storage/arithmetic.
A "step" is one full steady iteration of the Driver loop (source, flux
sweeps, local dt, DPLUR, update, limit_flow, BCs, properties; the residual
every 100 steps), with the state resident in HBM.

Prints ONE JSON line.  value = Mcell-iter/s; roofline = the dominant
kernel's algorithmic bytes / its measured average duration (HIP events on the
launch stream over the timed region) vs the 8 TB/s HBM peak; cpu_baseline =
the FP64 C oracle (OpenMP) timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of a stage's kernel, measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE on this workload (FETCH_SIZE scaled x2.0 and
# WRITE_SIZE x1.0, the factors a 512 MiB copy kernel gives).  Keyed by
# (workload, stage); the summary with the commands is
# profiles/kernels_rans_256x128x64_fr_192x96x96_fourblock_4x256x192x128.md.  Re-measured whenever
# the kernel changes; a stage without an entry reports null.
MEASURED_TRAFFIC = {
    ("sst_flat_plate_dplur", "dplur_sweep"): 478.8e6,
    ("sst_flat_plate_dplur", "dplur_sweep_first"): 494.3e6,
    ("sst_flat_plate_dplur", "flux_ij"): 402.7e6,
    ("sst_flat_plate_dplur", "flux_k"): 402.3e6,
    ("sst_flat_plate_dplur", "source"): 354.7e6,
    ("sst_flat_plate_dplur", "update"): 273.5e6,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-chem", action="store_true", help="skip the finite-rate / surrogate leg")
    ap.add_argument("--chem-steps", type=int, default=10)
    ap.add_argument("--chem-eval-iter", type=int, default=200,
                    help="solver iteration at which surrogate and plain chemistry are compared")
    ap.add_argument("--chem-reps", type=int, default=5)
    ap.add_argument("--surrogate", default=None, help="surrogate weights file (default: the packaged model)")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-scale", action="store_true", help="skip the four-block 25 M-cell leg")
    ap.add_argument("--scale-steps", type=int, default=5)
    ap.add_argument("--profile-json", default=None, help="also write per-stage timings here")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="instead: run the streaming-copy calibration kernel on N MiB and exit")
    return ap.parse_args()


def copy_probe(mib: int, reps: int = 5):
    """Known-size copy kernel: the HBM-counter calibration and achievable-bandwidth reference."""
    from openstc_amd import _lib as L

    lib = L.load()
    nbytes = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    L.check(lib.ostc_device_alloc(C.byref(a), nbytes), "alloc")
    L.check(lib.ostc_device_alloc(C.byref(b), nbytes), "alloc")
    L.check(lib.ostc_stream_copy(b, a, nbytes, None), "copy")
    L.check(lib.ostc_device_sync(), "sync")
    t0 = time.perf_counter()
    for _ in range(reps):
        L.check(lib.ostc_stream_copy(b, a, nbytes, None), "copy")
    L.check(lib.ostc_device_sync(), "sync")
    dt = (time.perf_counter() - t0) / reps
    lib.ostc_device_free(a)
    lib.ostc_device_free(b)
    print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes,
                      "avg_s": dt, "GBs_read_plus_write": 2 * nbytes / dt / 1e9}), flush=True)


# Algorithmic scalars per cell (distinct values read + written, stencil
# neighbours counted once) for the air / SST / DPLUR path, per launch.
def alg_scalars(nv: int, rans: bool, dplur: bool):
    nsc = 2 if rans else 0
    turb = 5 if rans else 0  # mut, lambda_t, wall distance (+ rho, k, omega already in bv/sv)
    flux = 6 + nsc + 9 + 1 + 2 + (turb - 2 if rans else 0) + 2 * nv  # bv, sv, metric, J, mu, lambda (+mut, lt, d), dq RMW
    return {
        "flux_i": flux, "flux_j": flux, "flux_k": flux, "flux_ij": flux,
        # SST: + the local time step formed in the same kernel (sound speed
        # read, spectral radii 3 + dt written)
        "source": (4 + nsc + 9 + 1 + 1 + 1) + (2 + nv + 2) + (1 + 3 + 1) if rans else nv,
        "local_dt": 9 + 3 + 3 + (1 if rans else 0) + 4,
        "dplur_dq0": nv + 1 + 3 + (2 if rans else 0) + nv,
        "dplur_sweep": nv + 5 + 1 + nsc + 9 + 3 + 1 + nv + (2 if rans else 0) + nv,
        # first sweep of a non-reacting zone: reads dq in place of dq0 and
        # writes dq0 in place of reading it -- the same count
        "dplur_sweep_first": nv + 5 + 1 + nsc + 9 + 3 + 1 + nv + (2 if rans else 0) + nv,
        "update": nv + nv + 2 + nv + 6 + nsc + 1,
        "limit_flow": 5 + nsc,
        "properties": 6,
    }


def cpu_baseline(case_builder, steps: int):
    """FP64 oracle, OpenMP over (k,j), same configuration, `steps` iterations."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ffi as O

    case = case_builder()
    sim = O.OracleSimulation(case)
    blk, ob = case.blocks[0], sim.blocks[0]
    if case.wall_points is not None:
        # nearest wall node on the tensor-product wall grid (equal to the
        # brute-force minimum of src/Driver.cu:461-482 for this grid)
        xs = np.unique(blk.x[0, 0, :])
        zs = np.unique(blk.z[:, 0, 0])
        X, Y, Z = ob.arr["x"][0], ob.arr["y"][0], ob.arr["z"][0]
        nx_ = xs[np.clip(np.searchsorted(xs, X), 0, len(xs) - 1)]
        nx2 = xs[np.clip(np.searchsorted(xs, X) - 1, 0, len(xs) - 1)]
        dx = np.minimum(np.abs(X - nx_), np.abs(X - nx2))
        nz_ = zs[np.clip(np.searchsorted(zs, Z), 0, len(zs) - 1)]
        nz2 = zs[np.clip(np.searchsorted(zs, Z) - 1, 0, len(zs) - 1)]
        dz = np.minimum(np.abs(Z - nz_), np.abs(Z - nz2))
        ob.arr["wall_distance"][0] = np.sqrt(Y * Y + dx * dx + dz * dz)
    # initialisation as the Driver does (src/Driver.cu:43-104), wall distance set above
    lib, p = sim.lib, C.byref(sim.param)
    z0 = C.byref(sim.zones[0])
    lib.orc_apply_bc(z0, p, C.byref(sim.bcs[0]))
    lib.orc_compute_cv_from_bv(z0, p)
    lib.orc_initialize_mut(z0, p)
    lib.orc_update_physical_properties(z0, p)
    sim.step()  # first iteration (residual scale) outside the timing
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step(output_screen=case.output_screen)
    dt = time.perf_counter() - t0
    cells = sum(b.mx * b.my * b.mz for b in case.blocks)
    return {
        "value": cells * steps / dt / 1e6,
        "unit": "Mcell-iter/s",
        "cores": int(lib.orc_omp_threads()),
        "kind": "port",
        "sample": f"{steps} iterations of the same {case.name} block ({cells} cells), FP64 C oracle with OpenMP "
                  f"over (k,j); setup excluded",
        "seconds": dt,
    }


SUR_FLOP_PER_CELL = 2 * (12 * 128 + 3 * 128 * 128 + 128 * 10)  # algorithmic MLP forward FLOP per cell


def chemistry_leg(steps: int, warmup: int, eval_iter: int, reps: int, weights: str | None = None):
    """Finite-rate workload (192x96x96 mixing layer, FP32 bulk + FP64
    chemistry, DPLUR x4, CFL 0.2): whole-iteration throughput, then, on the
    solver's own state at iteration `eval_iter`, the chemistry step of every
    cell by the plain FP64 point-implicit integrator vs the surrogate with
    fallback (same zone, same states)."""
    import numpy as np

    from openstc_amd import cases, surrogate as S
    from openstc_amd import _lib as L
    from openstc_amd.runtime import Simulation

    case = cases.mixing_layer(cfl=0.2, seed=5)
    cells = sum(b.mx * b.my * b.mz for b in case.blocks)
    print(f"[bench] {case.name}: {cells} cells, setup", file=sys.stderr, flush=True)
    sim = Simulation(case)
    sim.initialize_computation()
    lib = sim.ctx.lib
    for _ in range(warmup):
        sim.step(output_screen=case.output_screen)
    sim.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step(output_screen=case.output_screen)
    sim.sync()
    el = time.perf_counter() - t0
    out = {"workload": case.name, "cells": cells, "steps": steps,
           "value": round(cells * steps / el / 1e6, 3), "unit": "Mcell-iter/s",
           "ms_per_step": round(el / steps * 1e3, 4)}
    done = warmup + steps
    while done < eval_iter:
        sim.step(output_screen=case.output_screen)
        done += 1
    sim.sync()
    weights = weights or S.DEFAULT_WEIGHTS
    if not os.path.exists(weights):
        out["surrogate"] = None
        sim.close()
        return out
    sur = S.Surrogate.load(sim.ctx, weights)
    z = sim.zones[0]
    dy_p = S.DeviceArray(nbytes=cells * 9 * 8)
    dy_s = S.DeviceArray(nbytes=cells * 9 * 8)
    stats = (C.c_longlong * 4)()

    def run(mode, buf, st):
        L.check(lib.ostc_chem_zone_step(sim.ctx.h, z.h, sur.h, mode, buf.p, st, None), "zone chemistry")

    run(0, dy_p, None)
    run(1, dy_s, stats)
    times = {}
    for mode, buf in ((0, dy_p), (1, dy_s)):
        L.check(lib.ostc_device_sync(), "sync")
        t0 = time.perf_counter()
        for _ in range(reps):
            run(mode, buf, None)
        L.check(lib.ostc_device_sync(), "sync")
        times[mode] = (time.perf_counter() - t0) / reps
    run(1, dy_s, stats)
    ref = dy_p.get(np.float64, (cells, 9))
    got = dy_s.get(np.float64, (cells, 9))
    rep = S.error_report(got, ref, list(stats))
    out["surrogate"] = {
        "eval_iteration": eval_iter,
        "plain_ms": round(times[0] * 1e3, 4), "surrogate_ms": round(times[1] * 1e3, 4),
        "speedup": round(times[0] / times[1], 3),
        "plain_Mcell_s": round(cells / times[0] / 1e6, 2), "surrogate_Mcell_s": round(cells / times[1] / 1e6, 2),
        "surrogate_TFLOPs_incl_fallback": round(SUR_FLOP_PER_CELL * cells / times[1] / 1e12, 2),
        "mfma_peak_TFLOPs_bf16": 2500.0,
        **{k: (round(v, 6) if isinstance(v, float) else v) for k, v in rep.items()},
        "error_definition": "rel_l2 = ||dY_sur - dY_plain||_2 / ||dY_plain||_2 over all cells and species "
                            "(dY/dt is the effective production rate); cell_* = max_k|err_k| / max_k|dY_plain_k| "
                            "over cells with max|dY_plain| > 1e-8",
    }
    for b in (dy_p, dy_s):
        b.free()
    sur.close()
    sim.close()
    return out


def scale_leg(steps: int, warmup: int, reps: int, dims=(256, 192, 128), nblocks: int = 4,
              weights: str | None = None):
    """Full path at scale: the four-block SST + finite-rate H2-jet workload
    (4 x 256x192x128 = 25.2 M cells, FP32 bulk + FP64 chemistry/Newton, EPI,
    DPLUR x4, on-device inner-face halos).  Whole-iteration throughput, then
    the chemistry step of every cell by the plain FP64 point-implicit
    integrator vs the surrogate with fallback on the solver's state."""
    import numpy as np

    from openstc_amd import cases, surrogate as S
    from openstc_amd import _lib as L
    from openstc_amd.runtime import Simulation

    mx, my, mz = dims
    t0 = time.perf_counter()
    case = cases.jet_crossflow(mx=mx, my=my, mz=mz, nblocks=nblocks)
    cells = sum(b.mx * b.my * b.mz for b in case.blocks)
    print(f"[bench] {case.name}: {cells} cells, setup", file=sys.stderr, flush=True)
    sim = Simulation(case)
    for blk in case.blocks:  # host copies are no longer needed
        blk.bv0 = blk.sv0 = blk.xg = blk.yg = blk.zg = blk.jac = blk.metric = None
    sim.initialize_computation()
    setup_s = time.perf_counter() - t0
    lib = sim.ctx.lib
    for _ in range(warmup):
        sim.step(output_screen=case.output_screen)
    sim.sync()
    print(f"[bench] {case.name}: timing {steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step(output_screen=case.output_screen)
    sim.sync()
    el = time.perf_counter() - t0
    out = {"workload": case.name, "blocks": nblocks, "block": list(dims), "cells": cells, "steps": steps,
           "warmup": warmup, "value": round(cells * steps / el / 1e6, 3), "unit": "Mcell-iter/s",
           "ms_per_step": round(el / steps * 1e3, 3), "setup_s": round(setup_s, 1),
           "model": "9-species/19-reaction H2-air finite-rate (EPI) + k-omega SST, DPLUR x4, CFL %g, "
                    "FP32 bulk + FP64 chemistry/Newton" % case.param.cfl}
    weights = weights or (S.JET_WEIGHTS if os.path.exists(S.JET_WEIGHTS) else S.DEFAULT_WEIGHTS)
    if not os.path.exists(weights):
        out["surrogate"] = None
        sim.close()
        return out
    sur = S.Surrogate.load(sim.ctx, weights)
    ncell = [b.mx * b.my * b.mz for b in case.blocks]
    nmax = max(ncell)
    dy_p = S.DeviceArray(nbytes=nmax * 9 * 8)
    dy_s = S.DeviceArray(nbytes=nmax * 9 * 8)

    def run(z, mode, buf, st):
        L.check(lib.ostc_chem_zone_step(sim.ctx.h, z.h, sur.h, mode, buf.p, st, None), "zone chemistry")

    times = {0: 0.0, 1: 0.0}
    sq_err = sq_ref = 0.0
    fallbacks = 0
    for z, n in zip(sim.zones, ncell):
        run(z, 0, dy_p, None)
        run(z, 1, dy_s, None)
        for mode, buf in ((0, dy_p), (1, dy_s)):
            L.check(lib.ostc_device_sync(), "sync")
            t0 = time.perf_counter()
            for _ in range(reps):
                run(z, mode, buf, None)
            L.check(lib.ostc_device_sync(), "sync")
            times[mode] += (time.perf_counter() - t0) / reps
        stats = (C.c_longlong * 4)()
        run(z, 1, dy_s, stats)
        ref = dy_p.get(np.float64, (n, 9))
        got = dy_s.get(np.float64, (n, 9))
        sq_err += float(np.sum((got - ref) ** 2))
        sq_ref += float(np.sum(ref ** 2))
        fallbacks += int(stats[1] + stats[2] + stats[3])
    out["surrogate"] = {
        "plain_ms": round(times[0] * 1e3, 3), "surrogate_ms": round(times[1] * 1e3, 3),
        "speedup": round(times[0] / times[1], 3),
        "rel_l2": round(float(np.sqrt(sq_err / max(sq_ref, 1e-300))), 6),
        "fallback_rate": round(fallbacks / cells, 5),
        "weights": os.path.basename(weights),
        "note": ("model trained on this workload's states (other initial perturbations, its first 60 iterations)"
                 if os.path.basename(weights) == os.path.basename(S.JET_WEIGHTS) else
                 "model trained on the 192x96x96 mixing-layer states; this workload's vitiated-air / jet states "
                 "are partly outside that range and fall back to the plain integrator"),
    }
    for b in (dy_p, dy_s):
        b.free()
    sur.close()
    sim.close()
    return out


def main():
    a = parse()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only (one block per GPU is the scale-out "
                                   "design; multi-GPU halos are not built)", "n_gpus": a.gpus}))
        sys.exit(2)
    from openstc_amd import _lib as L
    from openstc_amd import cases
    from openstc_amd.runtime import Simulation

    if a.config == 2:
        build = lambda: cases.sst_plate()  # noqa: E731
    else:
        build = lambda: cases.warped_box(mx=256, my=256, mz=64)  # noqa: E731
    case = build()
    cells = sum(b.mx * b.my * b.mz for b in case.blocks)
    print(f"[bench] {case.name}: {cells} cells, setup", file=sys.stderr, flush=True)
    sim = Simulation(case)
    sim.initialize_computation()
    lib = sim.ctx.lib
    for _ in range(a.warmup):
        sim.step(output_screen=case.output_screen)
    sim.sync()
    print(f"[bench] timing {a.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        sim.step(output_screen=case.output_screen)
    sim.sync()
    elapsed = time.perf_counter() - t0
    # per-stage times (HIP events around each stage) from a second pass of the
    # same length, so that the events do not enter the timed region
    L.check(lib.ostc_profile(sim.ctx.h, 1), "profile")
    for _ in range(a.steps):
        sim.step(output_screen=case.output_screen)
    sim.sync()
    nst = lib.ostc_profile_stage_count()
    ms = (C.c_double * nst)()
    cnt = (C.c_longlong * nst)()
    L.check(lib.ostc_profile_read(sim.ctx.h, ms, cnt, nst, 1), "profile_read")
    L.check(lib.ostc_profile(sim.ctx.h, 0), "profile off")
    names = [lib.ostc_profile_stage_name(i).decode() for i in range(nst)]
    p = case.param
    rans = p.turb_method == L.RANS
    esize = 4 if case.precision == L.FP32 else 8
    scal = alg_scalars(p.n_var, rans, p.implicit_method == 1)
    stages = {}
    for i, n in enumerate(names):
        if cnt[i] == 0:
            continue
        avg_ms = ms[i] / cnt[i]
        st = {"total_ms": ms[i], "launches": int(cnt[i]), "avg_ms": avg_ms}
        if n in scal:
            byts = scal[n] * esize * cells
            st["alg_bytes_per_launch"] = byts
            st["achieved_GBs"] = byts / (avg_ms * 1e-3) / 1e9
            st["frac_hbm_peak"] = st["achieved_GBs"] / HBM_PEAK_GBS
        if (case.name, n) in MEASURED_TRAFFIC:
            st["traffic_bytes_per_launch"] = MEASURED_TRAFFIC[(case.name, n)]
        stages[n] = st
    dom = max((n for n in stages if "alg_bytes_per_launch" in stages[n]), key=lambda n: stages[n]["total_ms"])
    d = stages[dom]
    ms_per_step = elapsed / a.steps * 1e3
    value = cells * a.steps / elapsed / 1e6
    out = {
        "metric": "Mcell-iter/s (flux+SST+chem+DPLUR), %HBM peak; surrogate vs plain chem @1e-3",
        "value": round(value, 3),
        "unit": "Mcell-iter/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if esize == 4 else "f64",
        "data": "synthetic",
        "config": {
            "workload": case.name,
            "block": [case.blocks[0].mx, case.blocks[0].my, case.blocks[0].mz],
            "cells": cells,
            "model": "air k-omega SST, MUSCL+minmod+AUSM+, central viscous, DPLUR x%d, CFL %g"
                     % (p.dplur_inner_step, p.cfl) if rans else "air inviscid, explicit, CFL %g" % p.cfl,
            "precision": "FP32 bulk" if esize == 4 else "FP64",
        },
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": round(d["achieved_GBs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(d["frac_hbm_peak"], 4),
            "traffic": MEASURED_TRAFFIC.get((case.name, dom)),
            "traffic_unit": "bytes/launch (rocprofv3 FETCH_SIZE+WRITE_SIZE, calibrated)",
            "alg_bytes_per_cell": scal[dom] * esize,
            "avg_launch_ms": round(d["avg_ms"], 5),
        },
        "stages": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                   for k, v in stages.items()},
    }
    sim.close()
    if not a.no_chem:
        out["chemistry"] = chemistry_leg(a.chem_steps, 2, a.chem_eval_iter, a.chem_reps, a.surrogate)
    if not a.no_scale:
        out["full_path_at_scale"] = scale_leg(a.scale_steps, 2, a.chem_reps, weights=a.surrogate)
    if not a.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cb = cpu_baseline(build, a.cpu_steps)
        out["cpu_baseline"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in cb.items()}
    else:
        out["cpu_baseline"] = None
    if a.profile_json:
        with open(a.profile_json, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

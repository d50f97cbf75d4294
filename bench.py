#!/usr/bin/env python3
"""Throughput of the per-iteration residual-and-update path on one MI355X.

Headline workload: the full path at scale -- k-omega SST + 9-species /
19-reaction H2-air finite-rate chemistry (EPI), four blocks of 256x192x128
(25.2 M cells) joined by inner faces with on-device halo copies, a Mach 2.5
H2 jet into a Mach 2.0 vitiated-air crossflow, MUSCL+minmod+AUSM+ with the
central viscous flux, DPLUR with 4 sweeps, CFL 0.2; FP32 bulk storage and
arithmetic with FP64 chemistry (rates, Jacobian, its LU solve) and FP64
T-from-e Newton.  A "step" is one full steady iteration of the Driver loop
(source, flux sweeps, local dt, DQ0, DPLUR sweeps, update, limit_flow, BCs,
inner-face halos, properties; the residual every output_screen steps), with
the state resident in HBM.

The other benchmark configurations run as extra legs with their own
throughput, roofline and CPU baseline: the warped Euler box (64x64x8, FP64,
explicit), the SST flat plate (256x128x64, FP32 bulk, DPLUR x4, CFL 5) and the
finite-rate mixing layer (192x96x96), whose solver states also carry the
plain-vs-surrogate chemistry comparison.

Prints ONE JSON line.  value = Mcell-iter/s of the headline; roofline = the
dominant kernel's algorithmic bytes / its average launch duration (HIP events
recorded by the library on the launch stream, over a stage-timing pass of the
same workload) vs the 8 TB/s HBM peak; cpu_baseline = the FP64 C oracle
(OpenMP) timed on a bounded sample of the same workload.
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
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector
BF16_PEAK_TFLOPS = 2500.0    # MI355X dense BF16 MFMA
METRIC = "Mcell-iter/s (flux+SST+chem+DPLUR), %HBM peak; surrogate vs plain chem @1e-3"

# HBM bytes per launch of a stage's kernel, measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE on these workloads (FETCH_SIZE scaled x2.0 and
# WRITE_SIZE x1.0, the factors a 512 MiB copy kernel gives).  Keyed by
# (workload, stage); the summary with the commands is
# profiles/kernels_fourblock_jet_4x256x192x128_plate_mixing_layer.md.
# Re-measured whenever the kernel changes; a stage without an entry reports
# null.
MEASURED_TRAFFIC = {
    ("h2_jet_vitiated_crossflow_fr_sst_4x256x192x128", "dplur_sweep"): 2304.5e6,
    ("h2_jet_vitiated_crossflow_fr_sst_4x256x192x128", "dplur_chem"): 4276.7e6,
    ("h2_jet_vitiated_crossflow_fr_sst_4x256x192x128", "dplur_dq0"): 4582.7e6,
    ("h2_jet_vitiated_crossflow_fr_sst_4x256x192x128", "source"): 1061.4e6,
    ("h2_jet_vitiated_crossflow_fr_sst_4x256x192x128", "flux_i"): 3578.1e6,
    ("h2_jet_vitiated_crossflow_fr_sst_4x256x192x128", "flux_j"): 4034.7e6,
    ("h2_jet_vitiated_crossflow_fr_sst_4x256x192x128", "flux_k"): 3071.3e6,
    ("h2_jet_vitiated_crossflow_fr_sst_4x256x192x128", "update"): 1767.5e6,
    ("h2_jet_vitiated_crossflow_fr_sst_4x256x192x128", "properties"): 730.5e6,
    ("h2air_mixing_layer_fr_epi_dplur", "dplur_sweep"): 514.4e6,
    ("h2air_mixing_layer_fr_epi_dplur", "dplur_chem"): 1195.2e6,
    ("h2air_mixing_layer_fr_epi_dplur", "dplur_dq0"): 1257.6e6,
    ("h2air_mixing_layer_fr_epi_dplur", "source"): 144.8e6,
    ("h2air_mixing_layer_fr_epi_dplur", "flux_i"): 671.2e6,
    ("h2air_mixing_layer_fr_epi_dplur", "flux_j"): 838.4e6,
    ("h2air_mixing_layer_fr_epi_dplur", "flux_k"): 743.6e6,
    ("h2air_mixing_layer_fr_epi_dplur", "update"): 442.8e6,
    ("sst_flat_plate_dplur", "dplur_sweep"): 480.9e6,
    ("sst_flat_plate_dplur", "dplur_sweep_first"): 499.3e6,
    ("sst_flat_plate_dplur", "flux_ij"): 358.6e6,
    ("sst_flat_plate_dplur", "flux_k"): 402.3e6,
    ("sst_flat_plate_dplur", "source"): 300.8e6,
    ("sst_flat_plate_dplur", "update"): 273.5e6,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed iterations of the headline workload")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--stage-steps", type=int, default=4, help="iterations of the per-stage timing pass")
    ap.add_argument("--legs", default="warped_box,sst_plate,mixing_layer",
                    help="extra benchmark configurations (comma list; empty for none)")
    ap.add_argument("--no-headline", action="store_true", help="skip the four-block headline workload")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--block", default="256,192,128", help="headline block size (mx,my,mz)")
    ap.add_argument("--cpu-block", default="96,72,48", help="block size of the headline's CPU sample")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--chem-eval-iter", type=int, default=200,
                    help="mixing-layer iteration at which surrogate and plain chemistry are compared")
    ap.add_argument("--chem-reps", type=int, default=5)
    ap.add_argument("--surrogate", default=None, help="surrogate weights file (default: the packaged model)")
    ap.add_argument("--profile-json", default=None, help="also write the full record here")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="instead: run the streaming-copy calibration kernel on N MiB and exit")
    return ap.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


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


# ------------------------------------------------------------ algorithmic bytes

def alg_bytes_per_cell(p, esize: int) -> dict:
    """Algorithmic bytes per cell and launch of each stage: the distinct
    scalars a kernel must read plus write, stencil neighbours counted once,
    bulk fields at the zone's storage size (esize), the chemistry Jacobian and
    its LU factors at 8 B, LU pivots at 4 B.  Keyed by the stage names of
    ostc_profile_stage_name."""
    from openstc_amd import _lib as L

    ns, nt = p.n_spec, (2 if p.turb_method == L.RANS else 0)
    nv, nsc = 5 + ns + nt, ns + nt
    rans, visc = nt > 0, p.viscous_scheme != 0
    chem = p.mix_model == L.FR and p.chem_src_method in (1, 2)
    epi = p.mix_model == L.FR and p.chem_src_method == 1
    njac = ns * ns if epi else (ns if chem else 0)   # FP64 Jacobian entries per cell
    # the EPI block of DPLUR is factored in the space of the species that take
    # part in a reaction (na of them): na^2 LU entries, pivots packed 8 to a word
    na = sum(1 for l in range(ns) if any(p.stoi_f[r * L.MAX_SPEC + l] or p.stoi_b[r * L.MAX_SPEC + l]
                                          for r in range(p.n_reac)))
    npw = (na + 7) // 8
    # EPI under DPLUR on the 9-species kernels (at most 20 reactions): the
    # source kernel forms the SST source and the local time step only, and
    # the chemistry follows the fluxes in the dq0 stage, which also stores
    # the chemistry block's LU factors
    fac = epi and ns == 9 and p.n_reac <= 20 and p.implicit_method == 1 and na > 0
    out = {}

    def put(stage, n_bulk, n_f64=0, n_i32=0):
        out[stage] = n_bulk * esize + n_f64 * 8 + n_i32 * 4

    # flux, one pass over the cells per direction (inviscid + viscous fused):
    # bv 6, sv, metric 9, J, [mu, lambda, (mu_t, lambda_t, d), rho*D_k], dq RMW.
    # The first direction writes the rows of dq no source kernel wrote
    # before it: a source kernel that forms the time step (SST zones without
    # chemistry, EPI zones with the chemistry after the fluxes) writes only
    # the turbulence rows, and zones without a source term none.
    dt_src = (rans and p.mix_model != L.FR) or fac
    no_src = not rans and not chem
    first_reads = nt if dt_src else (0 if no_src else nv)
    flux = 6 + nsc + 9 + 1 + ((2 + (3 if rans else 0) + ns) if visc else 0) + 2 * nv
    for s in ("flux_j", "flux_k"):
        put(s, flux)
    for s in ("flux_i", "flux_ij"):
        put(s, flux - nv + first_reads)
    # source: SST reads rho,u,v,w, k,w, metric 9, J, mu, d (+T, cp for a
    # mixture) and writes mu_t, lambda_t, dq (first writer), J_t 2; the SST
    # source of an air zone also forms the local time step (reads c, writes
    # spectral radii 3 and dt); chemistry reads rho, T, Y and writes dq and
    # the FP64 Jacobian
    if rans:
        # an air zone's SST source also forms the time step and writes only
        # the turbulence rows of dq
        src = (4 + 2 + 9 + 1 + 1 + 1 + (ns + 2 if ns else 0)) + (2 + nv + 2)
        if ns == 0:
            src += 1 + 3 + 1 - (nv - nt)
    elif chem:
        src = 2 + ns + nv
    else:
        src = nv
    if fac:
        # no chemistry: SST (as above, without T and Y) or nothing, dq
        # written; the time step: c, gamma (+ metric 9, u, v, w, rho, mu
        # without SST) -> spectral radii 3, dt
        if rans:
            src = (4 + 2 + 9 + 1 + 1 + 1 + 2) + (2 + nt + 2)
        else:
            src = 9 + 3 + 1 + 1
        src += 2 + 4
        put("source", src)
    else:
        put("source", src, njac)
    # local time step: metric 9, u,v,w, rho, c [mu] [mu_t] [gamma] -> spectral radii 3, dt
    put("local_dt", 9 + 3 + 2 + (1 if visc else 0) + (1 if rans else 0) + (1 if ns else 0) + 4)
    # DQ0 (reacting zones; non-reacting zones form dq0 in the first sweep):
    # dq, dt, spectral radii 3, [J_t 2] -> dq0; the reacting species' block of
    # the Jacobian read, its LU factors + packed pivots written (fac: the
    # factors and pivots read)
    if fac:
        # the chemistry after the fluxes with dq0: rho, T, Y, J, dq, dt,
        # spectral radii 3, [J_t 2] -> dq0, the block's factors + packed pivots
        put("dplur_dq0", 2 + ns + 1 + nv + 1 + 3 + (2 if rans else 0) + nv, na * na, npw)
    elif epi:
        put("dplur_dq0", nv + 1 + 3 + (2 if rans else 0) + nv, 2 * na * na, npw)
    else:
        put("dplur_dq0", nv + 1 + 3 + (2 if rans else 0) + nv, njac)
    # one sweep: din, sv, u,v,w,|V|, (air p, rho | mixture T, gamma), metric 9,
    # spectral radii 3, dt, dq0, [J_t 2] -> dout.  With EPI the species rows'
    # block solve is its own kernel (dplur_chem): the sweep reads the flow and
    # turbulence rows of dq0 and leaves the species rows' sums in dout, which
    # dplur_chem reads back with dt, the species rows of dq0, the LU factors
    # and pivots (and the spectral radii 3 when a species takes part in no
    # reaction: its block row is the diagonal), and overwrites with the
    # solution.
    sweep = nv + nsc + 4 + 2 + 9 + 3 + 1 + nv + (2 if rans else 0) + nv
    if epi:
        put("dplur_sweep", sweep - ns)
        put("dplur_chem", 2 * ns + 1 + ns + (3 if na < ns else 0), na * na, npw)
    else:
        put("dplur_sweep", sweep, ns if chem else 0)
    put("dplur_sweep_first", sweep)   # reads dq in place of dq0 and writes dq0 in place of reading it
    # update: cv, dq, dt, J [T: the Newton start] -> cv, bv 6, sv, |V|
    put("update", nv + nv + 2 + (1 if ns else 0) + nv + 6 + nsc + 1)
    # properties: air T, |V| -> c, mu, lambda, Mach; mixture T, Y, |V| -> cp,
    # gamma, c, mu, lambda, rho*D_k, Mach
    put("properties", 6 if ns == 0 else 1 + ns + 1 + 5 + ns + 1)
    return out


def plain_chem_flop_per_cell(p) -> int:
    """Algorithmic FP64 FLOP count of one plain point-implicit chemistry step
    per cell, frozen from the mechanism's structure (transcendentals counted
    as 1 FLOP; pow(T, b) as exp(b ln T) with ln T shared):
      concentrations 3 per species; Gibbs g/RT of a NASA-7 fit 20 per species;
      per reaction: Arrhenius 6 (+7 for a duplicate pair), third-body sum 2
      per species, fall-off 12 (+26 for the Troe blending), backward rate
      4 + 2 per species with a net coefficient (or 6 for an explicit reverse
      Arrhenius), rate of progress 2 + one product per stoichiometric unit;
      production rates 2 per (reaction, species) pair with a net coefficient
      + 2 per species; EPI Jacobian 5 per (entry, reaction) pair where the row
      species changes and the column species takes part, + 3 per entry; the
      step's matrix assembly 2 per entry, right-hand side 1 per species, LU
      factorisation and solve 2/3 n^3 + 2 n^2."""
    ns, nr = p.n_spec, p.n_reac
    S = 20
    f = 3 * ns + S * ns + 1
    for r in range(nr):
        t = p.reac_type[r]
        nu_f = [p.stoi_f[r * S + l] for l in range(ns)]
        nu_b = [p.stoi_b[r * S + l] for l in range(ns)]
        f += 6 + (7 if t == 3 else 0)
        if t > 3:
            f += 2 * ns + (1 if t == 4 else 12 + (26 if t > 5 else 0))
        f += 6 if t == 2 else 4 + 2 * sum(1 for l in range(ns) if nu_b[l] != nu_f[l])
        f += 2 + sum(nu_f) + sum(nu_b)
        f += 2 * sum(1 for l in range(ns) if nu_b[l] != nu_f[l])
    f += 2 * ns
    for m in range(ns):
        for n in range(ns):
            f += 3
            for r in range(nr):
                if p.stoi_b[r * S + m] != p.stoi_f[r * S + m] and (p.stoi_f[r * S + n] or p.stoi_b[r * S + n]):
                    f += 5
    f += 2 * ns * ns + ns + (2 * ns ** 3) // 3 + 2 * ns * ns
    return int(f)


SUR_FLOP_PER_CELL = 2 * (12 * 128 + 3 * 128 * 128 + 128 * 10)  # algorithmic MLP forward FLOP per cell


# ------------------------------------------------------------------ legs

def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def counters(sim, reset=False):
    tot = [0, 0, 0, 0]
    for z in sim.zones:
        c = z.counters(reset=reset)
        tot = [a + b for a, b in zip(tot, c)]
    return {"newton_cap_hits": int(tot[0]), "limit_flow_repairs": int(tot[1]), "surrogate_fallbacks": int(tot[2])}


def time_leg(case, steps: int, warmup: int, stage_steps: int):
    """Whole-iteration throughput of one case (state resident in HBM, device
    synchronised on both sides of the timed loop), then a stage-timing pass
    of `stage_steps` iterations with the library's HIP events on the launch
    stream (outside the timed region: the events cost ~0.1 ms per iteration).
    Returns (record, sim)."""
    from openstc_amd import _lib as L
    from openstc_amd.runtime import Simulation

    cells = sum(b.mx * b.my * b.mz for b in case.blocks)
    t0 = time.perf_counter()
    sim = Simulation(case)
    for blk in case.blocks:  # host copies are no longer needed
        blk.bv0 = blk.sv0 = blk.xg = blk.yg = blk.zg = blk.jac = blk.metric = None
    sim.initialize_computation()
    setup_s = time.perf_counter() - t0
    lib = sim.ctx.lib
    for _ in range(warmup):
        sim.step(output_screen=case.output_screen)
    sim.sync()
    counters(sim, reset=True)
    log(f"{case.name}: {cells} cells, timing {steps} steps")
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step(output_screen=case.output_screen)
    sim.sync()
    elapsed = time.perf_counter() - t0
    cnt = counters(sim)
    stages = {}
    if stage_steps > 0:
        L.check(lib.ostc_profile(sim.ctx.h, 1), "profile")
        for _ in range(stage_steps):
            sim.step(output_screen=case.output_screen)
        sim.sync()
        nst = lib.ostc_profile_stage_count()
        ms = (C.c_double * nst)()
        cnt_s = (C.c_longlong * nst)()
        L.check(lib.ostc_profile_read(sim.ctx.h, ms, cnt_s, nst, 1), "profile_read")
        L.check(lib.ostc_profile(sim.ctx.h, 0), "profile off")
        names = [lib.ostc_profile_stage_name(i).decode() for i in range(nst)]
        p = case.param
        esize = 4 if case.precision == L.FP32 else 8
        alg = alg_bytes_per_cell(p, esize)
        for i, n in enumerate(names):
            if cnt_s[i] == 0:
                continue
            avg_ms = ms[i] / cnt_s[i]
            st = {"total_ms": ms[i] / stage_steps, "launches": round(cnt_s[i] / stage_steps, 2), "avg_ms": avg_ms}
            if n in alg:
                # a bracket covers one zone (or all zones, for the stages the
                # library brackets once per iteration: properties); launches
                # per iteration = zones x repetitions (DPLUR sweeps)
                per_it = cnt_s[i] / stage_steps
                reps = 1
                if n in ("dplur_sweep", "dplur_chem"):
                    reps = p.dplur_inner_step - (1 if any(x == "dplur_sweep_first" and cnt_s[j] for j, x in
                                                          enumerate(names)) and n == "dplur_sweep" else 0)
                byts = alg[n] * cells * reps / per_it
                st["alg_bytes_per_cell"] = alg[n]
                st["alg_bytes_per_launch"] = byts
                st["achieved_GBs"] = byts / (avg_ms * 1e-3) / 1e9
                st["frac_hbm_peak"] = st["achieved_GBs"] / HBM_PEAK_GBS
            if (case.name, n) in MEASURED_TRAFFIC:
                st["traffic_bytes_per_launch"] = MEASURED_TRAFFIC[(case.name, n)]
            stages[n] = st
    rec = {"workload": case.name, "blocks": len(case.blocks),
           "block": [case.blocks[0].mx, case.blocks[0].my, case.blocks[0].mz], "cells": cells,
           "steps": steps, "warmup": warmup,
           "value": round(cells * steps / elapsed / 1e6, 3), "unit": "Mcell-iter/s",
           "ms_per_step": round(elapsed / steps * 1e3, 4), "setup_s": round(setup_s, 1),
           "counters": cnt}
    if stages:
        dom = max((n for n in stages if "alg_bytes_per_launch" in stages[n]), key=lambda n: stages[n]["total_ms"])
        d = stages[dom]
        rec["roofline"] = {
            "kernel": dom, "bound": "hbm", "achieved": round(d["achieved_GBs"], 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(d["frac_hbm_peak"], 4),
            "traffic": MEASURED_TRAFFIC.get((case.name, dom)),
            "traffic_unit": "bytes/launch (rocprofv3 FETCH_SIZE+WRITE_SIZE, calibrated)",
            "alg_bytes_per_cell": d["alg_bytes_per_cell"], "avg_launch_ms": round(d["avg_ms"], 5)}
        rec["stages"] = {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                         for k, v in stages.items()}
    return rec, sim


def plate_wall_distance(case, sim):
    """Wall distance of a single-block flat plate from the nearest node of its
    tensor-product wall grid (equal to the brute-force minimum of
    src/Driver.cu:461-482 for this grid, without its O(N * N_wall) cost)."""
    import numpy as np

    blk, ob = case.blocks[0], sim.blocks[0]
    xs = np.unique(blk.x[0, 0, :])
    zs = np.unique(blk.z[:, 0, 0])
    X, Y, Z = ob.arr["x"][0], ob.arr["y"][0], ob.arr["z"][0]
    dx = np.minimum(np.abs(X - xs[np.clip(np.searchsorted(xs, X), 0, len(xs) - 1)]),
                    np.abs(X - xs[np.clip(np.searchsorted(xs, X) - 1, 0, len(xs) - 1)]))
    dz = np.minimum(np.abs(Z - zs[np.clip(np.searchsorted(zs, Z), 0, len(zs) - 1)]),
                    np.abs(Z - zs[np.clip(np.searchsorted(zs, Z) - 1, 0, len(zs) - 1)]))
    ob.arr["wall_distance"][0] = np.sqrt(Y * Y + dx * dx + dz * dz)


def cpu_baseline(case, steps: int, sample: str, plate: bool = False):
    """FP64 oracle, OpenMP over (k,j), the same configuration; one untimed
    iteration (sets the residual scale), then `steps` timed iterations."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ffi as O

    sim = O.OracleSimulation(case)
    if plate:
        # initialisation as the Driver does (src/Driver.cu:43-104), wall distance from the tensor-product wall
        plate_wall_distance(case, sim)
        lib, p, z0 = sim.lib, C.byref(sim.param), C.byref(sim.zones[0])
        lib.orc_apply_bc(z0, p, C.byref(sim.bcs[0]))
        lib.orc_compute_cv_from_bv(z0, p)
        lib.orc_initialize_mut(z0, p)
        lib.orc_update_physical_properties(z0, p)
    else:
        sim.initialize_computation()
    sim.step()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step(output_screen=case.output_screen)
    dt = time.perf_counter() - t0
    cells = sum(b.mx * b.my * b.mz for b in case.blocks)
    return {"value": round(cells * steps / dt / 1e6, 4), "unit": "Mcell-iter/s",
            "cores": int(sim.lib.orc_omp_threads()), "kind": "port", "cpu": cpu_model(),
            "sample": f"{sample}: {steps} iterations of {cells} cells after one untimed iteration, FP64 C "
                      f"oracle with OpenMP over (k,j); setup excluded",
            "seconds": round(dt, 3)}


def headline(a):
    from openstc_amd import cases

    mx, my, mz = (int(v) for v in a.block.split(","))
    case = cases.jet_crossflow(mx=mx, my=my, mz=mz, nblocks=4)
    log(f"{case.name}: setup")
    rec, sim = time_leg(case, a.steps, a.warmup, a.stage_steps)
    rec["model"] = ("9-species/19-reaction H2-air finite-rate (EPI) + k-omega SST, MUSCL+minmod+AUSM+, central "
                    "viscous, DPLUR x%d, CFL %g, inner-face halos on device" % (case.param.dplur_inner_step,
                                                                                 case.param.cfl))
    rec["precision"] = "FP32 bulk + FP64 chemistry (rates, Jacobian, LU solve) and T-from-e Newton"
    rec["surrogate"] = surrogate_at_scale(sim, case, a.chem_reps, a.surrogate)
    sim.close()
    return rec


def surrogate_at_scale(sim, case, reps: int, weights: str | None):
    """Plain FP64 point-implicit chemistry vs the surrogate with fallback on
    every cell of the four blocks (the solver's state after the timed steps)."""
    import numpy as np

    from openstc_amd import surrogate as S
    from openstc_amd import _lib as L

    weights = weights or S.DEFAULT_WEIGHTS
    if not os.path.exists(weights):
        return None
    lib = sim.ctx.lib
    sur = S.Surrogate.load(sim.ctx, weights)
    ncell = [b.mx * b.my * b.mz for b in case.blocks]
    dy_p = S.DeviceArray(nbytes=max(ncell) * 9 * 8)
    dy_s = S.DeviceArray(nbytes=max(ncell) * 9 * 8)

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
    cells = sum(ncell)
    out = {"plain_ms": round(times[0] * 1e3, 3), "surrogate_ms": round(times[1] * 1e3, 3),
           "speedup": round(times[0] / times[1], 3),
           "rel_l2": round(float(np.sqrt(sq_err / max(sq_ref, 1e-300))), 6),
           "fallback_rate": round(fallbacks / cells, 5), "weights": os.path.basename(weights)}
    for b in (dy_p, dy_s):
        b.free()
    sur.close()
    return out


def leg_warped_box(a):
    from openstc_amd import cases

    case = cases.warped_box(mx=64, my=64, mz=8, steps=100)
    rec, sim = time_leg(case, 100, 5, 10)
    rec["model"] = "air (gamma 1.4) inviscid, MUSCL+minmod+AUSM+, explicit Euler, CFL 0.5, FP64"
    sim.close()
    if not a.no_cpu:
        rec["cpu_baseline"] = cpu_baseline(cases.warped_box(mx=64, my=64, mz=8, steps=100), 20,
                                           "the same 64x64x8 block")
    return rec


def leg_sst_plate(a):
    from openstc_amd import cases

    case = cases.sst_plate()
    rec, sim = time_leg(case, 20, 3, 5)
    rec["model"] = "air k-omega SST, MUSCL+minmod+AUSM+, central viscous, DPLUR x4, CFL 5, FP32 bulk"
    sim.close()
    if not a.no_cpu:
        rec["cpu_baseline"] = cpu_baseline(cases.sst_plate(), 3, "the same 256x128x64 block", plate=True)
    return rec


def leg_mixing_layer(a):
    """Finite-rate mixing layer (192x96x96, FP32 bulk + FP64 chemistry, EPI,
    DPLUR x4, CFL 0.2), then, on the solver's own state at iteration
    chem_eval_iter, the chemistry step of every cell by the plain FP64
    point-implicit integrator vs the surrogate with fallback."""
    import numpy as np

    from openstc_amd import cases, surrogate as S
    from openstc_amd import _lib as L

    case = cases.mixing_layer(seed=5)
    rec, sim = time_leg(case, 10, 2, 3)
    rec["model"] = ("9-species/19-reaction H2-air finite-rate (EPI), laminar, MUSCL+minmod+AUSM+, central "
                    "viscous, DPLUR x4, CFL %g, FP32 bulk + FP64 chemistry/Newton" % case.param.cfl)
    lib = sim.ctx.lib
    cells = rec["cells"]
    done = 2 + 10 + 3
    while done < a.chem_eval_iter:
        sim.step(output_screen=case.output_screen)
        done += 1
    sim.sync()
    z = sim.zones[0]
    dy_p = S.DeviceArray(nbytes=cells * 9 * 8)
    dy_s = S.DeviceArray(nbytes=cells * 9 * 8)

    def timed(fn):
        fn()
        L.check(lib.ostc_device_sync(), "sync")
        t0 = time.perf_counter()
        for _ in range(a.chem_reps):
            fn()
        L.check(lib.ostc_device_sync(), "sync")
        return (time.perf_counter() - t0) / a.chem_reps

    def zone_step(sur, mode, buf, st=None):
        L.check(lib.ostc_chem_zone_step(sim.ctx.h, z.h, sur.h if sur else None, mode, buf.p, st, None),
                "zone chemistry")

    t_plain = timed(lambda: zone_step(None, 0, dy_p))
    flop = plain_chem_flop_per_cell(case.param)
    rec["plain_chemistry"] = {
        "eval_iteration": a.chem_eval_iter, "ms": round(t_plain * 1e3, 4),
        "Mcell_s": round(cells / t_plain / 1e6, 2), "flop_per_cell": flop,
        "fp64_TFLOPs": round(flop * cells / t_plain / 1e12, 3), "fp64_peak_TFLOPs": FP64_PEAK_TFLOPS}
    ref = dy_p.get(np.float64, (cells, 9))

    def compare(weights):
        """the surrogate with fallback against the plain step on the same cells"""
        sur = S.Surrogate.load(sim.ctx, weights)
        t_sur = timed(lambda: zone_step(sur, 1, dy_s))
        stats = (C.c_longlong * 4)()
        zone_step(sur, 1, dy_s, stats)
        rep = S.error_report(dy_s.get(np.float64, (cells, 9)), ref, list(stats))
        act = "hard-swish" if sur.desc.activation == S.ACT_HSWISH else "GELU"
        sur.close()
        return {"weights": os.path.basename(weights), "activation": act,
                "plain_ms": round(t_plain * 1e3, 4), "surrogate_ms": round(t_sur * 1e3, 4),
                "speedup": round(t_plain / t_sur, 3), "surrogate_Mcell_s": round(cells / t_sur / 1e6, 2),
                "surrogate_TFLOPs_incl_fallback": round(SUR_FLOP_PER_CELL * cells / t_sur / 1e12, 2),
                "mfma_peak_TFLOPs_bf16": BF16_PEAK_TFLOPS,
                **{k: (round(v, 6) if isinstance(v, float) else v) for k, v in rep.items()}}

    weights = a.surrogate or S.DEFAULT_WEIGHTS
    if os.path.exists(weights):
        rec["surrogate"] = compare(weights)
        rec["surrogate"]["error_definition"] = (
            "rel_l2 = ||dY_sur - dY_plain||_2 / ||dY_plain||_2 over all cells and species (dY/dt is the effective "
            "production rate); cell_* = max_k|err_k| / max_k|dY_plain_k| over cells with max|dY_plain| > 1e-8")
    if not a.surrogate and os.path.exists(S.HSWISH_WEIGHTS):
        rec["surrogate_hswish_mixing_only"] = compare(S.HSWISH_WEIGHTS)
    for b in (dy_p, dy_s):
        b.free()
    sim.close()
    if not a.no_cpu:
        rec["cpu_baseline"] = cpu_baseline(cases.mixing_layer(mx=96, my=48, mz=48, seed=5), 2,
                                           "the mixing layer at 96x48x48 (same profile and seed)")
    return rec


LEGS = {"warped_box": ("1_warped_box", leg_warped_box), "sst_plate": ("2_sst_plate", leg_sst_plate),
        "mixing_layer": ("3_mixing_layer", leg_mixing_layer)}


def main():
    a = parse()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only (one block per GPU is the scale-out "
                                   "design; multi-GPU halos are not built)", "n_gpus": a.gpus}))
        sys.exit(2)
    from openstc_amd import cases

    out = {"metric": METRIC, "value": None, "unit": "Mcell-iter/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic"}
    if not a.no_headline:
        h = headline(a)
        out["value"] = h["value"]
        out["ms_per_step"] = h["ms_per_step"]
        out["config"] = {"workload": h["workload"], "blocks": h["blocks"], "block": h["block"], "cells": h["cells"],
                         "model": h["model"], "precision": h["precision"]}
        out["roofline"] = h.get("roofline")
        out["stages"] = h.get("stages")
        out["counters"] = h["counters"]
        out["setup_s"] = h["setup_s"]
        out["surrogate_at_scale"] = h["surrogate"]
        if not a.no_cpu:
            cx, cy, cz = (int(v) for v in a.cpu_block.split(","))
            log("cpu baseline (headline sample)")
            out["cpu_baseline"] = cpu_baseline(cases.jet_crossflow(mx=cx, my=cy, mz=cz, nblocks=4), a.cpu_steps,
                                               f"the same four-block jet workload at 4 x {cx}x{cy}x{cz}")
        else:
            out["cpu_baseline"] = None
    out["configs"] = {}
    for leg in [s for s in a.legs.split(",") if s]:
        key, fn = LEGS[leg]
        log(f"leg {key}")
        out["configs"][key] = fn(a)
    if a.profile_json:
        with open(a.profile_json, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

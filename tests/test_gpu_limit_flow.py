"""GPU parity of limit_flow's repair path (src/TimeAdvanceFunc.cu:34-224)
against the oracle, with the device repair counter.

Bad cells are planted on the host -- NaN, density/pressure/velocity outside
the setup limits, negative k and omega -- isolated and adjacent, with and
without valid neighbours (the reference bounds the i-neighbours by mz, so a
cell with i > mz averages nothing and takes the clamped state), then:

  * stand-alone: ``ostc_limit_flow`` (classification kernel + repair list)
    against ``orc_limit_flow`` on the same planted state;
  * in the time step: bad conserved variables planted in CV, so that the
    update produces the bad cells and the classification folded into the
    update kernel lists them; ``ostc_steady_step`` is called twice with the
    same step number (a caller that repeats a step) against two oracle steps.

Tolerances: FP64 zones 1e-13 scaled after the stand-alone repair (plain
averages, same summation order) and 1e-9 after a time step (the iteration
bound of the other FP64 tests); FP32 zones 2e-5 scaled (the state is copied
into the oracle after every planting, so only the step's own FP32 arithmetic
differs).  ``ostc_zone_counters``[1] must equal the oracle's count of cells
with a flow repair (``orc_limit_flow_total``), which in the first step is the
number planted.
"""
import ctypes as C

import numpy as np
import pytest

from harness import Pair, interior, scaled_err
from openstc_amd import _lib as L
from openstc_amd import cases

pytestmark = pytest.mark.gpu

STATE = ["BV", "SV", "CV", "VEL", "SOUND", "MUL", "LAMBDA", "MACH", "MUT", "LAMBDA_T", "GAMMA", "CP", "RHO_D"]


def say(*a):
    print("[gpu-limit]", *a, flush=True)


def at(blk_shape_g, i, j, k, g=2):
    return (k + g, j + g, i + g)


def plant_bv(pr: Pair, plants):
    """plants: [(field, plane, (i, j, k), value)] written into the GPU zone
    and the oracle (the oracle gets the GPU's stored values back, so FP32
    rounding of the planted state is the same on both sides)."""
    z = pr.zg()
    arrs = {}
    for field, _, _, _ in plants:
        if field not in arrs:
            arrs[field] = z.download(field)
    for field, l, (i, j, k), v in plants:
        arrs[field][(l,) + at(None, i, j, k)] = v
    for field, a in arrs.items():
        z.upload(field, a)
    pr.sync_to_oracle(STATE)


def compare(pr: Pair, tol, label):
    errs = {}
    ns = pr.case.param.n_spec
    for f in ("BV", "SV", "CV"):
        a = interior(pr.zg().download(f))
        o = interior(pr.oarr(f))
        assert np.array_equal(np.isnan(a), np.isnan(o)), (label, f, "NaN pattern differs")
        a = np.where(np.isnan(a), 0.0, a)
        o = np.where(np.isnan(o), 0.0, o)
        if f == "SV" and ns:
            # mass fractions: absolute; k, omega: scaled
            errs["SV:Y"] = float(np.abs(a[:ns] - o[:ns]).max())
            if a.shape[0] > ns:
                errs["SV:turb"] = scaled_err(a[ns:], o[ns:])
        else:
            errs[f] = scaled_err(a, o)
    say(label, {k: f"{v:.2e}" for k, v in errs.items()})
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, (label, bad)


def flow_plants(pr: Pair, mixture=False):
    """Planted BV/SV cells; returns (plants, number of flow repairs)."""
    p = pr.case.param
    ll, ul = list(p.limit_ll), list(p.limit_ul)
    mz = pr.case.blocks[0].mz
    plants = [
        ("BV", 0, (3, 4, 2), np.nan),              # NaN density, neighbours valid
        ("BV", 4, (5, 10, 3), -1.0),               # negative pressure ...
        ("BV", 4, (6, 10, 3), 0.5 * ll[4]),        # ... next to one below the limit
        ("BV", 0, (mz + 6, 12, 4), 2.0 * ul[0]),   # i > mz: no neighbour counted, clamped
        ("BV", 1, (0, 0, 0), 2.0 * ll[1]),         # corner cell, u below its limit
        ("BV", 3, (2, 3, 1), 2.0 * ul[3]),         # w above its limit, diagonal to the NaN cell
    ]
    return plants, 6


def turb_plants(pr: Pair):
    ns = pr.case.param.n_spec
    return [
        ("SV", ns, (4, 5, 6), -1.0),        # negative k
        ("SV", ns + 1, (5, 5, 6), -3.0),    # negative omega next to it
        ("SV", ns, (3, 8, 2), np.nan),      # NaN k with valid neighbours
        ("SV", ns + 1, (pr.case.blocks[0].mz + 8, 3, 3), -1.0),  # i > mz: freestream omega
        ("BV", 4, (8, 8, 8), -1.0),          # flow and ...
        ("SV", ns, (8, 8, 8), -1.0),         # ... turbulence repair in one cell
    ]


def standalone(case, precision, turb=False):
    pr = Pair(case, precision=precision)
    pr.init()
    pr.zg().counters(reset=True)
    plants, nflow = flow_plants(pr)
    if turb:
        plants += turb_plants(pr)
        nflow += 1
    plant_bv(pr, plants)
    L.check(pr.lib.ostc_limit_flow(pr.c, pr.zg().h, None), "limit_flow")
    n_orc = pr.orc.lib.orc_limit_flow(pr.zo(), pr.p)
    cnt = pr.zg().counters(reset=True)
    say("flow repairs: gpu", cnt[1], "oracle", n_orc, "planted", nflow)
    assert n_orc == nflow
    assert cnt[1] == nflow
    return pr


def test_limit_flow_standalone_air_fp64():
    case = cases.warped_box(mx=40, my=24, mz=8)
    say("warped_box 40x24x8 FP64, stand-alone limit_flow")
    pr = standalone(case, L.FP64)
    compare(pr, 1e-13, "air FP64")
    pr.close()


def test_limit_flow_standalone_sst_fp64_fp32():
    for prec, tol in ((L.FP64, 1e-13), (L.FP32, 2e-5)):
        case = cases.sst_plate(mx=48, my=32, mz=12, perturb=True)
        say("sst_plate 48x32x12", "FP64" if prec == L.FP64 else "FP32", "stand-alone limit_flow with k/omega")
        pr = standalone(case, prec, turb=True)
        compare(pr, tol, "SST")
        pr.close()


def test_limit_flow_standalone_mixture_fp64():
    case = cases.mixing_layer(mx=28, my=20, mz=10, cfl=0.2)
    say("mixing_layer 28x20x10 FP64, stand-alone limit_flow (9 species)")
    pr = standalone(case, L.FP64)
    compare(pr, 1e-13, "mixture FP64")
    pr.close()


def cv_plants(pr: Pair, mixture: bool):
    """Bad conserved variables that the update turns into bad cells."""
    cv = interior(pr.zg().download("CV"))
    bv = interior(pr.zg().download("BV"))
    p = pr.case.param
    ns = p.n_spec
    out = []
    if mixture:
        # velocity above its limit with the same internal energy (the Newton
        # solve for T then converges as usual)
        for (i, j, k) in ((5, 5, 5), (6, 5, 5)):
            rho, u = bv[0, k, j, i], bv[1, k, j, i]
            u2 = 2.0 * p.limit_ul[1]
            out.append(("CV", 1, (i, j, k), rho * u2))
            out.append(("CV", 4, (i, j, k), cv[4, k, j, i] + 0.5 * rho * (u2 * u2 - u * u)))
        out.append(("CV", 4, (12, 9, 3), np.nan))
        return out, 3
    out = [
        ("CV", 4, (5, 10, 3), -abs(cv[4, 3, 10, 5])),   # negative pressure after the update
        ("CV", 4, (6, 10, 3), -abs(cv[4, 3, 10, 6])),   # adjacent
        ("CV", 4, (20, 12, 4), np.nan),                 # NaN energy, i > mz
    ]
    if p.turb_method == L.RANS:
        out.append(("CV", ns + 5, (4, 5, 6), -abs(cv[ns + 5, 6, 5, 4])))  # negative rho*k
    return out, 3


def in_step(case, precision, tol, mixture=False):
    pr = Pair(case, precision=precision)
    pr.init()
    pr.zg().counters(reset=True)
    pr.orc.lib.orc_limit_flow_total(1)
    total = 0
    sums = (C.c_double * 4)()
    computed = C.c_int(0)
    for rep in range(2):
        plants, n = cv_plants(pr, mixture)
        plant_bv(pr, plants)
        total += n
        # the same step number twice: the zone alternates its repair list itself
        L.check(pr.lib.ostc_steady_step(pr.c, pr.gpu.zarr, 1, None, None, 2, 1 << 30, sums, C.byref(computed),
                                        None), "steady_step")
        pr.orc.step()
        cnt = pr.zg().counters()
        n_orc = pr.orc.lib.orc_limit_flow_total(0)
        say(f"step {rep + 1}: flow repairs gpu {cnt[1]}, oracle {n_orc}, planted so far {total}")
        # the first step repairs exactly the planted cells; in the second the
        # first step's repaired states perturb their neighbours, and more
        # cells leave the limits (the oracle counts the same cells)
        if rep == 0:
            assert n_orc == total
        assert cnt[1] == n_orc
        compare(pr, tol, f"after step {rep + 1}")
        if precision == L.FP32:
            pr.sync_to_oracle(STATE)
    pr.close()


def test_limit_flow_in_step_sst_fp64():
    case = cases.sst_plate(mx=48, my=32, mz=12, perturb=True)
    say("sst_plate 48x32x12 FP64: repairs from the update-folded classification, 2 steps")
    in_step(case, L.FP64, 1e-9)


def test_limit_flow_in_step_sst_fp32():
    case = cases.sst_plate(mx=48, my=32, mz=12, perturb=True)
    say("sst_plate 48x32x12 FP32: repairs from the update-folded classification, 2 steps")
    in_step(case, L.FP32, 2e-5)


def test_limit_flow_in_step_mixture_fp64():
    case = cases.mixing_layer(mx=28, my=20, mz=10, cfl=0.2)
    say("mixing_layer 28x20x10 FP64 (finite rate): repairs in the time step, 2 steps")
    in_step(case, L.FP64, 1e-9, mixture=True)

"""GPU parity of the finite-rate chemistry path (FP64 rates, EPI, DPLUR with
the chemistry block, T-from-e Newton) against the FP64 oracle.

Tolerances: FP64 point chemistry 1e-10 relative (pow/exp/log ulp budgets of
the device math library vs glibc); FP64 zone stages 1e-10 scaled; FP32-bulk
zone stages 1e-4 scaled with the GPU state copied into the oracle first
(mixture thermodynamics in the flux are FP32; chemistry and Newton FP64)."""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi as O
from harness import Pair
from openstc_amd import _lib as L
from openstc_amd import cases
from test_gpu_parity import stage_check

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-chem]", *a, flush=True)


class Dev:
    """Device copy of a host float64 array."""

    def __init__(self, a: np.ndarray):
        self.lib = L.load()
        self.a = np.ascontiguousarray(a, dtype=np.float64)
        self.p = C.c_void_p()
        L.check(self.lib.ostc_device_alloc(C.byref(self.p), max(8, self.a.nbytes)), "alloc")
        L.check(self.lib.ostc_memcpy_h2d(self.p, self.a.ctypes.data_as(C.c_void_p), self.a.nbytes), "h2d")

    def get(self):
        out = np.empty_like(self.a)
        L.check(self.lib.ostc_device_sync(), "sync")
        L.check(self.lib.ostc_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes), "d2h")
        return out

    def free(self):
        self.lib.ostc_device_free(self.p)


def states(n, seed):
    rng = np.random.default_rng(seed)
    t = rng.uniform(600, 2800, n)
    y = rng.dirichlet(np.ones(9) * 0.6, n)
    y = np.maximum(y, 1e-8)
    y /= y.sum(1, keepdims=True)
    p = rng.uniform(0.5, 3.0, n) * 101325.0
    dt = 10 ** rng.uniform(-9, -5, n)
    return t, p, y, dt


def test_chem_source_and_plain_step_vs_oracle():
    mech, _ = cases.load_mechanism(True)
    param = cases.mixture_param(mech, False, True, 4, 1.0, 1,
                                cases.InflowState(1.0, 1.0, 0, 0, 1e5, 1000.0, 1.0, 0.0, [0.0] * 9))
    from openstc_amd.runtime import Context

    ctx = Context(param)
    n = 4096
    t, p, y, dt = states(n, 11)
    mw = np.array(mech.mw[:9])
    rho = p / (8314.462618 * t * np.sum(y / mw, 1))
    say("source batch n=%d" % n)
    dr, dtm, dy_ = Dev(rho), Dev(t), Dev(y)
    dom, djac = Dev(np.zeros((n, 9))), Dev(np.zeros((n, 81)))
    L.check(ctx.lib.ostc_chem_source_batch(ctx.h, n, dr.p, dtm.p, dy_.p, dom.p, djac.p, None), "source_batch")
    om_g, jac_g = dom.get(), djac.get()
    op = O.copy_struct(O.OrcParam, param)
    lib = O.load()
    err_om = err_j = 0.0
    for c in range(0, n, 7):
        om, jac = np.zeros(20), np.zeros(400)
        yy = np.ascontiguousarray(np.pad(y[c], (0, 11)))
        lib.orc_chem_source_point(C.byref(op), rho[c], t[c], O.dp(yy), O.dp(om), None, O.dp(jac), None, None, None)
        s = np.abs(om[:9]).max()
        err_om = max(err_om, np.abs(om_g[c] - om[:9]).max() / s)
        sj = np.abs(jac[:81]).max()
        err_j = max(err_j, np.abs(jac_g[c] - jac[:81]).max() / sj)
    say("omega err", err_om, "jac err", err_j)
    assert err_om < 1e-10 and err_j < 1e-10
    say("plain step batch")
    dp_, ddt, ddy = Dev(p), Dev(dt), Dev(np.zeros((n, 9)))
    L.check(ctx.lib.ostc_chem_plain_step(ctx.h, n, dtm.p, dp_.p, dy_.p, ddt.p, ddy.p, None), "plain")
    dy_g = ddy.get()
    dy_o = np.zeros((n, 9))
    for c in range(n):
        yy = np.ascontiguousarray(np.pad(y[c], (0, 11)))
        out = np.zeros(20)
        lib.orc_plain_chem_step(C.byref(op), t[c], p[c], O.dp(yy), dt[c], O.dp(out))
        dy_o[c] = out[:9]
    e = np.abs(dy_g - dy_o).max(1) / (np.abs(dy_o).max(1) + 1e-300)
    say("plain step max rel err", e.max())
    assert e.max() < 1e-9
    for d in (dr, dtm, dy_, dom, djac, dp_, ddt, ddy):
        d.free()
    ctx.close()


def test_mixing_layer_nonreacting_mixture_fp64():
    """The 9-species mixture without reactions (mix_model = MIXTURE): species
    transport with DPLUR, whose first sweep forms dq0 itself -- stages and 3
    iterations against the oracle."""
    case = cases.mixing_layer(mx=28, my=20, mz=10, cfl=0.2)
    case.param.mix_model = L.MIXTURE
    say("mixing_layer 28x20x10 FP64 non-reacting mixture, DPLUR, stages + 3 steps")
    pr = Pair(case, precision=L.FP64)
    pr.init()
    stage_check(pr, 1e-10, sync32=False)
    for _ in range(3):
        pr.gpu.step()
        pr.orc.step()
    for f in ("BV", "SV"):
        e = pr.err(f, halo=2)
        say("after the stages and 3 steps", f, e)
        assert e < 1e-8, (f, e)
    pr.close()


@pytest.mark.parametrize("name,chem_method", [("mixing_layer", 1), ("mixing_layer", 2), ("sst_plate", 0)])
def test_explicit_point_implicit_fp64(name, chem_method):
    """Explicit time stepping with the point-implicit treatment of the stiff
    sources (src/ImplicitTreatmentHPP.cuh:10-59): EPI (chemSrcMethod 1) and
    DA (2) chemistry on the finite-rate mixing layer, the SST rows on the
    flat plate -- stages and 3 iterations against the oracle."""
    if name == "mixing_layer":
        case = cases.mixing_layer(mx=28, my=20, mz=10, cfl=0.2, chem_method=chem_method)
    else:
        case = cases.sst_plate(mx=48, my=32, mz=12, cfl=0.5, perturb=True)
    case.param.implicit_method = 0
    case.param.dplur_inner_step = 0
    say(f"{name} FP64 explicit + point-implicit (chem method {chem_method}), stages + 3 steps")
    pr = Pair(case, precision=L.FP64)
    pr.init()
    stage_check(pr, 1e-10, sync32=False)
    for _ in range(3):
        pr.gpu.step()
        pr.orc.step()
    for f in ("BV", "SV"):
        e = pr.err(f, halo=2)
        say("after the stages and 3 steps", f, e)
        assert e < 1e-8, (f, e)
    pr.close()


def test_mixing_layer_fp64_stages():
    case = cases.mixing_layer(mx=28, my=20, mz=10)
    say("mixing_layer 28x20x10 FP64 stages (FR, EPI, DPLUR)")
    pr = Pair(case, precision=L.FP64)
    pr.init()
    for f in ("BV", "SV", "CV", "MUL", "GAMMA", "RHO_D"):
        e = pr.err(f, halo=1)
        say("init", f, e)
        assert e < 1e-11, (f, e)
    stage_check(pr, 1e-10, sync32=False)
    pr.close()


def test_mixing_layer_fp64_iterations():
    case = cases.mixing_layer(mx=28, my=20, mz=10, cfl=0.2)
    say("mixing_layer 28x20x10 FP64 4 steps")
    pr = Pair(case, precision=L.FP64)
    pr.init()
    for s in range(4):
        pr.gpu.step()
        pr.orc.step()
    for f in ("BV", "SV"):
        e = pr.err(f, halo=2)
        say(f, "err after 4 steps", e)
        assert e < 1e-8, (f, e)
    pr.close()


def test_mixing_layer_fp32_stages():
    case = cases.mixing_layer(mx=28, my=20, mz=10)
    say("mixing_layer 28x20x10 FP32 stages")
    pr = Pair(case, precision=L.FP32)
    pr.init()
    stage_check(pr, 1e-4, sync32=True)
    pr.close()

"""ostc_residual_sums_noalloc (the entry point that allocates nothing) against
ostc_residual_sums on the same zones: two calls in a row, the second over a
smaller input, each bit-identical to the allocating entry point (same kernels,
same summation order); a request beyond the scratch array is refused before
anything is launched."""
import ctypes as C

import numpy as np
import pytest

from openstc_amd import _lib as L
from openstc_amd import cases
from openstc_amd.runtime import Simulation, Zone

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-noalloc]", *a, flush=True)


def sums_alloc(sim, zarr, n):
    out = (C.c_double * 4)()
    L.check(sim.ctx.lib.ostc_residual_sums(sim.ctx.h, zarr, n, out, None), "residual_sums")
    return np.array(out[:], dtype=np.float64)


def sums_noalloc(sim, zarr, n, dbuf):
    lib = sim.ctx.lib
    L.check(lib.ostc_residual_sums_noalloc(sim.ctx.h, zarr, n, dbuf, None), "residual_sums_noalloc")
    L.check(lib.ostc_device_sync(), "sync")
    host = np.zeros(4, dtype=np.float64)
    L.check(lib.ostc_memcpy_d2h(host.ctypes.data_as(C.c_void_p), dbuf, 32), "d2h")
    return host


@pytest.mark.parametrize("precision", [L.FP64, L.FP32])
def test_residual_sums_noalloc_matches(precision):
    case = cases.warped_box_split(mx=64, my=48, mz=8, steps=3, precision=precision)
    say("warped_box_split 64x48x8 in", len(case.blocks), "blocks", "FP64" if precision == L.FP64 else "FP32")
    sim = Simulation(case)
    sim.initialize_computation()
    lib = sim.ctx.lib
    for _ in range(2):
        sim.step()
    for z in sim.zones:
        L.check(lib.ostc_store_last_step(sim.ctx.h, z.h, None), "store_last_step")
    sim.step()
    sim.sync()
    dbuf = C.c_void_p()
    L.check(lib.ostc_device_alloc(C.byref(dbuf), 64), "alloc")
    try:
        n = len(sim.zones)
        # first call: every zone
        ref = sums_alloc(sim, sim.zarr, n)
        got = sums_noalloc(sim, sim.zarr, n, dbuf)
        say("all zones: alloc", ref, "noalloc", got)
        assert np.all(ref > 0)
        assert np.array_equal(ref, got)
        # second call in a row: a smaller input (the first zone only); the
        # scratch array still holds the first call's partial sums
        one = (C.c_void_p * 1)(sim.zones[0].h)
        ref1 = sums_alloc(sim, one, 1)
        got1 = sums_noalloc(sim, one, 1, dbuf)
        say("first zone: alloc", ref1, "noalloc", got1)
        assert np.array_equal(ref1, got1)
        assert np.all(ref1 < ref)
    finally:
        lib.ostc_device_free(dbuf)
        sim.close()


def test_residual_sums_noalloc_rejects_oversize():
    """262145 one-tile zones need 4 doubles more than the scratch array holds:
    refused on the host, nothing launched."""
    case = cases.warped_box(mx=8, my=8, mz=4)
    sim = Simulation(case)
    lib = sim.ctx.lib
    tiny = Zone(sim.ctx, 2, 2, 1, L.FP64, 2, 0)
    dbuf = C.c_void_p()
    L.check(lib.ostc_device_alloc(C.byref(dbuf), 64), "alloc")
    try:
        n = (1 << 18) + 1
        zarr = (C.c_void_p * n)(*([tiny.h] * n))
        r = lib.ostc_residual_sums_noalloc(sim.ctx.h, zarr, n, dbuf, None)
        say("oversize request:", r, lib.ostc_last_error().decode())
        assert r == L.ERR_ARG
        ok = (C.c_void_p * 1)(tiny.h)
        assert lib.ostc_residual_sums_noalloc(sim.ctx.h, ok, 1, dbuf, None) == 0
        L.check(lib.ostc_device_sync(), "sync")
    finally:
        lib.ostc_device_free(dbuf)
        tiny.close()
        sim.close()

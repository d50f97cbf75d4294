"""Python handles over the C-ABI: contexts, zones and the steady time loop.

``Simulation`` follows the reference Driver: ``initialize_computation``
(src/Driver.cu:43-104) then ``steady_step`` per iteration
(src/Driver.cu:218-287), with the residual normalisation of
src/Driver.cu:291-380.  All compute runs in libostc.so on the GPU.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Context:
    def __init__(self, param: L.Param):
        self.lib = L.load()
        self.param = param
        h = C.c_void_p()
        L.check(self.lib.ostc_ctx_create(C.byref(param), C.byref(h)), "ostc_ctx_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.ostc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Zone:
    def __init__(self, ctx: Context, mx: int, my: int, mz: int, precision: int = L.FP64, ngg: int = 2,
                 block_id: int = 0):
        self.ctx, self.lib = ctx, ctx.lib
        self.mx, self.my, self.mz, self.ngg = mx, my, mz, ngg
        self.precision = precision
        d = L.ZoneDesc(mx, my, mz, ngg, precision, block_id)
        h = C.c_void_p()
        L.check(self.lib.ostc_zone_create(ctx.h, C.byref(d), C.byref(h)), "ostc_zone_create")
        self.h = h
        self.shape = (mz + 2 * ngg, my + 2 * ngg, mx + 2 * ngg)
        self.plane = int(np.prod(self.shape))

    def planes(self, field: str) -> int:
        n = C.c_int()
        L.check(self.lib.ostc_zone_field_planes(self.h, L.F[field], C.byref(n)), "field_planes")
        return n.value

    def upload(self, field: str, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        n = a.size // self.plane
        assert n * self.plane == a.size, (field, a.shape)
        L.check(self.lib.ostc_zone_upload(self.h, L.F[field], _dptr(a), n), f"upload {field}")

    def download(self, field: str, n_planes: int | None = None) -> np.ndarray:
        n = self.planes(field) if n_planes is None else n_planes
        out = np.empty((n,) + self.shape, dtype=np.float64)
        if n:
            L.check(self.lib.ostc_zone_download(self.h, L.F[field], _dptr(out), n), f"download {field}")
        return out

    def set_bc(self, bc: L.BC):
        self._bc = bc
        L.check(self.lib.ostc_zone_set_bc(self.h, C.byref(bc)), "ostc_zone_set_bc")

    def counters(self, reset=False):
        out = (C.c_longlong * 4)()
        L.check(self.lib.ostc_zone_counters(self.h, out, 1 if reset else 0, None), "counters")
        return list(out)

    def close(self):
        if self.h:
            self.lib.ostc_zone_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def block_geometry(x: np.ndarray, y: np.ndarray, z: np.ndarray, ngg: int = 2):
    """Ghost coordinates + metrics of one block on the host (libostc,
    bit-exact with the reference's operation order).  x/y/z: (mz,my,mx)."""
    lib = L.load()
    mz, my, mx = x.shape
    shape = (mz + 2 * ngg, my + 2 * ngg, mx + 2 * ngg)
    xo, yo, zo = (np.zeros(shape) for _ in range(3))
    jac = np.zeros(shape)
    met = np.zeros((9,) + shape)
    nneg = C.c_int()
    xs, ys, zs = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    L.check(lib.ostc_block_geometry(mx, my, mz, ngg, _dptr(xs), _dptr(ys), _dptr(zs), _dptr(xo), _dptr(yo),
                                    _dptr(zo), _dptr(jac), _dptr(met), C.byref(nneg)), "ostc_block_geometry")
    return xo, yo, zo, jac, met, nneg.value


def register_inner_face(raw, ngg: int = 2, dim: int = 3) -> L.InnerFace:
    """One record of the reference's inner{rank}.txt -> registered face
    (libostc; src/Mesh.cpp:44-178)."""
    lib = L.load()
    r = (C.c_int * 13)(*raw)
    out = L.InnerFace()
    L.check(lib.ostc_register_inner_face(r, ngg, dim, C.byref(out)), "ostc_register_inner_face")
    return out


def mesh_geometry(blocks, ngg: int = 2):
    """Ghost grids (incl. across inner faces) + metrics of all blocks on the
    host (libostc).  blocks: objects with mx, my, mz, x, y, z, inner.
    Returns [(xg, yg, zg, jac, metric, n_neg)] per block."""
    lib = L.load()
    n = len(blocks)
    keep, outs = [], []
    arr = (L.BlockCoords * n)()
    for b, blk in enumerate(blocks):
        shape = (blk.mz + 2 * ngg, blk.my + 2 * ngg, blk.mx + 2 * ngg)
        xs, ys, zs = (np.ascontiguousarray(a, dtype=np.float64) for a in (blk.x, blk.y, blk.z))
        xo, yo, zo, jac = (np.zeros(shape) for _ in range(4))
        met = np.zeros((9,) + shape)
        keep += [xs, ys, zs]
        outs.append([xo, yo, zo, jac, met])
        c = arr[b]
        c.mx, c.my, c.mz = blk.mx, blk.my, blk.mz
        c.x, c.y, c.z = xs.ctypes.data, ys.ctypes.data, zs.ctypes.data
        c.x_out, c.y_out, c.z_out = xo.ctypes.data, yo.ctypes.data, zo.ctypes.data
        c.jac, c.metric = jac.ctypes.data, met.ctypes.data
    faces = [f for blk in blocks for f in blk.inner]
    inner = (L.InnerFace * max(1, len(faces)))(*faces)
    n_inner = (C.c_int * n)(*[len(blk.inner) for blk in blocks])
    L.check(lib.ostc_mesh_geometry(n, ngg, arr, inner, n_inner), "ostc_mesh_geometry")
    return [tuple(o) + (arr[b].n_neg_jac,) for b, o in enumerate(outs)]


class Simulation:
    """One multi-block steady case on the GPU (Driver<mix,turb> equivalent)."""

    def __init__(self, case, precision: int | None = None):
        self.case = case
        self.ctx = Context(case.param)
        prec = precision if precision is not None else case.precision
        self.zones = []
        for b, blk in enumerate(case.blocks):
            z = Zone(self.ctx, blk.mx, blk.my, blk.mz, prec, case.ngg, b)
            z.upload("X", blk.xg)
            z.upload("Y", blk.yg)
            z.upload("Z", blk.zg)
            z.upload("JAC", blk.jac)
            z.upload("METRIC", blk.metric)
            z.upload("BV", blk.bv0)
            if case.param.n_scalar:
                z.upload("SV", blk.sv0)
            z.set_bc(blk.bc)
            self.zones.append(z)
        self.zarr = (C.c_void_p * len(self.zones))(*[z.h for z in self.zones])
        n_inner = [len(blk.inner) for blk in case.blocks]
        self.n_inner = (C.c_int * len(n_inner))(*n_inner)
        flat = [f for blk in case.blocks for f in blk.inner]
        self.inner = (L.InnerFace * max(1, len(flat)))(*flat) if flat else None
        self.res_scale = [1.0] * 4
        self.step_no = 0
        self.n_grid = sum(b.mx * b.my * b.mz for b in case.blocks)

    # src/Driver.cu:43-104
    def initialize_computation(self):
        lib, c = self.ctx.lib, self.ctx.h
        p = self.case.param
        if p.turb_method == L.RANS:
            wp = self.case.wall_points
            for z in self.zones:
                if wp is not None and len(wp):
                    a = np.ascontiguousarray(wp, dtype=np.float64)
                    L.check(lib.ostc_wall_distance(c, z.h, _dptr(a), a.shape[0], None), "wall_distance")
                else:
                    z.upload("WALL_DIST", np.full((1,) + z.shape, 1e3))
        for z in self.zones:
            if z.mz == 1:  # 2-D blocks (src/Driver.cu:59-65)
                L.check(lib.ostc_eliminate_k_gradient(c, z.h, None), "eliminate_k_gradient")
        for z in self.zones:
            L.check(lib.ostc_apply_bc(c, z.h, None), "apply_bc")
        for z in self.zones:
            L.check(lib.ostc_compute_cv_from_bv(c, z.h, None), "cv_from_bv")
            if p.turb_method == L.RANS:
                L.check(lib.ostc_initialize_mut(c, z.h, None), "initialize_mut")
        self._inner_halo()
        for z in self.zones:
            L.check(lib.ostc_update_physical_properties(c, z.h, None), "properties")
        L.check(lib.ostc_device_sync(), "sync")

    def _inner_halo(self):
        if self.inner is None:
            return
        off = 0
        for b, blk in enumerate(self.case.blocks):
            for f in range(len(blk.inner)):
                L.check(self.ctx.lib.ostc_inner_halo(self.ctx.h, self.zarr, b, C.byref(self.inner[off + f]), None),
                        "inner_halo")
            off += len(blk.inner)

    def step(self, output_screen: int = 1 << 30):
        """One iteration; returns the max normalised residual or None."""
        self.step_no += 1
        sums = (C.c_double * 4)()
        computed = C.c_int(0)
        L.check(self.ctx.lib.ostc_steady_step(self.ctx.h, self.zarr, len(self.zones), self.inner, self.n_inner,
                                              self.step_no, output_screen, sums, C.byref(computed), None),
                "ostc_steady_step")
        if not computed.value:
            return None
        res = [math.sqrt(s / self.n_grid) for s in sums]
        if self.step_no == 1:
            self.res_scale = [r if r >= 1e-20 else 1e-20 for r in res]
        res = [r / s for r, s in zip(res, self.res_scale)]
        err = res[0]
        for r in res[1:]:
            if abs(r) > err:
                err = r
        if math.isnan(err):
            raise L.OstcError(f"NaN residual at step {self.step_no}")
        return err

    def sync(self):
        L.check(self.ctx.lib.ostc_device_sync(), "sync")

    def close(self):
        for z in self.zones:
            z.close()
        self.ctx.close()

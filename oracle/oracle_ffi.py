"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product.  The structures below are
declared independently of the product's binding; a CPU test checks that the
two layouts agree field by field.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import math
import os
import subprocess

import numpy as np

MAX_SPEC = 20
MAX_REAC = 50
MAX_FACE = 32
MAX_BCDATA = 8

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrcParam(C.Structure):
    _fields_ = [
        ("mix_model", C.c_int), ("turb_method", C.c_int),
        ("inviscid_scheme", C.c_int), ("reconstruction", C.c_int), ("limiter", C.c_int),
        ("viscous_scheme", C.c_int), ("rans_model", C.c_int), ("turb_implicit", C.c_int),
        ("chem_src_method", C.c_int), ("implicit_method", C.c_int), ("dplur_inner_step", C.c_int),
        ("n_spec", C.c_int), ("n_scalar", C.c_int), ("n_reac", C.c_int), ("n_var", C.c_int),
        ("Pr", C.c_double), ("cfl", C.c_double), ("Sc", C.c_double), ("Prt", C.c_double), ("Sct", C.c_double),
        ("mw", C.c_double * MAX_SPEC), ("t_low", C.c_double * MAX_SPEC), ("t_mid", C.c_double * MAX_SPEC),
        ("t_high", C.c_double * MAX_SPEC),
        ("low_temp_coeff", C.c_double * (MAX_SPEC * 7)), ("high_temp_coeff", C.c_double * (MAX_SPEC * 7)),
        ("LJ_potent_inv", C.c_double * MAX_SPEC), ("vis_coeff", C.c_double * MAX_SPEC),
        ("WjDivWi_to_One4th", C.c_double * (MAX_SPEC * MAX_SPEC)),
        ("sqrt_WiDivWjPl1Mul8", C.c_double * (MAX_SPEC * MAX_SPEC)),
        ("reac_type", C.c_int * MAX_REAC), ("reac_order", C.c_int * MAX_REAC),
        ("stoi_f", C.c_int * (MAX_REAC * MAX_SPEC)), ("stoi_b", C.c_int * (MAX_REAC * MAX_SPEC)),
        ("A", C.c_double * MAX_REAC), ("b", C.c_double * MAX_REAC), ("Ea", C.c_double * MAX_REAC),
        ("A2", C.c_double * MAX_REAC), ("b2", C.c_double * MAX_REAC), ("Ea2", C.c_double * MAX_REAC),
        ("third_body_coeff", C.c_double * (MAX_REAC * MAX_SPEC)),
        ("troe_alpha", C.c_double * MAX_REAC), ("troe_t3", C.c_double * MAX_REAC),
        ("troe_t1", C.c_double * MAX_REAC), ("troe_t2", C.c_double * MAX_REAC),
        ("limit_ll", C.c_double * 7), ("limit_ul", C.c_double * 7), ("sv_inf", C.c_double * (MAX_SPEC + 2)),
    ]


class OrcFace(C.Structure):
    _fields_ = [("range_start", C.c_int * 3), ("range_end", C.c_int * 3), ("face", C.c_int),
                ("direction", C.c_int), ("kind", C.c_int), ("slot", C.c_int)]


class OrcInflow(C.Structure):
    _fields_ = [("density", C.c_double), ("u", C.c_double), ("v", C.c_double), ("w", C.c_double),
                ("pressure", C.c_double), ("temperature", C.c_double), ("mut", C.c_double),
                ("sv", C.c_double * (MAX_SPEC + 2))]


class OrcWall(C.Structure):
    _fields_ = [("isothermal", C.c_int), ("temperature", C.c_double)]


class OrcBC(C.Structure):
    _fields_ = [("n_face", C.c_int), ("face", OrcFace * MAX_FACE), ("inflow", OrcInflow * MAX_BCDATA),
                ("wall", OrcWall * MAX_BCDATA)]


class OrcInnerFace(C.Structure):
    _fields_ = [("range_start", C.c_int * 3), ("range_end", C.c_int * 3), ("face", C.c_int),
                ("direction", C.c_int), ("target_start", C.c_int * 3), ("target_end", C.c_int * 3),
                ("target_block", C.c_int), ("target_face", C.c_int), ("target_direction", C.c_int),
                ("src_tar", C.c_int * 3), ("loop_dir", C.c_int * 3), ("target_loop_dir", C.c_int * 3),
                ("n_point", C.c_int * 3)]


_DP = C.POINTER(C.c_double)
ZONE_ARRAYS = ["x", "y", "z", "jac", "metric", "wall_distance", "bv", "sv", "cv", "bv_last", "vel",
               "acoustic_speed", "mach", "mul", "thermal_conductivity", "rho_D", "gamma", "cp", "chem_src_jac",
               "mut", "turb_therm_cond", "turb_src_jac", "dq", "dq0", "dqk", "inv_spectr_rad", "visc_spectr_rad",
               "dt_local"]


class OrcZone(C.Structure):
    _fields_ = ([("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int), ("ngg", C.c_int), ("nx", C.c_int),
                 ("ny", C.c_int), ("nz", C.c_int), ("plane", C.c_long)]
                + [(n, _DP) for n in ZONE_ARRAYS])


class OrcSurDesc(C.Structure):
    _fields_ = [("n_out", C.c_int), ("kind", C.c_int), ("n_elem", C.c_int), ("activation", C.c_int),
                ("in_mean", C.c_float * 12), ("in_std", C.c_float * 12),
                ("out_mean", C.c_float * 16), ("out_std", C.c_float * 16),
                ("loss_w", C.c_float * 16), ("elem_w", C.c_float * 36),
                ("floor", C.c_float), ("t_min", C.c_float), ("t_max", C.c_float), ("p_min", C.c_float),
                ("p_max", C.c_float), ("dt_min", C.c_float), ("dt_max", C.c_float),
                ("y_tol", C.c_float), ("elem_tol", C.c_float), ("elem_abs", C.c_float), ("dy_max", C.c_float)]


def build(quiet=True):
    """Compile the oracle with its Makefile (gcc)."""
    r = subprocess.run(["make", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _bind(C.CDLL(LIB_PATH))
    return _lib


@contextlib.contextmanager
def use_library(path: str):
    """Bind load() to another build of the oracle for the duration of the
    block (the planted-mistake tests build mutated copies of the oracle)."""
    global _lib
    saved = _lib
    try:
        _lib = _bind(C.CDLL(path))
        yield _lib
    finally:
        _lib = saved


def _bind(lib):
    """Declare the signatures of the oracle's entry points on a loaded library."""
    P, I, D = C.c_void_p, C.c_int, C.c_double
    Z = C.POINTER(OrcZone)
    Pp = C.POINTER(OrcParam)
    sig = {
        "orc_ghost_grid": (None, [I, I, I, I, _DP, _DP, _DP]),
        "orc_ghost_faces": (None, [I, I, I, I, _DP, _DP, _DP]),
        "orc_ghost_edges": (None, [I, I, I, I, _DP, _DP, _DP]),
        "orc_jac_metric": (I, [I, I, I, I, _DP, _DP, _DP, _DP, _DP, C.POINTER(I)]),
        "orc_trim_ghost": (None, [I, I, I, I, _DP, _DP]),
        "orc_inner_ghost_grid": (None, [C.POINTER(OrcInnerFace), I, I, I, I, _DP, _DP, _DP, I, I, I, _DP, _DP,
                                        _DP]),
        "orc_enthalpy": (None, [D, _DP, Pp]),
        "orc_enthalpy_and_cp": (None, [D, _DP, _DP, Pp]),
        "orc_cp": (None, [D, _DP, Pp]),
        "orc_gibbs_div_rt": (None, [D, Pp, _DP]),
        "orc_sutherland": (D, [D]),
        "orc_transport_point": (None, [Pp, D, D, _DP, _DP, _DP, _DP, _DP]),
        "orc_mixture_viscosity": (D, [Pp, D, D, _DP]),
        "orc_compute_cv_from_bv": (None, [Z, Pp]),
        "orc_update_physical_properties": (None, [Z, Pp]),
        "orc_initialize_mut": (None, [Z, Pp]),
        "orc_compute_source": (None, [Z, Pp]),
        "orc_inviscid_flux": (None, [Z, Pp]),
        "orc_inviscid_flux_dir": (None, [Z, Pp, I]),
        "orc_ausm_face": (None, [Pp, _DP, _DP, _DP, D, D, _DP]),
        "orc_viscous_flux": (None, [Z, Pp]),
        "orc_local_time_step": (None, [Z, Pp]),
        "orc_implicit_treatment": (None, [Z, Pp]),
        "orc_jac_times_dq_point": (None, [Z, Pp, I, I, I, I, D, _DP]),
        "orc_update_cv_and_bv": (I, [Z, Pp]),
        "orc_limit_flow": (I, [Z, Pp]),
        "orc_limit_flow_total": (C.c_long, [I]),
        "orc_apply_bc": (None, [Z, Pp, C.POINTER(OrcBC)]),
        "orc_inner_communication": (None, [Z, I, C.POINTER(OrcInnerFace), Pp]),
        "orc_eliminate_k_gradient": (None, [Z, Pp]),
        "orc_store_last_step": (None, [Z]),
        "orc_square_of_dbv": (None, [Z]),
        "orc_sum_dbv": (None, [Z, _DP]),
        "orc_wall_distance": (None, [Z, _DP, I]),
        "orc_steady_step": (D, [Z, I, Pp, C.POINTER(OrcBC), C.POINTER(OrcInnerFace), C.POINTER(I), I, I, _DP,
                                C.c_long, C.POINTER(I)]),
        "orc_chem_source_point": (None, [Pp, D, D, _DP, _DP, _DP, _DP, _DP, _DP, _DP]),
        "orc_solve_lu": (None, [_DP, _DP, I]),
        "orc_plain_chem_step": (None, [Pp, D, D, _DP, D, _DP]),
        "orc_plain_chem_batch": (None, [Pp, C.c_long, _DP, _DP, _DP, _DP, _DP]),
        "orc_newton_temperature": (I, [Pp, _DP, D, D, _DP]),
        "orc_omp_threads": (I, []),
        "orc_sur_n_params": (I, [I]),
        "orc_bf16": (D, [D]),
        "orc_sur_features": (None, [C.POINTER(OrcSurDesc), D, D, _DP, D, C.POINTER(C.c_float)]),
        "orc_sur_forward": (None, [C.POINTER(OrcSurDesc), C.POINTER(C.c_float), I, C.c_long, _DP, _DP, _DP,
                                   _DP, _DP]),
        "orc_sur_loss_grad": (D, [C.POINTER(OrcSurDesc), C.POINTER(C.c_float), I, C.c_long, _DP, _DP, _DP,
                                  _DP, C.POINTER(C.c_float), _DP]),
        "orc_sur_decode_check": (I, [C.POINTER(OrcSurDesc), C.POINTER(C.c_float), D, D, _DP, D, _DP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def copy_struct(dst_type, src):
    """Field-by-field copy between layout-identical ctypes structures."""
    dst = dst_type()
    C.memmove(C.addressof(dst), C.addressof(src), C.sizeof(dst_type))
    return dst


class OracleBlock:
    """Numpy-backed oracle zone (every array padded by ngg)."""

    def __init__(self, mx, my, mz, ngg, param: OrcParam):
        self.mx, self.my, self.mz, self.ngg = mx, my, mz, ngg
        self.shape = (mz + 2 * ngg, my + 2 * ngg, mx + 2 * ngg)
        p = param
        ns, nv, nsc = p.n_spec, p.n_var, p.n_scalar
        counts = dict(x=1, y=1, z=1, jac=1, metric=9, wall_distance=1, bv=6, sv=max(nsc, 1), cv=nv, bv_last=4,
                      vel=1, acoustic_speed=1, mach=1, mul=1, thermal_conductivity=1, rho_D=max(ns, 1), gamma=1,
                      cp=1, chem_src_jac=max(ns * ns, 1), mut=1, turb_therm_cond=1, turb_src_jac=2, dq=nv,
                      dq0=nv, dqk=nv, inv_spectr_rad=3, visc_spectr_rad=1, dt_local=1)
        self.arr = {k: np.zeros((n,) + self.shape) for k, n in counts.items()}
        self.s = OrcZone()
        self.s.mx, self.s.my, self.s.mz, self.s.ngg = mx, my, mz, ngg
        self.s.nx, self.s.ny, self.s.nz = self.shape[2], self.shape[1], self.shape[0]
        self.s.plane = int(np.prod(self.shape))
        for k in ZONE_ARRAYS:
            setattr(self.s, k, dp(self.arr[k]))

    def ghost_grid(self, x, y, z, edges=True):
        """Padded (ngg+1) coordinates with the single-block ghost grid (the
        face layers only with edges=False)."""
        lib = load()
        mx, my, mz, g = self.mx, self.my, self.mz, self.ngg
        G = g + 1
        big = (mz + 2 * G, my + 2 * G, mx + 2 * G)
        X, Y, Z = (np.zeros(big) for _ in range(3))
        X[G:G + mz, G:G + my, G:G + mx] = x
        Y[G:G + mz, G:G + my, G:G + mx] = y
        Z[G:G + mz, G:G + my, G:G + mx] = z
        if edges:
            lib.orc_ghost_grid(mx, my, mz, g, dp(X), dp(Y), dp(Z))
        else:
            lib.orc_ghost_faces(mx, my, mz, g, dp(X), dp(Y), dp(Z))
        self.big = (X, Y, Z)
        return self.big

    def ghost_edges(self):
        X, Y, Z = self.big
        load().orc_ghost_edges(self.mx, self.my, self.mz, self.ngg, dp(X), dp(Y), dp(Z))

    def metrics(self):
        """Jacobian/metrics from the padded coordinates; trims the ghosts."""
        lib = load()
        mx, my, mz, g = self.mx, self.my, self.mz, self.ngg
        X, Y, Z = self.big
        nneg = C.c_int()
        r = lib.orc_jac_metric(mx, my, mz, g, dp(X), dp(Y), dp(Z), dp(self.arr["jac"]), dp(self.arr["metric"]),
                               C.byref(nneg))
        if r != 0:
            raise ValueError("non-positive interior Jacobian")
        for name, src in (("x", X), ("y", Y), ("z", Z)):
            lib.orc_trim_ghost(mx, my, mz, g, dp(src), dp(self.arr[name]))
        return nneg.value

    def geometry(self, x, y, z):
        """Ghost grid + metrics by the oracle's own restatement."""
        self.ghost_grid(x, y, z)
        return self.metrics()

    @property
    def ptr(self):
        return C.byref(self.s)


def oracle_bc(bc):
    return copy_struct(OrcBC, bc)


class OracleSimulation:
    """The oracle's steady loop over the same case (src/Driver.cu:43-104, 218-287)."""

    def __init__(self, case):
        self.lib = load()
        self.case = case
        self.param = copy_struct(OrcParam, case.param)
        self.blocks = []
        for blk in case.blocks:
            ob = OracleBlock(blk.mx, blk.my, blk.mz, case.ngg, self.param)
            ob.ghost_grid(blk.x, blk.y, blk.z, edges=False)
            self.blocks.append(ob)
        # ghost layers across inner faces from the neighbours (src/Mesh.cpp:940-985)
        lib = self.lib
        for b, blk in enumerate(case.blocks):
            ob = self.blocks[b]
            for f in blk.inner:
                t = self.blocks[f.target_block]
                X, Y, Z = ob.big
                TX, TY, TZ = t.big
                lib.orc_inner_ghost_grid(C.byref(copy_struct(OrcInnerFace, f)), ob.mx, ob.my, ob.mz, case.ngg,
                                         dp(X), dp(Y), dp(Z), t.mx, t.my, t.mz, dp(TX), dp(TY), dp(TZ))
        # edges and corners after the inner-face layers (src/Mesh.cpp:859-937)
        for ob in self.blocks:
            ob.ghost_edges()
        for blk, ob in zip(case.blocks, self.blocks):
            ob.metrics()
            ob.arr["bv"][:] = blk.bv0
            if case.param.n_scalar:
                ob.arr["sv"][:case.param.n_scalar] = blk.sv0
        self.bcs = (OrcBC * len(case.blocks))(*[oracle_bc(b.bc) for b in case.blocks])
        self.zones = (OrcZone * len(self.blocks))(*[b.s for b in self.blocks])
        n_inner = [len(blk.inner) for blk in case.blocks]
        self.n_inner = (C.c_int * len(n_inner))(*n_inner)
        flat = [copy_struct(OrcInnerFace, f) for blk in case.blocks for f in blk.inner]
        self.inner = (OrcInnerFace * max(1, len(flat)))(*flat)
        self.res_scale = (C.c_double * 4)(1, 1, 1, 1)
        self.step_no = 0
        self.n_grid = sum(b.mx * b.my * b.mz for b in case.blocks)

    def initialize_computation(self):
        lib, p = self.lib, C.byref(self.param)
        if self.param.turb_method == 1:
            wp = self.case.wall_points
            for b in range(len(self.blocks)):
                if wp is not None and len(wp):
                    a = np.ascontiguousarray(wp, dtype=np.float64)
                    lib.orc_wall_distance(C.byref(self.zones[b]), dp(a), a.shape[0])
                else:
                    self.blocks[b].arr["wall_distance"][:] = 1e3
        for b in range(len(self.blocks)):
            if self.blocks[b].mz == 1:  # 2-D blocks (src/Driver.cu:59-65)
                lib.orc_eliminate_k_gradient(C.byref(self.zones[b]), p)
        for b in range(len(self.blocks)):
            lib.orc_apply_bc(C.byref(self.zones[b]), p, C.byref(self.bcs[b]))
        for b in range(len(self.blocks)):
            lib.orc_compute_cv_from_bv(C.byref(self.zones[b]), p)
            lib.orc_initialize_mut(C.byref(self.zones[b]), p)
        off = 0
        for b in range(len(self.blocks)):
            for f in range(self.n_inner[b]):
                lib.orc_inner_communication(self.zones, b, C.byref(self.inner[off + f]), p)
            off += self.n_inner[b]
        for b in range(len(self.blocks)):
            lib.orc_update_physical_properties(C.byref(self.zones[b]), p)

    def step(self, output_screen=1 << 30):
        self.step_no += 1
        nf = C.c_int()
        r = self.lib.orc_steady_step(self.zones, len(self.blocks), C.byref(self.param), self.bcs, self.inner,
                                     self.n_inner, self.step_no, output_screen, self.res_scale, self.n_grid,
                                     C.byref(nf))
        return None if r < 0 else r

"""ctypes binding of libostc.so (the C-ABI declared in include/ostc.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C openstc_amd/csrc``).  Loading fails loudly when it is missing: there
is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

MAX_SPEC = 20
MAX_REAC = 50
MAX_FACE = 32
MAX_BCDATA = 8

OK = 0
ERR_ARG = 1
AIR, MIXTURE, FR = 0, 1, 2
LAMINAR, RANS = 0, 1
FP32, FP64 = 4, 8

# field identifiers (enum ostc_field)
FIELDS = [
    "X", "Y", "Z", "JAC", "METRIC", "WALL_DIST",
    "BV", "SV", "CV", "BV_LAST",
    "VEL", "SOUND", "MACH", "MUL", "LAMBDA",
    "RHO_D", "GAMMA", "CP", "CHEM_JAC",
    "MUT", "LAMBDA_T", "TURB_JAC",
    "DQ", "DQ0", "DQK", "SPECTR", "DT",
]
F = {name: i for i, name in enumerate(FIELDS)}


class Param(C.Structure):
    """struct ostc_param (mirror of the reference's DParameter)."""
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


class Face(C.Structure):
    _fields_ = [("range_start", C.c_int * 3), ("range_end", C.c_int * 3), ("face", C.c_int),
                ("direction", C.c_int), ("kind", C.c_int), ("slot", C.c_int)]


class Inflow(C.Structure):
    _fields_ = [("density", C.c_double), ("u", C.c_double), ("v", C.c_double), ("w", C.c_double),
                ("pressure", C.c_double), ("temperature", C.c_double), ("mut", C.c_double),
                ("sv", C.c_double * (MAX_SPEC + 2))]


class Wall(C.Structure):
    _fields_ = [("isothermal", C.c_int), ("temperature", C.c_double)]


class BC(C.Structure):
    _fields_ = [("n_face", C.c_int), ("face", Face * MAX_FACE), ("inflow", Inflow * MAX_BCDATA),
                ("wall", Wall * MAX_BCDATA)]


class InnerFace(C.Structure):
    _fields_ = [("range_start", C.c_int * 3), ("range_end", C.c_int * 3), ("face", C.c_int),
                ("direction", C.c_int), ("target_start", C.c_int * 3), ("target_end", C.c_int * 3),
                ("target_block", C.c_int), ("target_face", C.c_int), ("target_direction", C.c_int),
                ("src_tar", C.c_int * 3), ("loop_dir", C.c_int * 3), ("target_loop_dir", C.c_int * 3),
                ("n_point", C.c_int * 3)]


MAX_ELEM = 8


class BlockCoords(C.Structure):
    """struct ostc_block_coords."""
    _fields_ = [("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int),
                ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
                ("x_out", C.c_void_p), ("y_out", C.c_void_p), ("z_out", C.c_void_p),
                ("jac", C.c_void_p), ("metric", C.c_void_p), ("n_neg_jac", C.c_int)]


class MechanismInfo(C.Structure):
    _fields_ = [("n_spec", C.c_int), ("n_elem", C.c_int), ("spec_names", (C.c_char * 24) * MAX_SPEC),
                ("elem_names", (C.c_char * 4) * MAX_ELEM), ("elem_comp", C.c_int * (MAX_SPEC * MAX_ELEM))]


SUR_IN, SUR_MAX_OUT, SUR_MAX_ELEM = 12, 16, 4


class SurDesc(C.Structure):
    """struct ostc_surrogate_desc."""
    _fields_ = [("n_out", C.c_int), ("kind", C.c_int), ("n_elem", C.c_int), ("activation", C.c_int),
                ("in_mean", C.c_float * SUR_IN), ("in_std", C.c_float * SUR_IN),
                ("out_mean", C.c_float * SUR_MAX_OUT), ("out_std", C.c_float * SUR_MAX_OUT),
                ("loss_w", C.c_float * SUR_MAX_OUT), ("elem_w", C.c_float * (SUR_MAX_ELEM * 9)),
                ("floor", C.c_float), ("t_min", C.c_float), ("t_max", C.c_float), ("p_min", C.c_float),
                ("p_max", C.c_float), ("dt_min", C.c_float), ("dt_max", C.c_float),
                ("y_tol", C.c_float), ("elem_tol", C.c_float), ("elem_abs", C.c_float), ("dy_max", C.c_float)]


MAX_BLOCK = 16


class CaseInfo(C.Structure):
    """struct ostc_case_info."""
    _fields_ = [("n_block", C.c_int), ("ngg", C.c_int), ("total_step", C.c_int), ("output_screen", C.c_int),
                ("output_file", C.c_int), ("n_wall_point", C.c_int), ("convergence_criteria", C.c_double),
                ("grid_scale", C.c_double), ("mx", C.c_int * MAX_BLOCK), ("my", C.c_int * MAX_BLOCK),
                ("mz", C.c_int * MAX_BLOCK), ("n_inner", C.c_int * MAX_BLOCK), ("n_neg_jac", C.c_int * MAX_BLOCK),
                ("start_step", C.c_int), ("reserved", C.c_int), ("res_scale", C.c_double * 4),
                ("rho_inf", C.c_double), ("v_inf", C.c_double)]


class ZoneDesc(C.Structure):
    _fields_ = [("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int), ("ngg", C.c_int),
                ("precision", C.c_int), ("block_id", C.c_int)]


_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libostc.so")

# every symbol include/ostc.h declares, with (restype, argtypes)
_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
SIGNATURES = {
    "ostc_ctx_create": (C.c_int, [C.POINTER(Param), C.POINTER(_P)]),
    "ostc_ctx_destroy": (C.c_int, [_P]),
    "ostc_zone_create": (C.c_int, [_P, C.POINTER(ZoneDesc), C.POINTER(_P)]),
    "ostc_zone_destroy": (C.c_int, [_P]),
    "ostc_zone_field_planes": (C.c_int, [_P, C.c_int, _I]),
    "ostc_zone_plane": (C.c_long, [_P]),
    "ostc_zone_upload": (C.c_int, [_P, C.c_int, _D, C.c_int]),
    "ostc_zone_download": (C.c_int, [_P, C.c_int, _D, C.c_int]),
    "ostc_zone_field_ptr": (C.c_int, [_P, C.c_int, C.POINTER(_P), _I]),
    "ostc_zone_set_bc": (C.c_int, [_P, C.POINTER(BC)]),
    "ostc_sizeof_param": (C.c_int, []),
    "ostc_sizeof_bc": (C.c_int, []),
    "ostc_block_geometry": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _D, _D, _D, _D, _D, _D, _D, _D, _I]),
    "ostc_compute_cv_from_bv": (C.c_int, [_P, _P, _P]),
    "ostc_initialize_mut": (C.c_int, [_P, _P, _P]),
    "ostc_wall_distance": (C.c_int, [_P, _P, _D, C.c_int, _P]),
    "ostc_wall_friction_heatflux": (C.c_int, [_P, C.c_double, _P, _P, _P]),
    "ostc_compute_source": (C.c_int, [_P, _P, _P]),
    "ostc_inviscid_flux": (C.c_int, [_P, _P, _P]),
    "ostc_viscous_flux": (C.c_int, [_P, _P, _P]),
    "ostc_local_time_step": (C.c_int, [_P, _P, _P]),
    "ostc_implicit_treatment": (C.c_int, [_P, _P, _P]),
    "ostc_update_cv_bv": (C.c_int, [_P, _P, _P]),
    "ostc_limit_flow": (C.c_int, [_P, _P, _P]),
    "ostc_apply_bc": (C.c_int, [_P, _P, _P]),
    "ostc_inner_halo": (C.c_int, [_P, C.POINTER(_P), C.c_int, C.POINTER(InnerFace), _P]),
    "ostc_update_physical_properties": (C.c_int, [_P, _P, _P]),
    "ostc_store_last_step": (C.c_int, [_P, _P, _P]),
    "ostc_residual_sums": (C.c_int, [_P, C.POINTER(_P), C.c_int, _D, _P]),
    "ostc_residual_sums_noalloc": (C.c_int, [_P, C.POINTER(_P), C.c_int, _P, _P]),
    "ostc_eliminate_k_gradient": (C.c_int, [_P, _P, _P]),
    "ostc_zone_counters": (C.c_int, [_P, C.POINTER(C.c_longlong), C.c_int, _P]),
    "ostc_steady_step": (C.c_int, [_P, C.POINTER(_P), C.c_int, C.POINTER(InnerFace), _I, C.c_int, C.c_int, _D,
                                   _I, _P]),
    "ostc_chem_plain_step": (C.c_int, [_P, C.c_long, _P, _P, _P, _P, _P, _P]),
    "ostc_chem_source_batch": (C.c_int, [_P, C.c_long, _P, _P, _P, _P, _P, _P]),
    "ostc_profile": (C.c_int, [_P, C.c_int]),
    "ostc_profile_read": (C.c_int, [_P, _D, C.POINTER(C.c_longlong), C.c_int, C.c_int]),
    "ostc_profile_stage_name": (C.c_char_p, [C.c_int]),
    "ostc_profile_stage_count": (C.c_int, []),
    "ostc_chem_zone_step": (C.c_int, [_P, _P, _P, C.c_int, _P, C.POINTER(C.c_longlong), _P]),
    "ostc_surrogate_n_params": (C.c_int, [C.c_int]),
    "ostc_surrogate_create": (C.c_int, [_P, C.POINTER(SurDesc), C.POINTER(C.c_float), C.c_long, C.POINTER(_P)]),
    "ostc_surrogate_destroy": (C.c_int, [_P]),
    "ostc_surrogate_get_weights": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "ostc_surrogate_set_weights": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "ostc_surrogate_step": (C.c_int, [_P, C.c_long, _P, _P, _P, _P, _P, C.POINTER(C.c_longlong), _P]),
    "ostc_surrogate_forward": (C.c_int, [_P, C.c_long, _P, _P, _P, _P, _P, _P, _P]),
    "ostc_surrogate_gradients": (C.c_int, [_P, C.c_long, _P, _P, _P, _P, _P, _P, _P, _D, _P]),
    "ostc_surrogate_train_step": (C.c_int, [_P, C.c_long, _P, _P, _P, _P, _P, _P, C.c_float, _D, _P]),
    "ostc_device_alloc": (C.c_int, [C.POINTER(_P), C.c_long]),
    "ostc_device_free": (C.c_int, [_P]),
    "ostc_memcpy_h2d": (C.c_int, [_P, _P, C.c_long]),
    "ostc_memcpy_d2h": (C.c_int, [_P, _P, C.c_long]),
    "ostc_stream_copy": (C.c_int, [_P, _P, C.c_long, _P]),
    "ostc_register_inner_face": (C.c_int, [C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(InnerFace)]),
    "ostc_mesh_geometry": (C.c_int, [C.c_int, C.c_int, C.POINTER(BlockCoords), C.POINTER(InnerFace), _I]),
    "ostc_read_chemkin": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.POINTER(Param),
                                    C.POINTER(MechanismInfo)]),
    "ostc_chemkin_error": (C.c_char_p, []),
    "ostc_case_read": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "ostc_case_destroy": (C.c_int, [_P]),
    "ostc_case_get_info": (C.c_int, [_P, C.POINTER(CaseInfo)]),
    "ostc_case_get_param": (C.c_int, [_P, C.POINTER(Param)]),
    "ostc_case_get_mechanism": (C.c_int, [_P, C.POINTER(MechanismInfo)]),
    "ostc_case_get_bc": (C.c_int, [_P, C.c_int, C.POINTER(BC)]),
    "ostc_case_get_inner": (C.c_int, [_P, C.c_int, C.POINTER(InnerFace), C.c_int, _I]),
    "ostc_case_field": (C.c_int, [_P, C.c_int, C.c_int, _D, C.c_int, C.c_int]),
    "ostc_case_wall_points": (C.c_int, [_P, _D, C.c_int]),
    "ostc_case_error": (C.c_char_p, []),
    "ostc_driver_create": (C.c_int, [_P, C.c_int, C.POINTER(_P)]),
    "ostc_driver_initialize": (C.c_int, [_P]),
    "ostc_driver_step": (C.c_int, [_P, _D, _I, _D]),
    "ostc_driver_simulate": (C.c_int, [_P, C.c_int, _I, _D]),
    "ostc_driver_write_flowfield": (C.c_int, [_P, C.c_char_p]),
    "ostc_driver_post_process": (C.c_int, [_P]),
    "ostc_driver_get_zone": (C.c_int, [_P, C.c_int, C.POINTER(_P)]),
    "ostc_driver_get_ctx": (C.c_int, [_P, C.POINTER(_P)]),
    "ostc_driver_get_step": (C.c_int, [_P, _I]),
    "ostc_driver_destroy": (C.c_int, [_P]),
    "ostc_driver_error": (C.c_char_p, []),
    "ostc_device_sync": (C.c_int, []),
    "ostc_last_error": (C.c_char_p, []),
    "ostc_build_info": (C.c_char_p, []),
}

_lib = None


def load(path: str | None = None):
    """Load libostc.so and bind every exported entry point."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libostc.so is not built ({p}); run __graft_entry__.build() first -- "
                           "the HIP path has no fallback")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class OstcError(RuntimeError):
    pass


def check(status: int, what: str = ""):
    if status != OK:
        msg = load().ostc_last_error().decode(errors="replace")
        raise OstcError(f"{what} failed with status {status}: {msg}")

"""Reference-format case directories for the synthetic workloads, and the C++
case reader / time-loop driver over them.

``write_case(case, directory)`` lays out ``directory/input_files`` the way the
reference reads its working directory: the nine setup files
(src/Parameter.h:64-74, grammar of src/Parameter.cpp:59-168), the grid
(src/Mesh.cpp:597-640: block count, per-block extents, then x, y, z of every
point with i fastest), the boundary records (src/Mesh.cpp:676-696, 0-based),
the inner-face records (src/Mesh.cpp:698-735, 1-based) and the CHEMKIN files
(src/ChemData.cpp:14,75,81).  Coordinates and inflow values are written with
17 significant digits, so they read back bit-exactly.

``CaseDir`` reads such a directory with libostc's C++ reader
(``ostc_case_read``) and ``Driver`` runs it with the C++ driver
(``ostc_driver_*``); fields that setup files cannot express (seeded
perturbations, tanh profiles) are injected with ``CaseDir.set_field``.
"""
from __future__ import annotations

import ctypes as C
import os
import shutil

import numpy as np

from . import _lib as L
from .cases import DATA


def _num(v) -> str:
    if isinstance(v, bool):
        return "1" if v else "0"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return repr(float(v))


def setup_files(setup: dict) -> dict:
    """File name -> text of the nine setup files."""
    s = setup
    files = {}
    files["0_global_control.txt"] = "\n".join([
        "bool parallel = 0",
        "int initial = 0",
        f"int total_step = {s['total_step']}",
        f"int output_file = {s['output_file']}",
        f"int output_screen = {s['output_screen']}", ""])
    files["1_grid_information.txt"] = "\n".join([
        "int dimension = 3", "real gridScale = 1.0", "real referenceLength = 1.0", "int wall_distance = 0", ""])
    files["2_scheme.txt"] = "\n".join([
        "bool    steady = 1",
        f"real    cfl = {_num(s['cfl'])}",
        f"int     implicit_method = {s['implicit_method']}",
        f"int     DPLUR_inner_step = {s['DPLUR_inner_step']}",
        f"real convergence_criteria = {_num(s['convergence_criteria'])}",
        "int\t temporal_scheme = 0",
        "int  inviscid_scheme = 3",
        f"int  reconstruction = {s['reconstruction']}",
        "int  limiter = 0",
        f"int viscous_order = {s['viscous_order']}", ""])
    files["3_species_reactions.txt"] = "\n".join([
        f"bool species = {_num(bool(s['species']))}",
        f"bool reaction = {_num(bool(s['reaction']))}",
        "string mechanism_file = chemistry/combustion_mechanism.inp",
        "string therm_file = chemistry/therm.dat",
        "string transport_file = chemistry/tran.dat",
        f"int chemSrcMethod = {s['chemSrcMethod']}", ""])
    files["4_laminar_turbulent.txt"] = "\n".join([
        f"bool    turbulence = {_num(bool(s['turbulence']))}",
        "int     turbulence_method = 1",
        "int     RANS_model = 2",
        "int     turb_implicit   =   1", ""])
    lines = ["array   string  boundary_conditions {", "    " + " ".join(n for n, _ in s["bcs"]), "}", ""]
    for name, d in s["bcs"]:
        lines.append(f"struct  {name} {{")
        lines.append(f"    string  type = {d['type']}")
        lines.append(f"    int     label = {int(d['label'])}")
        for k, v in d.items():
            if k in ("type", "label"):
                continue
            if isinstance(v, str):
                lines.append(f"    string  {k} = {v}")
            else:
                lines.append(f"    real    {k} = {_num(float(v))}")
        lines.append("}")
        lines.append("")
    files["5_boundary_condition.txt"] = "\n".join(lines)
    files["6_post_process.txt"] = "array   int post_process {\n}\n"
    files["8_initialization.txt"] = f"int groups_init = 1\nstring  default_init    =   {s['default_init']}\n"
    files["9_transport_property.txt"] = "\n".join([
        "int diffusivity_method = 1", "real schmidt_number = 0.5", "real prandtl_number = 0.72",
        "real turbulent_prandtl_number = 0.9", "real turbulent_schmidt_number = 0.9", ""])
    return files


def _rank(stem, ext):
    return f"{stem}{0:>4}.{ext}"


def write_grid(path: str, blocks) -> None:
    with open(path, "w") as f:
        f.write(f"{len(blocks)}\n")
        for b in blocks:
            f.write(f"{b.mx} {b.my} {b.mz}\n")
        for b in blocks:
            for arr in (b.x, b.y, b.z):
                vals = np.ascontiguousarray(arr, dtype=np.float64).ravel()
                f.write("\n".join(map(repr, vals.tolist())))
                f.write("\n")


def write_boundary(path: str, blocks) -> None:
    with open(path, "w") as f:
        for b in blocks:
            f.write(f"{len(b.rows)}\n")
            for r in b.rows:
                f.write(" ".join(f"{int(v):6d}" for v in r) + "\n")


def write_inner(path: str, blocks) -> None:
    with open(path, "w") as f:
        f.write(f"{len(blocks)}\n")
        f.write(" ".join(str(len(b.inner_raw)) for b in blocks) + "\n")
        f.write(" imin  imax  jmin  jmax  kmin  kmax   id\n")
        for bi, b in enumerate(blocks):
            for r in b.inner_raw:
                f.write(" ".join(f"{int(v):6d}" for v in list(r[:6]) + [bi + 1]) + "\n")
                f.write(" ".join(f"{int(v):6d}" for v in r[6:13]) + "\n")


def write_case(case, directory: str) -> str:
    """Write ``case`` (with ``case.meta['setup']`` and per-block ``rows`` /
    ``inner_raw``) as a reference-format case directory; returns it."""
    setup = case.meta["setup"]
    inp = os.path.join(directory, "input_files")
    for sub in ("setup", "grid", "boundary_condition", "chemistry"):
        os.makedirs(os.path.join(inp, sub), exist_ok=True)
    for name, text in setup_files(setup).items():
        with open(os.path.join(inp, "setup", name), "w") as f:
            f.write(text)
    write_grid(os.path.join(inp, "grid", _rank("grid", "grd")), case.blocks)
    write_boundary(os.path.join(inp, "boundary_condition", _rank("boundary", "txt")), case.blocks)
    write_inner(os.path.join(inp, "boundary_condition", _rank("inner", "txt")), case.blocks)
    for fn in ("combustion_mechanism.inp", "therm.dat", "tran.dat"):
        shutil.copyfile(os.path.join(DATA, fn), os.path.join(inp, "chemistry", fn))
    return directory


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


class CaseDir:
    """A case directory read by libostc's C++ reader (host only)."""

    def __init__(self, directory: str):
        self.lib = L.load()
        h = C.c_void_p()
        st = self.lib.ostc_case_read(directory.encode(), C.byref(h))
        if st != 0:
            raise L.OstcError(f"ostc_case_read({directory}): {self.lib.ostc_case_error().decode()}")
        self.h = h
        self.info = L.CaseInfo()
        L.check(self.lib.ostc_case_get_info(h, C.byref(self.info)), "ostc_case_get_info")
        self.ngg = self.info.ngg

    def shape(self, b: int):
        g = self.ngg
        return (self.info.mz[b] + 2 * g, self.info.my[b] + 2 * g, self.info.mx[b] + 2 * g)

    def param(self) -> L.Param:
        p = L.Param()
        L.check(self.lib.ostc_case_get_param(self.h, C.byref(p)), "ostc_case_get_param")
        return p

    def bc(self, b: int) -> L.BC:
        out = L.BC()
        L.check(self.lib.ostc_case_get_bc(self.h, b, C.byref(out)), "ostc_case_get_bc")
        return out

    def inner(self, b: int):
        n = C.c_int()
        L.check(self.lib.ostc_case_get_inner(self.h, b, None, 0, C.byref(n)), "ostc_case_get_inner")
        arr = (L.InnerFace * max(1, n.value))()
        L.check(self.lib.ostc_case_get_inner(self.h, b, arr, n.value, C.byref(n)), "ostc_case_get_inner")
        return [arr[q] for q in range(n.value)]

    def field(self, b: int, field: str, n_planes: int) -> np.ndarray:
        out = np.empty((n_planes,) + self.shape(b))
        L.check(self.lib.ostc_case_field(self.h, b, L.F[field], _dptr(out), n_planes, 0), f"case field {field}")
        return out

    def set_field(self, b: int, field: str, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64)
        n = a.size // int(np.prod(self.shape(b)))
        L.check(self.lib.ostc_case_field(self.h, b, L.F[field], _dptr(a), n, 1), f"case set_field {field}")

    def wall_points(self) -> np.ndarray:
        n = self.info.n_wall_point
        out = np.zeros((max(n, 1), 3))
        L.check(self.lib.ostc_case_wall_points(self.h, _dptr(out), n), "ostc_case_wall_points")
        return out[:n]

    def close(self):
        if self.h:
            self.lib.ostc_case_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Driver:
    """libostc's C++ time-loop driver over a case directory."""

    def __init__(self, case_dir: CaseDir, precision: int = L.FP64):
        self.lib = case_dir.lib
        h = C.c_void_p()
        st = self.lib.ostc_driver_create(case_dir.h, precision, C.byref(h))
        if st != 0:
            raise L.OstcError(f"ostc_driver_create: {self.lib.ostc_driver_error().decode()}")
        self.h = h
        self.n_block = case_dir.info.n_block
        self.shapes = [case_dir.shape(b) for b in range(self.n_block)]

    def _check(self, st, what):
        if st != 0:
            raise L.OstcError(f"{what}: {self.lib.ostc_driver_error().decode()}")

    def initialize(self):
        self._check(self.lib.ostc_driver_initialize(self.h), "ostc_driver_initialize")

    def step(self):
        """One iteration; (err_max, res[4]) on output steps, else None."""
        res = (C.c_double * 4)()
        computed = C.c_int()
        err = C.c_double()
        self._check(self.lib.ostc_driver_step(self.h, res, C.byref(computed), C.byref(err)), "ostc_driver_step")
        return (err.value, list(res)) if computed.value else None

    def simulate(self, io: bool = False):
        done = C.c_int()
        err = C.c_double()
        self._check(self.lib.ostc_driver_simulate(self.h, 1 if io else 0, C.byref(done), C.byref(err)),
                    "ostc_driver_simulate")
        return done.value, err.value

    def write_flowfield(self, path: str | None = None):
        """The reference's flow-field file (Tecplot binary); default
        <case>/output/field/flowfield.plt."""
        self._check(self.lib.ostc_driver_write_flowfield(self.h, path.encode() if path else None),
                    "ostc_driver_write_flowfield")

    def post_process(self):
        """The setup's post_process procedures (0: wall friction and heat
        flux -> <case>/output/wall/friction_heatflux.dat)."""
        self._check(self.lib.ostc_driver_post_process(self.h), "ostc_driver_post_process")

    def zone(self, b: int):
        z = C.c_void_p()
        self._check(self.lib.ostc_driver_get_zone(self.h, b, C.byref(z)), "ostc_driver_get_zone")
        return z

    def download(self, b: int, field: str, n_planes: int) -> np.ndarray:
        out = np.empty((n_planes,) + self.shapes[b])
        L.check(self.lib.ostc_zone_download(self.zone(b), L.F[field], _dptr(out), n_planes), f"download {field}")
        return out

    def close(self):
        if self.h:
            self.lib.ostc_driver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""Case directories in the reference's format, read by libostc's C++ reader
(Parameter src/Parameter.cpp:8-168, Mesh src/Mesh.cpp:546-735, inflow states
src/BoundCond.cpp:6-106, boundary programs src/BoundCond.cu:73-232,
initialize_from_start src/Initialize.cu:4-43), and the C++ time-loop driver
(Driver src/Driver.cu:14-380).

CPU: the synthetic configurations written with case_files.write_case read
back bit-exactly equal to the Python-built cases (model constants, boundary
programs, geometry, inner faces, wall points, initial field); the
reference's own sample setup files (tests/golden/reference_setup, copied
from its input_files/setup) parse to the values they state, with inflow
states equal to an independent Python evaluation; grammar edge cases.
GPU: the C++ driver reproduces the Python-driven loop bit-for-bit, and the
command-line runner completes a case."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from openstc_amd import _lib as L
from openstc_amd import case_files, cases

HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLE = os.path.join(HERE, "golden", "reference_setup")


def struct_diffs(a, b, name=""):
    out = []
    for f, _ in a._fields_:
        va, vb = getattr(a, f), getattr(b, f)
        if hasattr(va, "_fields_"):
            out += struct_diffs(va, vb, f"{name}{f}.")
        elif hasattr(va, "_length_"):
            la, lb = list(va), list(vb)
            if la and hasattr(la[0], "_fields_"):
                for q, (x, y) in enumerate(zip(la, lb)):
                    out += struct_diffs(x, y, f"{name}{f}[{q}].")
            else:
                la = [bytes(x) if hasattr(x, "_length_") else x for x in la]
                lb = [bytes(x) if hasattr(x, "_length_") else x for x in lb]
                if la != lb:
                    out.append(name + f)
        elif va != vb:
            out.append(name + f)
    return out


BUILDERS = {
    "warped_box": lambda: cases.warped_box(mx=20, my=16, mz=8, steps=3),
    "sst_plate": lambda: cases.sst_plate(mx=24, my=20, mz=8, steps=3),
    "mixing_layer": lambda: cases.mixing_layer(mx=16, my=12, mz=8, steps=3),
    "warped_box_split": lambda: cases.warped_box_split(mx=20, my=16, mz=8, steps=3),
    "jet_crossflow": lambda: cases.jet_crossflow(mx=16, my=12, mz=10, nblocks=4, steps=3),
}


@pytest.mark.parametrize("name", sorted(BUILDERS))
def test_written_case_reads_back_bit_exact(tmp_path, name):
    case = BUILDERS[name]()
    d = case_files.write_case(case, str(tmp_path / name))
    cd = case_files.CaseDir(d)
    assert cd.info.n_block == len(case.blocks)
    assert cd.info.total_step == case.meta["setup"]["total_step"]
    assert struct_diffs(case.param, cd.param()) == []
    for b, blk in enumerate(case.blocks):
        assert struct_diffs(blk.bc, cd.bc(b)) == [], b
        for fld, arr, n in (("X", blk.xg, 1), ("Y", blk.yg, 1), ("Z", blk.zg, 1), ("JAC", blk.jac, 1),
                            ("METRIC", blk.metric, 9)):
            assert np.array_equal(cd.field(b, fld, n).reshape(arr.shape), arr), (b, fld)
        inner = cd.inner(b)
        assert len(inner) == len(blk.inner)
        for x, y in zip(blk.inner, inner):
            assert struct_diffs(x, y) == []
        # initialize_from_start: the default inflow ("freestream", inflow slot
        # 0 in every configuration) everywhere, ghosts included
        bv = cd.field(b, "BV", 6)
        st = blk.bc.inflow[0]
        ref = [st.density, st.u, st.v, st.w, st.pressure, st.temperature]
        for l in range(6):
            assert np.all(bv[l] == ref[l]), l
        nsc = case.param.n_scalar
        if nsc:
            sv = cd.field(b, "SV", nsc)
            for l in range(nsc):
                assert np.all(sv[l] == st.sv[l]), l
    if case.wall_points is not None:
        assert np.array_equal(cd.wall_points(), case.wall_points)


def sample_dir(tmp_path, mx=20, my=16, mz=8):
    """The reference's sample setup (9-species FR + SST, DPLUR 3, NND-2) on a
    3-D box with the sample's boundary labels: wall 2 at j=0, freestream 5 and
    fuel 9 on i=0, outflow 6 elsewhere."""
    d = tmp_path / "sample"
    inp = d / "input_files"
    (inp / "setup").mkdir(parents=True)
    for fn in os.listdir(SAMPLE):
        if fn.endswith(".txt") and fn[0].isdigit():
            shutil.copyfile(os.path.join(SAMPLE, fn), inp / "setup" / fn)
    grid = (inp / "setup" / "1_grid_information.txt")
    grid.write_text(grid.read_text().replace("int dimension = 2", "int dimension = 3"))
    (inp / "grid").mkdir()
    (inp / "boundary_condition").mkdir()
    (inp / "chemistry").mkdir()
    for fn in ("combustion_mechanism.inp", "therm.dat", "tran.dat"):
        shutil.copyfile(os.path.join(cases.DATA, fn), inp / "chemistry" / fn)
    k, j, i = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    blk = cases.Block(mx, my, mz, i * 1e-3 * 1.0, j * 1e-3 * 0.5, k * 1e-3 * 0.7)
    js = my // 2
    blk.rows = [(0, mx - 1, 0, 0, 0, mz - 1, 2), (0, 0, js, my - 1, 0, mz - 1, 5), (0, 0, 0, js, 0, mz - 1, 9),
                (mx - 1, mx - 1, 0, my - 1, 0, mz - 1, 6), (0, mx - 1, my - 1, my - 1, 0, mz - 1, 6),
                (0, mx - 1, 0, my - 1, 0, 0, 6), (0, mx - 1, 0, my - 1, mz - 1, mz - 1, 6)]
    case_files.write_grid(str(inp / "grid" / "grid   0.grd"), [blk])
    case_files.write_boundary(str(inp / "boundary_condition" / "boundary   0.txt"), [blk])
    return str(d), blk


def test_reference_sample_setup(tmp_path):
    d, blk = sample_dir(tmp_path)
    cd = case_files.CaseDir(d)
    p = cd.param()
    info = cd.info
    # 0_global_control / 2_scheme / 3_species_reactions / 4_laminar_turbulent / 9_transport_property
    assert (info.total_step, info.output_file, info.output_screen) == (10000, 5000, 1000)
    assert info.convergence_criteria == 1e-5 and info.grid_scale == 0.001
    assert (p.mix_model, p.turb_method, p.rans_model, p.turb_implicit) == (L.FR, L.RANS, 2, 1)
    assert (p.inviscid_scheme, p.reconstruction, p.limiter, p.viscous_scheme) == (3, 3, 0, 2)
    assert (p.implicit_method, p.dplur_inner_step, p.chem_src_method) == (1, 3, 1)
    assert (p.n_spec, p.n_reac, p.n_scalar, p.n_var) == (9, 19, 11, 16)
    assert (p.cfl, p.Pr, p.Sc, p.Prt, p.Sct) == (0.5, 0.72, 0.5, 0.9, 0.9)
    # gridScale 0.001 scales the coordinates
    assert np.array_equal(cd.field(0, "X", 1)[0, 2:-2, 2:-2, 2:-2], blk.x * 0.001)
    # 5_boundary_condition: freestream (label 5), fuel (label 9), isothermal wall 294.4 K
    mech, minfo = cases.load_mechanism(True)
    names = cases.species_names(minfo)
    y_fs = np.zeros(9)
    for n, v in (("O2", 0.258), ("N2", 0.486), ("H2O", 0.256)):
        y_fs[names.index(n)] = v
    y_fu = np.zeros(9)
    y_fu[names.index("H2")] = 1.0
    fs = cases.mixture_inflow(mech, y_fs, 2.44, 105431.0, 1270.0, rans=True, intensity=0.01, mut_ratio=1.0)
    fu = cases.mixture_inflow(mech, y_fu, 1.0, 105431.0, 254.0, rans=True, intensity=0.01, mut_ratio=1.0)
    bc = cd.bc(0)
    for slot, st in ((0, fs), (1, fu)):
        iv = bc.inflow[slot]
        got = [iv.density, iv.u, iv.v, iv.w, iv.pressure, iv.temperature, iv.mut] + list(iv.sv[:11])
        ref = [st.density, st.u, st.v, st.w, st.pressure, st.temperature, st.mut] + list(st.sv)
        assert got == ref, slot
    assert bc.wall[0].isothermal == 1 and bc.wall[0].temperature == 294.4
    # application order: outflow faces, inflow 5 then 9, then the wall (trimmed)
    kinds = [(bc.face[q].kind, bc.face[q].slot) for q in range(bc.n_face)]
    assert kinds == [(6, 0)] * 4 + [(5, 0), (5, 1), (2, 0)]
    w = bc.face[6]
    assert list(w.range_start) == [0, 0, 0] and list(w.range_end) == [blk.mx - 1, 0, blk.mz - 1]
    # DParameter limits from the freestream struct
    assert p.limit_ll[0] == 1e-6 * fs.density and p.limit_ul[4] == 1e+3 * fs.pressure
    assert list(p.sv_inf[:11]) == list(fs.sv)
    assert cd.info.n_wall_point == blk.mx * blk.mz
    # 8_initialization: groups_init = 1 -> the freestream everywhere
    bv = cd.field(0, "BV", 6)
    assert np.all(bv[0] == fs.density) and np.all(bv[5] == fs.temperature)


def write_setup(tmp_path, edit):
    case = cases.warped_box(mx=12, my=10, mz=6, steps=3)
    d = case_files.write_case(case, str(tmp_path / "c"))
    for fn, fun in edit.items():
        path = os.path.join(d, "input_files", "setup", fn)
        with open(path) as f:
            text = f.read()
        with open(path, "w") as f:
            f.write(fun(text))
    return d, case


def test_setup_grammar_rules(tmp_path):
    # first definition wins; a BOM-prefixed comment line is ignored; "//"
    # continues an array on the next line; the value on a string array's
    # opening line is dropped (the reference reads two tokens there)
    d, case = write_setup(tmp_path, {
        "2_scheme.txt": lambda t: "﻿// header\nreal    cfl = 0.25\n" + t,
        "6_post_process.txt": lambda t: "array int post_process { 0 //\n 3 }\n",
        "5_boundary_condition.txt": lambda t: t.replace(
            "array   string  boundary_conditions {\n    freestream\n}",
            "array string boundary_conditions { bogus\n freestream\n}"),
    })
    p = case_files.CaseDir(d).param()
    assert p.cfl == 0.25
    # init_cond patch: a box taking the state of another struct
    d, case = write_setup(tmp_path / "b", {
        "8_initialization.txt": lambda t: t.replace("int groups_init = 1", "int groups_init = 2") +
        "struct init_cond_0 {\n string label = slow\n real x0 = -1\n real x1 = 0.5\n real y0 = -1\n"
        " real y1 = 10\n real z0 = -1\n real z1 = 10\n}\n",
        "5_boundary_condition.txt": lambda t: t + "struct slow {\n string type = inflow\n int label = 7\n"
        " real mach = 0.5\n real pressure = 90000\n real temperature = 250\n}\n",
    })
    cd = case_files.CaseDir(d)
    bv = cd.field(0, "BV", 6)
    x = cd.field(0, "X", 1)[0]
    slow = cases.air_inflow(0.5, 90000.0, 250.0)
    fast = cases.air_inflow(2.0, 101325.0, 300.0)
    inside = x <= 0.5
    assert inside.any() and (~inside).any()
    assert np.all(bv[0][inside] == slow.density) and np.all(bv[1][inside] == slow.u)
    assert np.all(bv[0][~inside] == fast.density)


@pytest.mark.parametrize("edit,msg", [
    ({"1_grid_information.txt": lambda t: t.replace("int dimension = 3", "int dimension = 2")}, "dimension"),
    ({"0_global_control.txt": lambda t: t.replace("bool parallel = 0", "bool parallel = 1")}, "parallel"),
    ({"0_global_control.txt": lambda t: t.replace("int initial = 0", "int initial = 2")}, "initial"),
    ({"0_global_control.txt": lambda t: t.replace("int initial = 0", "int initial = 1")}, "restart"),
    ({"2_scheme.txt": lambda t: t.replace("real    cfl", "real    cfl_typo")}, "cfl"),
    ({"5_boundary_condition.txt": lambda t: t.replace("struct  freestream", "struct  other")}, "freestream"),
])
def test_unsupported_or_incomplete_setup_is_rejected(tmp_path, edit, msg):
    d, _ = write_setup(tmp_path, edit)
    with pytest.raises(L.OstcError, match=msg):
        case_files.CaseDir(d)


# ---- the reference's flow-field file (Tecplot binary V112, MPIIO layout),
# written and read here independently of the library (src/MPIIO.hpp:60-548)
def plt_write(path, names, zones):
    """zones: list of (mx, my, mz, [var arrays, i fastest])"""
    def s(txt):
        return np.array([ord(ch) for ch in txt] + [0], "<i4").tobytes()
    out = [b"#!TDV112", np.array([1, 0], "<i4").tobytes(), s("Solution file"), np.int32(len(names)).tobytes()]
    out += [s(n) for n in names]
    for b, (mx, my, mz, _) in enumerate(zones):
        out += [np.float32(299.0).tobytes(), s(f"zone {b}"), np.array([-1, -2], "<i4").tobytes(),
                np.float64(0).tobytes(), np.array([-1, 0, 0, 0, 0, mx, my, mz, 0], "<i4").tobytes()]
    out.append(np.float32(357.0).tobytes())
    for mx, my, mz, vs in zones:
        out += [np.float32(299.0).tobytes(), np.full(len(names), 2, "<i4").tobytes(),
                np.array([0, 0, -1], "<i4").tobytes()]
        out += [np.array([v.min(), v.max()], "<f8").tobytes() for v in vs]
        out += [np.ascontiguousarray(v, "<f8").tobytes() for v in vs]
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "wb") as f:
        f.write(b"".join(out))


def plt_read(path):
    buf = open(path, "rb").read()
    at = [0]

    def take(fmt, n=1):
        a = np.frombuffer(buf, fmt, n, at[0])
        at[0] += a.nbytes
        return a

    def s():
        chars = []
        while True:
            c = int(take("<i4")[0])
            if c == 0:
                return "".join(chars)
            chars.append(chr(c))
    assert buf[:8] == b"#!TDV112"
    at[0] = 8
    assert list(take("<i4", 2)) == [1, 0]
    assert s() == "Solution file"
    nv = int(take("<i4")[0])
    names = [s() for _ in range(nv)]
    dims = []
    while True:
        m = float(take("<f4")[0])
        if m == 357.0:
            break
        assert m == 299.0
        name = s()
        head = take("<i4", 2)
        take("<f8")
        rest = take("<i4", 9)
        assert list(head) == [-1, -2] and list(rest[:5]) == [-1, 0, 0, 0, 0] and rest[8] == 0
        dims.append((name, tuple(int(v) for v in rest[5:8])))
    zones = []
    for name, (mx, my, mz) in dims:
        assert float(take("<f4")[0]) == 299.0
        assert list(take("<i4", nv)) == [2] * nv and list(take("<i4", 3)) == [0, 0, -1]
        mm = take("<f8", 2 * nv).reshape(nv, 2)
        vs = [take("<f8", mx * my * mz).reshape(mz, my, mx) for _ in range(nv)]
        for v, (lo, hi) in zip(vs, mm):
            assert lo == v.min() and hi == v.max()
        zones.append((name, vs))
    assert at[0] == len(buf)
    return names, zones


def test_restart_reads_the_flowfield_file(tmp_path):
    """initial = 1 (src/Initialize.cuh:58-200): interior bv/sv from
    output/field/flowfield.plt by variable label (any case, aliases rho/p/T/k;
    unknown labels skipped), ghosts and missing variables from the setup's
    initialisation; a mismatching zone is an error."""
    case = cases.sst_plate(mx=12, my=10, mz=6)
    d = case_files.write_case(case, str(tmp_path / "c"))
    base = case_files.CaseDir(d)
    bv0, sv0 = base.field(0, "BV", 6), base.field(0, "SV", 2)
    path = os.path.join(d, "input_files", "setup", "0_global_control.txt")
    text = open(path).read()
    open(path, "w").write(text.replace("int initial = 0", "int initial = 1"))
    rng = np.random.default_rng(11)
    mx, my, mz = 12, 10, 6
    vals = {n: rng.uniform(0.5, 2.0, (mz, my, mx)) for n in ("RHO", "u", "V", "w", "p", "Temperature", "k", "OMEGA")}
    names = ["x", "y", "z"] + list(vals) + ["mach", "mut", "bogus"]
    arrays = [rng.normal(size=(mz, my, mx)) for _ in range(3)] + list(vals.values()) + \
        [rng.normal(size=(mz, my, mx)) for _ in range(3)]
    plt_write(os.path.join(d, "output", "field", "flowfield.plt"), names, [(mx, my, mz, arrays)])
    # no step.txt / residual_scale.txt: step 0, scales 1
    cd = case_files.CaseDir(d)
    assert cd.info.start_step == 0 and list(cd.info.res_scale) == [1.0] * 4
    msg = os.path.join(d, "output", "message")
    os.makedirs(msg, exist_ok=True)
    open(os.path.join(msg, "step.txt"), "w").write("120")
    open(os.path.join(msg, "residual_scale.txt"), "w").write("0.5\n2e-05\n1234.5\n0.25\n")
    cd = case_files.CaseDir(d)
    assert cd.info.start_step == 120 and list(cd.info.res_scale) == [0.5, 2e-05, 1234.5, 0.25]
    bv = cd.field(0, "BV", 6)
    sv = cd.field(0, "SV", 2)
    g = cd.ngg
    inner = (slice(g, -g),) * 3
    for l, n in enumerate(["RHO", "u", "V", "w", "p", "Temperature"]):
        assert np.array_equal(bv[l][inner], vals[n]), n
    assert np.array_equal(sv[0][inner], vals["k"]) and np.array_equal(sv[1][inner], vals["OMEGA"])
    # ghosts keep the setup's initialisation
    ghost = np.ones(bv0[0].shape, bool)
    ghost[inner] = False
    assert np.array_equal(bv[:, ghost], bv0[:, ghost]) and np.array_equal(sv[:, ghost], sv0[:, ghost])
    # a file without the turbulence variables: k, omega keep the inflow values
    plt_write(os.path.join(d, "output", "field", "flowfield.plt"), names[:9], [(mx, my, mz, arrays[:9])])
    cd2 = case_files.CaseDir(d)
    assert np.array_equal(cd2.field(0, "SV", 2), sv0)
    assert np.array_equal(cd2.field(0, "BV", 6)[0][inner], vals["RHO"])
    # wrong zone size
    plt_write(os.path.join(d, "output", "field", "flowfield.plt"), names,
              [(mx, my, mz - 1, [a[:-1] for a in arrays])])
    with pytest.raises(L.OstcError, match="does not match"):
        case_files.CaseDir(d)


def inject(cd, case):
    for b, blk in enumerate(case.blocks):
        cd.set_field(b, "BV", blk.bv0)
        if case.param.n_scalar:
            cd.set_field(b, "SV", blk.sv0)


@pytest.mark.gpu
@pytest.mark.parametrize("name,precision,steps", [("warped_box", L.FP64, 5), ("sst_plate", L.FP32, 3),
                                                   ("mixing_layer", L.FP32, 2), ("warped_box_split", L.FP64, 3),
                                                   ("jet_crossflow", L.FP32, 2)])
def test_driver_matches_python_loop(tmp_path, name, precision, steps):
    from openstc_amd.runtime import Simulation
    print(f"[gpu-driver] {name} precision {precision} steps {steps}", flush=True)
    case = BUILDERS[name]()
    d = case_files.write_case(case, str(tmp_path / name))
    cd = case_files.CaseDir(d)
    inject(cd, case)
    drv = case_files.Driver(cd, precision)
    drv.initialize()
    sim = Simulation(case, precision)
    sim.initialize_computation()
    for s in range(steps):
        r = drv.step()
        e = sim.step(output_screen=case.meta["setup"]["output_screen"])
        if r is not None or e is not None:
            assert r[0] == e, (s, r, e)
    sim.sync()
    for b in range(len(case.blocks)):
        got = drv.download(b, "BV", 6)
        ref = sim.zones[b].download("BV")
        assert np.array_equal(got, ref), b
    drv.close()
    sim.close()


@pytest.mark.gpu
def test_command_line_runner(tmp_path):
    case = cases.warped_box(mx=24, my=16, mz=8, steps=4)
    d = case_files.write_case(case, str(tmp_path / "cli"))
    exe = os.path.join(os.path.dirname(L.LIB_PATH), "ostc_run")
    print("[gpu-driver] ostc_run", d, flush=True)
    r = subprocess.run([exe, d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Mcell-iter/s" in r.stdout and "converged to" in r.stdout
    hist = open(os.path.join(d, "history.dat")).read().split("\n")
    assert hist[0] == "step\terror_max" and hist[1].startswith("1\t")
    assert os.path.exists(os.path.join(d, "output", "message", "residual_scale.txt"))
    # the flow-field file of the last output step (output_file divides total_step)
    names, zones = plt_read(os.path.join(d, "output", "field", "flowfield.plt"))
    assert names == ["x", "y", "z", "density", "u", "v", "w", "pressure", "temperature", "mach"]
    assert len(zones) == 1 and zones[0][1][0].shape == (8, 16, 24)


@pytest.mark.gpu
@pytest.mark.parametrize("name,precision", [("sst_plate", L.FP64), ("mixing_layer", L.FP32)])
def test_flowfield_file_and_restart_round_trip(tmp_path, name, precision):
    """The driver's flow-field file (MPIIO::print_field) read back by the
    independent reader above equals the zone's interior fields and grid; a
    restart (initial = 1) from it reproduces the zone's bv/sv exactly."""
    case = getattr(cases, name)(mx=12, my=10, mz=6)
    d = case_files.write_case(case, str(tmp_path / name))
    cd = case_files.CaseDir(d)
    inject(cd, case)
    print("[gpu-driver] flowfield", name, flush=True)
    drv = case_files.Driver(cd, precision)
    drv.initialize()
    for _ in range(3):
        drv.step()
    drv.write_flowfield()
    names, zones = plt_read(os.path.join(d, "output", "field", "flowfield.plt"))
    p = cd.param()
    ns, rans = p.n_spec, p.turb_method == L.RANS
    spec = cases.species_names(cases.load_mechanism(True)[1]) if ns else []
    assert names == ["x", "y", "z", "density", "u", "v", "w", "pressure", "temperature", "mach"] + spec + \
        (["tke", "omega", "mut"] if rans else [])
    g = cd.ngg
    inner = (slice(g, -g),) * 3
    vs = dict(zip(names, zones[0][1]))
    for l, n in enumerate(["density", "u", "v", "w", "pressure", "temperature"]):
        assert np.array_equal(vs[n], drv.download(0, "BV", 6)[l][inner]), n
    assert np.array_equal(vs["mach"], drv.download(0, "MACH", 1)[0][inner])
    sv = drv.download(0, "SV", p.n_scalar) if p.n_scalar else None
    for k, n in enumerate(spec + (["tke", "omega"] if rans else [])):
        assert np.array_equal(vs[n], sv[k][inner]), n
    for n, f in (("x", "X"), ("y", "Y"), ("z", "Z")):
        assert np.array_equal(vs[n], cd.field(0, f, 1)[0][inner])
    bv = drv.download(0, "BV", 6)
    drv.close()
    # restart
    path = os.path.join(d, "input_files", "setup", "0_global_control.txt")
    text = open(path).read()
    open(path, "w").write(text.replace("int initial = 0", "int initial = 1"))
    cd2 = case_files.CaseDir(d)
    assert cd2.info.start_step == 3  # output/message/step.txt of the written file
    assert np.array_equal(cd2.field(0, "BV", 6)[:, g:-g, g:-g, g:-g], bv[:, g:-g, g:-g, g:-g])
    drv2 = case_files.Driver(cd2, precision)
    step = C.c_int()
    L.check(drv2.lib.ostc_driver_get_step(drv2.h, C.byref(step)), "get_step")
    assert step.value == 3
    drv2.close()
    if p.n_scalar:
        assert np.array_equal(cd2.field(0, "SV", p.n_scalar)[:, g:-g, g:-g, g:-g], sv[:, g:-g, g:-g, g:-g])


def test_post_process_array_is_read(tmp_path):
    """A setup with post_process { 0 } reads, and the case exposes the
    freestream density and speed of the procedure's dynamic pressure (the
    procedure itself runs on the GPU)."""
    case = cases.jet_crossflow(mx=16, my=12, mz=10, nblocks=2, steps=2)
    d = case_files.write_case(case, str(tmp_path / "pp"))
    path = os.path.join(d, "input_files", "setup", "6_post_process.txt")
    open(path, "w").write(open(os.path.join(SAMPLE, "6_post_process.txt")).read().replace("// 0", "0"))
    cd = case_files.CaseDir(d)
    assert cd.info.n_block == 2
    # the freestream state behind the dynamic pressure is the one the flow
    # limits are formed from (src/DParameter.cu:82-87)
    p = cd.param()
    assert cd.info.rho_inf > 0 and cd.info.v_inf > 0
    assert p.limit_ll[0] == 1e-6 * cd.info.rho_inf and p.limit_ul[1] == 1e+3 * cd.info.v_inf


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [L.FP64, L.FP32])
def test_wall_friction_heatflux_post_process(tmp_path, precision):
    """Procedure 0 (src/PostProcess.cu:6-70) on an isothermal-wall case
    (jet_crossflow, 2 blocks): the device entry point equals the numpy oracle on
    the downloaded fields of every block, and simulate(io) leaves
    output/wall/friction_heatflux.dat holding the last block's line."""
    import orc_post
    case = cases.jet_crossflow(mx=16, my=12, mz=10, nblocks=2, steps=2)
    d = case_files.write_case(case, str(tmp_path / "pp"))
    setup = os.path.join(d, "input_files", "setup")
    open(os.path.join(setup, "6_post_process.txt"), "w").write(
        open(os.path.join(SAMPLE, "6_post_process.txt")).read().replace("// 0", "0"))
    ctl = os.path.join(setup, "0_global_control.txt")
    text = open(ctl).read()
    out_line = [ln for ln in text.split("\n") if "output_file" in ln][0]
    open(ctl, "w").write(text.replace(out_line, "int output_file = 1"))
    cd = case_files.CaseDir(d)
    inject(cd, case)
    print(f"[gpu-driver] post_process jet_crossflow 2 blocks precision {precision}", flush=True)
    drv = case_files.Driver(cd, precision)
    drv.initialize()
    drv.simulate(io=True)
    g = cd.ngg
    lib = drv.lib
    dev = C.c_void_p()
    mx = cd.info.mx[0]
    L.check(lib.ostc_device_alloc(C.byref(dev), 2 * mx * 8), "alloc")
    try:
        # dynamic pressure of the freestream struct (src/PostProcess.cu:25-27)
        q = 0.5 * cd.info.rho_inf * cd.info.v_inf ** 2
        assert q > 0
        for b in range(2):
            metric = drv.download(b, "METRIC", 9)
            bv = drv.download(b, "BV", 6)
            mul = drv.download(b, "MUL", 1)[0]
            lam = drv.download(b, "LAMBDA", 1)[0]
            cf_ref, qw_ref = orc_post.wall_friction_heatflux(metric, bv, mul, lam, g, q)
            L.check(lib.ostc_wall_friction_heatflux(drv.zone(b), q, dev, C.c_void_p(dev.value + mx * 8), None),
                    "ostc_wall_friction_heatflux")
            got = np.empty(2 * mx)
            L.check(lib.ostc_memcpy_d2h(got.ctypes.data_as(C.c_void_p), dev, 2 * mx * 8), "d2h")
            cf, qw = got[:mx], got[mx:]
            assert np.abs(qw_ref).max() > 0 and np.abs(cf_ref).max() > 0
            assert np.allclose(cf, cf_ref, rtol=1e-12, atol=1e-14 * np.abs(cf_ref).max()), b
            assert np.allclose(qw, qw_ref, rtol=1e-12, atol=1e-14 * np.abs(qw_ref).max()), b
        x_line = cd.field(1, "X", 1)[0][g, g, g:g + mx]
        want = orc_post.friction_heatflux_file(x_line, cf_ref, qw_ref).split("\n")
        have = open(os.path.join(d, "output", "wall", "friction_heatflux.dat")).read().split("\n")
        assert have[0] == want[0] and len(have) == len(want)
        a = np.array([[float(v) for v in ln.split("\t")] for ln in have[1:-1]])
        e = np.array([[float(v) for v in ln.split("\t")] for ln in want[1:-1]])
        # 6 significant digits: the last one may round the other way
        assert np.all(np.abs(a - e) <= 1e-5 * np.abs(e) + 1e-300)
    finally:
        lib.ostc_device_free(dev)
        drv.close()

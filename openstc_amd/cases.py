"""Synthetic configurations of the benchmark workloads.

Each builder returns a ``Case``: grid, model constants, boundary program and
initial field, laid out exactly as the reference's setup would produce them
(inflow states: src/BoundCond.cpp:6-158; boundary ranges:
src/Mesh.cpp:14-32 and src/BoundCond.cu:73-250; limit_flow bounds:
src/DParameter.cu:78-100).  Perturbations use a counter hash (splitmix64) of
(linear index, variable) so that any consumer of the same case sees the same
bits.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

RU = 8314.462618
MW_AIR = 28.9647
GAMMA_AIR = 1.4


def splitmix64(x: np.ndarray) -> np.ndarray:
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    return z ^ (z >> np.uint64(31))


def hash_uniform(seed: int, n: int, var: int) -> np.ndarray:
    """u in [-1, 1) for cells 0..n-1 and variable `var`."""
    with np.errstate(over="ignore"):
        idx = np.arange(n, dtype=np.uint64) * np.uint64(8) + np.uint64(var)
        h = splitmix64(idx ^ np.uint64(seed))
    return (h >> np.uint64(11)).astype(np.float64) / float(1 << 53) * 2.0 - 1.0


def sutherland(t):
    return 1.716e-5 * (t / 273) ** 1.5 * (273 + 111) / (t + 111)


@dataclass
class InflowState:
    density: float
    u: float
    v: float
    w: float
    pressure: float
    temperature: float
    velocity: float
    mut: float = 0.0
    sv: list = field(default_factory=list)


def air_inflow(mach, pressure, temperature, direction=(1.0, 0.0, 0.0), rans=False, intensity=0.01, mut_ratio=1.0):
    """Inflow::Inflow for single-species air given (Ma, p, T)
    (src/BoundCond.cpp:6-158)."""
    mw = MW_AIR
    viscosity = sutherland(temperature)
    c = math.sqrt(GAMMA_AIR * RU / mw * temperature)
    velocity = mach * c
    u, v, w = (d * velocity for d in direction)
    density = pressure * mw / (RU * temperature)
    st = InflowState(density, u, v, w, pressure, temperature, velocity)
    if rans:
        st.mut = mut_ratio * viscosity
        k = 1.5 * velocity * velocity * intensity * intensity
        st.sv = [k, density * k / st.mut]
    return st


@dataclass
class Block:
    mx: int
    my: int
    mz: int
    x: np.ndarray  # interior (mz, my, mx)
    y: np.ndarray
    z: np.ndarray
    bc: L.BC = None
    inner: list = field(default_factory=list)
    xg: np.ndarray = None  # padded geometry / metrics (from libostc)
    yg: np.ndarray = None
    zg: np.ndarray = None
    jac: np.ndarray = None
    metric: np.ndarray = None
    bv0: np.ndarray = None
    sv0: np.ndarray = None
    rows: list = field(default_factory=list)       # boundary-file records (i1 i2 j1 j2 k1 k2 label, 0-based)
    inner_raw: list = field(default_factory=list)  # inner-file records (1-based ranges + target block)


@dataclass
class Case:
    name: str
    param: L.Param
    blocks: list
    ngg: int = 2
    precision: int = L.FP64
    steps: int = 100
    output_screen: int = 100
    wall_points: np.ndarray = None
    seed: int = 0
    meta: dict = field(default_factory=dict)


def setup_spec(steps, output_screen, cfl, implicit, dplur_steps, reconstruction, viscous, rans=False,
               species=False, reaction=False, chem_method=0, bcs=(), default_init="freestream",
               convergence=1e-30):
    """The case in the terms of the reference's setup files (written by
    case_files.write_case).  bcs: (struct name, {type, label, values...}) in
    the order of the boundary_conditions array.  The convergence criterion
    defaults to 1e-30 so that runs take their full step count."""
    return dict(total_step=steps, output_file=steps, output_screen=output_screen, cfl=cfl,
                implicit_method=1 if implicit else 0, DPLUR_inner_step=dplur_steps, reconstruction=reconstruction,
                viscous_order=2 if viscous else 0, turbulence=rans, species=species, reaction=reaction,
                chemSrcMethod=chem_method, bcs=list(bcs), default_init=default_init,
                convergence_criteria=convergence)


def register_boundary(row, ngg=2):
    """Boundary::register_boundary (src/Mesh.cpp:14-32), 3-D."""
    i1, i2, j1, j2, k1, k2, label = row
    rs, re = [i1, j1, k1], [i2, j2, k2]
    face, direction = 0, 1
    for d in range(3):
        if rs[d] == re[d]:
            face = d
            if rs[d] == 0:
                direction = -1
            break
    for d in range(3):
        if d == face:
            continue
        if re[d] < rs[d]:
            rs[d], re[d] = re[d], rs[d]
        rs[d] -= ngg
        re[d] += ngg
    return dict(rs=rs, re=re, face=face, direction=direction, label=label)


def bc_program(rows, bc_structs, inflows, walls, ngg=2):
    """Face list in the reference's application order.

    rows: boundary-file rows of one block.  bc_structs: list of (label, type)
    in the order of the setup's boundary_conditions array.  Ordering follows
    DBoundCond::initialize_bc_on_GPU / apply_boundary_conditions
    (src/BoundCond.cu:73-155, src/BoundCond.cuh:321-412)."""
    faces = [register_boundary(r, ngg) for r in rows]
    labels = []
    for f in faces:
        if f["label"] not in labels:
            labels.append(f["label"])
    kind_of = {}
    order = {"wall": [], "inflow": [], "outflow": [], "symmetry": []}
    for lab, typ in bc_structs:
        if lab in labels:
            labels.remove(lab)
            order[typ].append(lab)
            kind_of[lab] = typ
    for lab in labels:
        typ = {2: "wall", 3: "symmetry", 5: "inflow", 6: "outflow"}.get(lab)
        if typ:
            order[typ].append(lab)
            kind_of[lab] = typ
    bc = L.BC()
    n = 0
    code = {"wall": 2, "symmetry": 3, "inflow": 5, "outflow": 6}
    for typ in ("outflow", "inflow", "symmetry", "wall"):
        for slot, lab in enumerate(order[typ]):
            for f in faces:
                if f["label"] != lab:
                    continue
                rs, re = list(f["rs"]), list(f["re"])
                if typ == "wall":  # src/BoundCond.cu:215-230
                    for q in range(3):
                        if q != f["face"]:
                            rs[q] += ngg
                            re[q] -= ngg
                fc = bc.face[n]
                for q in range(3):
                    fc.range_start[q] = rs[q]
                    fc.range_end[q] = re[q]
                fc.face, fc.direction, fc.kind, fc.slot = f["face"], f["direction"], code[typ], slot
                n += 1
    bc.n_face = n
    for s, st in enumerate(inflows):
        iv = bc.inflow[s]
        iv.density, iv.u, iv.v, iv.w = st.density, st.u, st.v, st.w
        iv.pressure, iv.temperature, iv.mut = st.pressure, st.temperature, st.mut
        for l, val in enumerate(st.sv):
            iv.sv[l] = val
    for s, (iso, tw) in enumerate(walls):
        bc.wall[s].isothermal = 1 if iso else 0
        bc.wall[s].temperature = tw
    return bc


def air_param(rans: bool, viscous: bool, implicit: bool, dplur_steps: int, cfl: float, reconstruction=2,
              freestream: InflowState = None) -> L.Param:
    p = L.Param()
    p.mix_model = L.AIR
    p.turb_method = L.RANS if rans else L.LAMINAR
    p.inviscid_scheme = 3
    p.reconstruction = reconstruction
    p.limiter = 0
    p.viscous_scheme = 2 if viscous else 0
    p.rans_model = 2 if rans else 0
    p.turb_implicit = 1
    p.chem_src_method = 0
    p.implicit_method = 1 if implicit else 0
    p.dplur_inner_step = dplur_steps
    p.n_spec = 0
    p.n_scalar = 2 if rans else 0
    p.n_reac = 0
    p.n_var = 5 + p.n_scalar
    p.Pr, p.cfl, p.Sc, p.Prt, p.Sct = 0.72, cfl, 0.5, 0.9, 0.9
    fs = freestream
    # src/DParameter.cu:78-100
    p.limit_ll[0] = 1e-6 * fs.density
    p.limit_ul[0] = 1e+3 * fs.density
    for l in range(1, 4):
        p.limit_ll[l] = -1e+3 * fs.velocity
        p.limit_ul[l] = 1e+3 * fs.velocity
    p.limit_ll[4] = 1e-6 * fs.pressure
    p.limit_ul[4] = 1e+3 * fs.pressure
    if rans:
        p.limit_ul[5] = np.finfo(np.float64).max
        p.limit_ul[6] = np.finfo(np.float64).max
    for l, v in enumerate(fs.sv):
        p.sv_inf[l] = v
    return p


def _pad(a, ngg):
    return np.pad(a, ((ngg, ngg), (ngg, ngg), (ngg, ngg)))


def finish_block(blk: Block, ngg: int = 2):
    """Attach geometry computed by libostc on the host."""
    from .runtime import block_geometry
    blk.xg, blk.yg, blk.zg, blk.jac, metric, nneg = block_geometry(blk.x, blk.y, blk.z, ngg)
    blk.metric = metric
    return blk


def uniform_field(blk: Block, st: InflowState, nsc: int, ngg: int):
    shape = (blk.mz + 2 * ngg, blk.my + 2 * ngg, blk.mx + 2 * ngg)
    bv = np.empty((6,) + shape)
    for l, v in enumerate((st.density, st.u, st.v, st.w, st.pressure, st.temperature)):
        bv[l] = v
    sv = np.empty((nsc,) + shape)
    for l in range(nsc):
        sv[l] = st.sv[l]
    return bv, sv


def perturb_interior(arr, blk: Block, seed: int, var: int, ngg: int, rel: float = None, absval: float = None):
    """arr[interior] *= 1+rel*u  (or += absval*u), u from the counter hash."""
    mz, my, mx = blk.mz, blk.my, blk.mx
    u = hash_uniform(seed, mx * my * mz, var).reshape(mz, my, mx)
    sl = (slice(ngg, ngg + mz), slice(ngg, ngg + my), slice(ngg, ngg + mx))
    if rel is not None:
        arr[sl] *= 1.0 + rel * u
    else:
        arr[sl] += absval * u


def warped_box(mx=64, my=64, mz=8, steps=100, precision=L.FP64, seed=1) -> Case:
    """Plain method, small: single-species air, laminar inviscid, one warped
    64x64x8 block, Mach 2 freestream (p=101325 Pa, T=300 K) with +-1% seeded
    perturbation of rho,u,v,w,p; MUSCL+minmod+AUSM+, explicit Euler, 100 steps;
    inflow at i=0, outflow at i=max, symmetry elsewhere."""
    ngg = 2
    k, j, i = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    xi, eta, zeta = i / (mx - 1), j / (my - 1), k / (mz - 1)
    tp = 2 * np.pi
    x = xi + 0.05 * np.sin(tp * eta) * np.sin(tp * zeta)
    y = eta + 0.05 * np.sin(tp * zeta) * np.sin(tp * xi)
    z = zeta + 0.05 * np.sin(tp * xi) * np.sin(tp * eta)
    fs = air_inflow(2.0, 101325.0, 300.0)
    param = air_param(rans=False, viscous=False, implicit=False, dplur_steps=0, cfl=0.5, reconstruction=2,
                      freestream=fs)
    blk = Block(mx, my, mz, x, y, z)
    rows = [(0, 0, 0, my - 1, 0, mz - 1, 5), (mx - 1, mx - 1, 0, my - 1, 0, mz - 1, 6),
            (0, mx - 1, 0, 0, 0, mz - 1, 3), (0, mx - 1, my - 1, my - 1, 0, mz - 1, 3),
            (0, mx - 1, 0, my - 1, 0, 0, 3), (0, mx - 1, 0, my - 1, mz - 1, mz - 1, 3)]
    blk.bc = bc_program(rows, [(5, "inflow")], [fs], [], ngg)
    blk.rows = rows
    finish_block(blk, ngg)
    bv, sv = uniform_field(blk, fs, 0, ngg)
    perturb_interior(bv[0], blk, seed, 0, ngg, rel=0.01)
    perturb_interior(bv[1], blk, seed, 1, ngg, rel=0.01)
    perturb_interior(bv[2], blk, seed, 2, ngg, absval=0.01 * fs.u)
    perturb_interior(bv[3], blk, seed, 3, ngg, absval=0.01 * fs.u)
    perturb_interior(bv[4], blk, seed, 4, ngg, rel=0.01)
    # T from the equation of state of the perturbed state
    bv[5] = bv[4] * MW_AIR / (bv[0] * RU)
    blk.bv0, blk.sv0 = bv, sv
    setup = setup_spec(steps, steps, 0.5, False, 0, 2, False,
                       bcs=[("freestream", dict(type="inflow", label=5, mach=2.0, pressure=101325.0,
                                                temperature=300.0, u=1.0, v=0.0, w=0.0))])
    return Case("warped_box_air_euler_explicit", param, [blk], ngg, precision, steps, steps, None, seed,
                dict(mach=2.0, p=101325.0, T=300.0, setup=setup))


def warped_box_split(mx=64, my=64, mz=8, split=None, steps=100, precision=L.FP64, seed=1) -> Case:
    """The warped box cut at i = split into two blocks that
    share the node plane i = split (block 1 = i <= split, block 2 = i >= split),
    connected by an inner face each way (the reference's inner{rank}.txt
    records, registered by libostc); inflow on block 1's i = 0, outflow on block
    2's i = max, symmetry elsewhere; same perturbed Mach 2 field."""
    ngg = 2
    split = mx // 2 if split is None else split
    full = warped_box(mx, my, mz, steps, precision, seed)
    gb = full.blocks[0]
    fs = air_inflow(2.0, 101325.0, 300.0)
    from .runtime import mesh_geometry, register_inner_face
    spans = [(0, split), (split, mx - 1)]
    blocks = []
    for b, (i0, i1) in enumerate(spans):
        n = i1 - i0 + 1
        blk = Block(n, my, mz, gb.x[:, :, i0:i1 + 1].copy(), gb.y[:, :, i0:i1 + 1].copy(), gb.z[:, :, i0:i1 + 1].copy())
        rows = [(0, n - 1, 0, 0, 0, mz - 1, 3), (0, n - 1, my - 1, my - 1, 0, mz - 1, 3),
                (0, n - 1, 0, my - 1, 0, 0, 3), (0, n - 1, 0, my - 1, mz - 1, mz - 1, 3)]
        if b == 0:
            rows.insert(0, (0, 0, 0, my - 1, 0, mz - 1, 5))
        else:
            rows.insert(0, (n - 1, n - 1, 0, my - 1, 0, mz - 1, 6))
        blk.bc = bc_program(rows, [(5, "inflow")], [fs], [], ngg)
        blk.rows = rows
        blocks.append(blk)
    n0, n1 = blocks[0].mx, blocks[1].mx
    raw = [  # i1 i2 j1 j2 k1 k2 | target ranges | target block (1-based)
        [n0, n0, -1, -my, 1, mz, 1, 1, -1, -my, 1, mz, 2],
        [1, 1, -1, -my, 1, mz, n0, n0, -1, -my, 1, mz, 1],
    ]
    for blk, r in zip(blocks, raw):
        blk.inner = [register_inner_face(r, ngg)]
        blk.inner_raw = [r]
    for blk, geo in zip(blocks, mesh_geometry(blocks, ngg)):
        blk.xg, blk.yg, blk.zg, blk.jac, blk.metric, _ = geo
    # the perturbed single-block field, restricted to each block (ghosts too)
    for blk, (i0, i1) in zip(blocks, spans):
        lo, hi = i0, i1 + 2 * ngg + 1
        blk.bv0 = full.blocks[0].bv0[:, :, :, lo:hi].copy()
        blk.sv0 = full.blocks[0].sv0[:, :, :, lo:hi].copy()
    return Case("warped_box_split_2blocks", full.param, blocks, ngg, precision, steps, steps, None, seed,
                dict(split=split, setup=full.meta["setup"]))


def _tanh_stretch(n, height, first):
    """y_j = H*(1 - tanh(beta*(1-s))/tanh(beta)), s=j/(n-1), beta such that y_1 = first."""
    s1 = 1.0 / (n - 1)

    def y1(beta):
        return height * (1 - math.tanh(beta * (1 - s1)) / math.tanh(beta))

    lo, hi = 1e-3, 20.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if y1(mid) > first:
            lo = mid
        else:
            hi = mid
    beta = 0.5 * (lo + hi)
    s = np.arange(n) / (n - 1)
    return height * (1 - np.tanh(beta * (1 - s)) / math.tanh(beta)), beta


def sst_plate(mx=256, my=128, mz=64, steps=1000, precision=L.FP32, dplur_steps=4, cfl=5.0, height=0.1,
            span=0.25, seed=2, perturb=False) -> Case:
    """Viscous RANS: air, k-omega SST, flat plate with tanh wall clustering
    (first cell y+~1 from the turbulent flat-plate Cf correlation at x=0.5 m),
    Mach 2.5, Re=1e7/m, T=220 K, adiabatic no-slip wall at j=0, inflow i=0,
    outflow i=max and j=max, symmetry in k; MUSCL+AUSM+ + central viscous
    flux, DPLUR with 4 sweeps, CFL 5."""
    ngg = 2
    t_inf, mach, re_m = 220.0, 2.5, 1e7
    mu = sutherland(t_inf)
    c = math.sqrt(GAMMA_AIR * RU / MW_AIR * t_inf)
    u_inf = mach * c
    rho = re_m * mu / u_inf
    p_inf = rho * RU * t_inf / MW_AIR
    fs = air_inflow(mach, p_inf, t_inf, rans=True, intensity=0.01, mut_ratio=1.0)
    cf = 0.0592 * (re_m * 0.5) ** -0.2
    u_tau = u_inf * math.sqrt(cf / 2)
    y_first = (mu / rho) / u_tau
    yline, beta = _tanh_stretch(my, height, y_first)
    k, j, i = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    x = i / (mx - 1) * 1.0
    y = yline[j]
    z = k / (mz - 1) * span
    param = air_param(rans=True, viscous=True, implicit=True, dplur_steps=dplur_steps, cfl=cfl, reconstruction=2,
                      freestream=fs)
    blk = Block(mx, my, mz, x, y, z)
    rows = [(0, 0, 0, my - 1, 0, mz - 1, 5), (mx - 1, mx - 1, 0, my - 1, 0, mz - 1, 6),
            (0, mx - 1, 0, 0, 0, mz - 1, 2), (0, mx - 1, my - 1, my - 1, 0, mz - 1, 6),
            (0, mx - 1, 0, my - 1, 0, 0, 3), (0, mx - 1, 0, my - 1, mz - 1, mz - 1, 3)]
    blk.bc = bc_program(rows, [(5, "inflow"), (2, "wall")], [fs], [(False, 300.0)], ngg)
    blk.rows = rows
    finish_block(blk, ngg)
    bv, sv = uniform_field(blk, fs, 2, ngg)
    if perturb:  # +-1% seeded perturbation (parity tests: no cancellation-only residuals)
        for l in (0, 1, 4):
            perturb_interior(bv[l], blk, seed, l, ngg, rel=0.01)
        perturb_interior(bv[2], blk, seed, 2, ngg, absval=0.01 * fs.u)
        perturb_interior(bv[3], blk, seed, 3, ngg, absval=0.01 * fs.u)
        perturb_interior(sv[0], blk, seed, 5, ngg, rel=0.01)
        perturb_interior(sv[1], blk, seed, 6, ngg, rel=0.01)
        bv[5] = bv[4] * MW_AIR / (bv[0] * RU)
    blk.bv0, blk.sv0 = bv, sv
    # wall points: the trimmed wall face (src/Driver.cu:128-149)
    wall = np.stack([x[:, 0, :].ravel(), y[:, 0, :].ravel(), z[:, 0, :].ravel()], axis=1)
    setup = setup_spec(steps, 100, cfl, True, dplur_steps, 2, True, rans=True,
                       bcs=[("freestream", dict(type="inflow", label=5, mach=mach, pressure=p_inf, temperature=t_inf,
                                                u=1.0, v=0.0, w=0.0, turbulence_intensity=0.01,
                                                turb_viscosity_ratio=1.0)),
                            ("wall", dict(type="wall", label=2, thermal_type="adiabatic"))])
    return Case("sst_flat_plate_dplur", param, [blk], ngg, precision, steps, 100, wall, seed,
                dict(mach=mach, p=p_inf, T=t_inf, rho=rho, u=u_inf, y_first=y_first, beta=beta, setup=setup))


# ---------------------------------------------------------------- mixtures

import os as _os

DATA = _os.path.join(_os.path.dirname(_os.path.abspath(__file__)), "data", "chemkin")


def load_mechanism(with_reactions=True, directory: str = DATA):
    """CHEMKIN inputs -> (Param with the chemistry tables filled, MechanismInfo)."""
    lib = L.load()
    p = L.Param()
    info = L.MechanismInfo()
    st = lib.ostc_read_chemkin(_os.path.join(directory, "combustion_mechanism.inp").encode(),
                               _os.path.join(directory, "therm.dat").encode(),
                               _os.path.join(directory, "tran.dat").encode(), 1 if with_reactions else 0,
                               C.byref(p), C.byref(info))
    if st != 0:
        raise L.OstcError("ostc_read_chemkin: " + lib.ostc_chemkin_error().decode())
    return p, info


def species_names(info):
    return [info.spec_names[i].value.decode() for i in range(info.n_spec)]


def nasa_cp(p: L.Param, t: float) -> np.ndarray:
    """cp_k(T) [J/kg/K] (Species::compute_cp, src/ChemData.cpp:100-117)."""
    ns = p.n_spec
    out = np.empty(ns)
    for i in range(ns):
        tt = max(t, p.t_low[i])
        c = p.low_temp_coeff if tt < p.t_mid[i] else p.high_temp_coeff
        c = [c[i * 7 + q] for q in range(5)]
        t2 = tt * tt
        t3 = t2 * tt
        t4 = t3 * tt
        out[i] = c[0] + c[1] * tt + c[2] * t2 + c[3] * t3 + c[4] * t4
        out[i] *= RU / p.mw[i]
    return out


def mixture_viscosity(p: L.Param, t: float, y: np.ndarray) -> float:
    """Wilke mixture viscosity (host compute_viscosity, src/Transport.cu:11-38)."""
    ns = p.n_spec
    mw = np.array(p.mw[:ns])
    mwt = 0.0
    for i in range(ns):
        mwt += y[i] / mw[i]
    mwt = 1.0 / mwt
    x = y * mwt / mw
    vis = np.empty(ns)
    for i in range(ns):
        tdl = t * p.LJ_potent_inv[i]
        ci = 1.147 * tdl ** -0.145 + (tdl + 0.5) ** -2
        vis[i] = p.vis_coeff[i] * math.sqrt(t) / ci
    mu = 0.0
    for i in range(ns):
        s = 0.0
        for j in range(ns):
            if i == j:
                pf = 1.0
            else:
                num = 1 + math.sqrt(vis[i] / vis[j]) * p.WjDivWi_to_One4th[i * L.MAX_SPEC + j]
                pf = num * num * p.sqrt_WiDivWjPl1Mul8[i * L.MAX_SPEC + j]
            s += pf * x[j]
        mu += vis[i] * x[i] / s
    return mu


def mixture_inflow(p: L.Param, y: np.ndarray, mach, pressure, temperature, direction=(1.0, 0.0, 0.0), rans=False,
                   intensity=0.01, mut_ratio=1.0):
    """Inflow::Inflow for a mixture given (Y, Ma, p, T) (src/BoundCond.cpp:6-158)."""
    ns = p.n_spec
    mwk = np.array(p.mw[:ns])
    mw = 0.0
    for i in range(ns):
        mw += y[i] / mwk[i]
    mw = 1.0 / mw
    viscosity = mixture_viscosity(p, temperature, y)
    cpi = nasa_cp(p, temperature)
    cp = cv = 0.0
    for i in range(ns):
        cp += y[i] * cpi[i]
        cv += y[i] * (cpi[i] - RU / mwk[i])
    gamma = cp / cv
    c = math.sqrt(gamma * RU / mw * temperature)
    velocity = mach * c
    u, v, w = (d * velocity for d in direction)
    density = pressure * mw / (RU * temperature)
    st = InflowState(density, u, v, w, pressure, temperature, velocity)
    st.sv = list(y)
    if rans:
        st.mut = mut_ratio * viscosity
        k = 1.5 * velocity * velocity * intensity * intensity
        st.sv += [k, density * k / st.mut]
    return st


def mixture_param(mech: L.Param, rans: bool, implicit: bool, dplur_steps: int, cfl: float, chem_method: int,
                  freestream: InflowState, reconstruction=2) -> L.Param:
    p = L.Param()
    C.memmove(C.addressof(p), C.addressof(mech), C.sizeof(L.Param))
    p.mix_model = L.FR if mech.n_reac > 0 else L.MIXTURE
    p.turb_method = L.RANS if rans else L.LAMINAR
    p.inviscid_scheme = 3
    p.reconstruction = reconstruction
    p.limiter = 0
    p.viscous_scheme = 2
    p.rans_model = 2 if rans else 0
    p.turb_implicit = 1
    p.chem_src_method = chem_method
    p.implicit_method = 1 if implicit else 0
    p.dplur_inner_step = dplur_steps
    p.n_scalar = p.n_spec + (2 if rans else 0)
    p.n_var = 5 + p.n_scalar
    # setup/9_transport_property.txt of the reference's sample case
    p.Pr, p.cfl, p.Sc, p.Prt, p.Sct = 0.72, cfl, 0.5, 0.9, 0.9
    fs = freestream
    p.limit_ll[0] = 1e-6 * fs.density
    p.limit_ul[0] = 1e+3 * fs.density
    for l in range(1, 4):
        p.limit_ll[l] = -1e+3 * fs.velocity
        p.limit_ul[l] = 1e+3 * fs.velocity
    p.limit_ll[4] = 1e-6 * fs.pressure
    p.limit_ul[4] = 1e+3 * fs.pressure
    if rans:
        p.limit_ul[5] = np.finfo(np.float64).max
        p.limit_ul[6] = np.finfo(np.float64).max
    for l, v in enumerate(fs.sv):
        p.sv_inf[l] = v
    return p


def mixing_layer(mx=192, my=96, mz=96, steps=500, precision=L.FP32, dplur_steps=4, cfl=0.2, chem_method=1,
            lx=0.2, ly=0.1, lz=0.1, seed=3, mech_dir: str = DATA) -> Case:
    """Finite-rate chemistry: 9-species/19-reaction H2-air, Cartesian box with
    a seeded mixing layer: phi = (1+tanh((y-y0)/(2*dy)))/2 between H2 (Y=1,
    1000 K, Ma 1.2, below y0) and air (O2 0.233 / N2 0.767, 1500 K, Ma 2.0,
    above y0), p = 1 atm, +-2% perturbation of rho,u,v,w,p; inflow as two
    labelled patches split at y0 (air label 5, fuel label 9), outflow at
    i=max and j=0/j=max, symmetry in k; laminar, EPI chemistry, DPLUR.
    CFL 0.2: the layer ignites (T up to ~2800 K) and burns stably through
    500 iterations; at CFL 1 the local time step overshoots the heat release
    at ignition within three iterations and limit_flow clamps the whole field
    (the oracle does the same: the reference's algorithm, not a GPU effect).
    mech_dir: another CHEMKIN mechanism with H2, O2 and N2 among its species."""
    ngg = 2
    mech, info = load_mechanism(True, mech_dir)
    names = species_names(info)
    ns = mech.n_spec
    y_air = np.zeros(ns)
    y_air[names.index("O2")] = 0.233
    y_air[names.index("N2")] = 0.767
    y_fuel = np.zeros(ns)
    y_fuel[names.index("H2")] = 1.0
    air = mixture_inflow(mech, y_air, 2.0, 101325.0, 1500.0)
    fuel = mixture_inflow(mech, y_fuel, 1.2, 101325.0, 1000.0)
    param = mixture_param(mech, rans=False, implicit=True, dplur_steps=dplur_steps, cfl=cfl, chem_method=chem_method,
                          freestream=air)
    k, j, i = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    x = i * (lx / (mx - 1))
    y = j * (ly / (my - 1))
    z = k * (lz / (mz - 1))
    blk = Block(mx, my, mz, x, y, z)
    jsplit = my // 2
    rows = [(0, 0, jsplit, my - 1, 0, mz - 1, 5), (0, 0, 0, jsplit, 0, mz - 1, 9),
            (mx - 1, mx - 1, 0, my - 1, 0, mz - 1, 6),
            (0, mx - 1, 0, 0, 0, mz - 1, 6), (0, mx - 1, my - 1, my - 1, 0, mz - 1, 6),
            (0, mx - 1, 0, my - 1, 0, 0, 3), (0, mx - 1, 0, my - 1, mz - 1, mz - 1, 3)]
    blk.bc = bc_program(rows, [(5, "inflow"), (9, "inflow")], [air, fuel], [], ngg)
    blk.rows = rows
    finish_block(blk, ngg)
    # mixing-layer initial field on the padded block (ghosts get the profile too)
    shape = (mz + 2 * ngg, my + 2 * ngg, mx + 2 * ngg)
    yc = blk.yg
    y0 = 0.5 * (y[0, jsplit, 0] + y[0, jsplit - 1, 0])
    dy = ly / (my - 1)
    phi = 0.5 * (1.0 + np.tanh((yc - y0) / (2 * dy)))
    sv = np.empty((ns,) + shape)
    for l in range(ns):
        sv[l] = phi * y_air[l] + (1 - phi) * y_fuel[l]
    bv = np.empty((6,) + shape)
    bv[1] = phi * air.u + (1 - phi) * fuel.u
    bv[2] = 0.0
    bv[3] = 0.0
    bv[4] = 101325.0
    bv[5] = phi * air.temperature + (1 - phi) * fuel.temperature
    perturb_interior(bv[1], blk, seed, 1, ngg, rel=0.02)
    perturb_interior(bv[2], blk, seed, 2, ngg, absval=0.02 * air.u)
    perturb_interior(bv[3], blk, seed, 3, ngg, absval=0.02 * air.u)
    perturb_interior(bv[4], blk, seed, 4, ngg, rel=0.02)
    mwk = np.array(mech.mw[:ns])
    mw = 1.0 / np.einsum("l...,l->...", sv, 1.0 / mwk)
    bv[0] = bv[4] * mw / (RU * bv[5])
    perturb_interior(bv[0], blk, seed, 0, ngg, rel=0.02)
    bv[5] = bv[4] * mw / (RU * bv[0])
    blk.bv0, blk.sv0 = bv, sv
    setup = setup_spec(steps, 100, cfl, True, dplur_steps, 2, True, species=True, reaction=True,
                       chem_method=chem_method,
                       bcs=[("freestream", dict(type="inflow", label=5, mach=2.0, pressure=101325.0, temperature=1500.0,
                                                u=1.0, v=0.0, w=0.0, O2=0.233, N2=0.767)),
                            ("fuel", dict(type="inflow", label=9, mach=1.2, pressure=101325.0, temperature=1000.0,
                                          u=1.0, v=0.0, w=0.0, H2=1.0))])
    return Case("h2air_mixing_layer_fr_epi_dplur", param, [blk], ngg, precision, steps, 100, None, seed,
                dict(species=names, air=air.__dict__, fuel=fuel.__dict__, mech_info=info, setup=setup))


def jet_crossflow(mx=256, my=192, mz=128, nblocks=4, steps=200, precision=L.FP32, dplur_steps=4, cfl=0.2, chem_method=1,
            lx=0.4, ly=0.1, lz=0.05, seed=5, mech_dir: str = DATA) -> Case:
    """Full path at scale: SST RANS + 9-species finite-rate chemistry on
    `nblocks` blocks of mx x my x mz joined along i by inner faces (shared node
    planes, the reference's inner{rank}.txt records).  Vitiated-air crossflow
    (O2 0.258 / N2 0.486 / H2O 0.256 by mass, 1270 K, Ma 2.0, p = 1 atm)
    enters at block 1's i = 0; an H2 jet (254 K, Ma 2.5, normal to the wall)
    enters through a slot in the j = 0 wall of block 1; isothermal wall
    (300 K) elsewhere at j = 0, outflow at j = max and the last block's
    i = max, symmetry in k.  Initial field: the crossflow with a seeded
    analytic H2 plume above the slot (tanh-blended ellipsoid, composition,
    temperature and velocity blended, +-2 % perturbation).  EPI chemistry,
    DPLUR 4 sweeps, CFL 0.2: the jet ignites (T up to ~2600 K) and burns
    stably through 200 iterations; at CFL 0.5 and above the heat release at
    ignition overshoots within 20-30 iterations and limit_flow clamps the
    field (the oracle does the same)."""
    ngg = 2
    mech, info = load_mechanism(True, mech_dir)
    names = species_names(info)
    ns = mech.n_spec
    y_air = np.zeros(ns)
    for n, v in (("O2", 0.258), ("N2", 0.486), ("H2O", 0.256)):
        y_air[names.index(n)] = v
    y_fuel = np.zeros(ns)
    y_fuel[names.index("H2")] = 1.0
    p_inf, t_air, t_jet = 101325.0, 1270.0, 254.0
    air = mixture_inflow(mech, y_air, 2.0, p_inf, t_air, rans=True, intensity=0.01, mut_ratio=1.0)
    jet = mixture_inflow(mech, y_fuel, 2.5, p_inf, t_jet, direction=(0.0, 1.0, 0.0), rans=True, intensity=0.01,
                         mut_ratio=1.0)
    param = mixture_param(mech, rans=True, implicit=True, dplur_steps=dplur_steps, cfl=cfl, chem_method=chem_method,
                          freestream=air)
    nx_tot = nblocks * (mx - 1) + 1
    dx, dy, dz = lx / (nx_tot - 1), ly / (my - 1), lz / (mz - 1)
    # jet slot on block 1's wall: i in [i0, i1], k in [k0, k1]
    i0, i1 = mx // 4, mx // 4 + max(1, mx // 16)
    k0, k1 = mz // 2 - max(1, mz // 16), mz // 2 + max(1, mz // 16)
    xj, zj = (0.5 * (i0 + i1)) * dx, (0.5 * (k0 + k1)) * dz
    blocks = []
    from .runtime import mesh_geometry, register_inner_face
    for b in range(nblocks):
        k, j, i = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
        x = (i + b * (mx - 1)) * dx
        blk = Block(mx, my, mz, x, j * dy, k * dz)
        rows = []
        if b == 0:
            rows.append((0, 0, 0, my - 1, 0, mz - 1, 5))
            rows += [(0, i0 - 1, 0, 0, 0, mz - 1, 2), (i1 + 1, mx - 1, 0, 0, 0, mz - 1, 2),
                     (i0, i1, 0, 0, 0, k0 - 1, 2), (i0, i1, 0, 0, k1 + 1, mz - 1, 2),
                     (i0, i1, 0, 0, k0, k1, 9)]
        else:
            rows.append((0, mx - 1, 0, 0, 0, mz - 1, 2))
        if b == nblocks - 1:
            rows.append((mx - 1, mx - 1, 0, my - 1, 0, mz - 1, 6))
        rows += [(0, mx - 1, my - 1, my - 1, 0, mz - 1, 6),
                 (0, mx - 1, 0, my - 1, 0, 0, 3), (0, mx - 1, 0, my - 1, mz - 1, mz - 1, 3)]
        blk.rows = rows
        blk.bc = bc_program(rows, [(5, "inflow"), (9, "inflow"), (2, "wall")], [air, jet], [(True, 300.0)], ngg)
        # inner faces: i = max of block b <-> i = 1 of block b+1 (1-based records)
        raw = []
        if b > 0:
            raw.append([1, 1, -1, -my, 1, mz, mx, mx, -1, -my, 1, mz, b])
        if b < nblocks - 1:
            raw.append([mx, mx, -1, -my, 1, mz, 1, 1, -1, -my, 1, mz, b + 2])
        blk.inner_raw = raw
        blk.inner = [register_inner_face(r, ngg) for r in raw]
        blocks.append(blk)
    for blk, geo in zip(blocks, mesh_geometry(blocks, ngg)):
        blk.xg, blk.yg, blk.zg, blk.jac, blk.metric, _ = geo
    mwk = np.array(mech.mw[:ns])
    for b, blk in enumerate(blocks):
        # seeded analytic plume: tanh-blended ellipsoid above the slot
        r = np.sqrt(((blk.xg - xj) / (6 * (i1 - i0 + 1) * dx)) ** 2 + (blk.yg / (0.25 * ly)) ** 2 +
                    ((blk.zg - zj) / (2 * (k1 - k0 + 1) * dz)) ** 2)
        phi = 0.5 * (1.0 - np.tanh((r - 1.0) / 0.15))
        shape = blk.xg.shape
        sv = np.empty((ns + 2,) + shape)
        for l in range(ns):
            sv[l] = phi * y_fuel[l] + (1 - phi) * y_air[l]
        sv[ns] = air.sv[ns]
        sv[ns + 1] = air.sv[ns + 1]
        bv = np.empty((6,) + shape)
        bv[1] = (1 - phi) * air.u
        bv[2] = phi * 0.3 * jet.v
        bv[3] = 0.0
        bv[4] = p_inf
        bv[5] = phi * t_jet + (1 - phi) * t_air
        s = seed * 16 + b
        perturb_interior(bv[1], blk, s, 1, ngg, rel=0.02)
        perturb_interior(bv[2], blk, s, 2, ngg, absval=0.02 * air.u)
        perturb_interior(bv[3], blk, s, 3, ngg, absval=0.02 * air.u)
        perturb_interior(bv[4], blk, s, 4, ngg, rel=0.02)
        mw = 1.0 / np.einsum("l...,l->...", sv[:ns], 1.0 / mwk)
        bv[0] = bv[4] * mw / (RU * bv[5])
        perturb_interior(bv[0], blk, s, 0, ngg, rel=0.02)
        bv[5] = bv[4] * mw / (RU * bv[0])
        blk.bv0, blk.sv0 = bv, sv
    # wall points: the wall faces in condition order, blocks in order, faces in file order
    pts = []
    for blk in blocks:
        for r in blk.rows:
            if r[6] != 2:
                continue
            sl = (slice(r[4], r[5] + 1), slice(r[2], r[3] + 1), slice(r[0], r[1] + 1))
            pts.append(np.stack([blk.x[sl].ravel(), blk.y[sl].ravel(), blk.z[sl].ravel()], axis=1))
    wall = np.concatenate(pts, axis=0)
    setup = setup_spec(steps, 100, cfl, True, dplur_steps, 2, True, rans=True, species=True, reaction=True,
                       chem_method=chem_method,
                       bcs=[("freestream", dict(type="inflow", label=5, mach=2.0, pressure=p_inf, temperature=t_air,
                                                u=1.0, v=0.0, w=0.0, O2=0.258, N2=0.486, H2O=0.256,
                                                turbulence_intensity=0.01, turb_viscosity_ratio=1.0)),
                            ("jet", dict(type="inflow", label=9, mach=2.5, pressure=p_inf, temperature=t_jet,
                                         u=0.0, v=1.0, w=0.0, H2=1.0, turbulence_intensity=0.01,
                                         turb_viscosity_ratio=1.0)),
                            ("wall", dict(type="wall", label=2, thermal_type="isothermal", temperature=300.0))])
    return Case(f"h2_jet_vitiated_crossflow_fr_sst_{nblocks}x{mx}x{my}x{mz}", param, blocks, ngg, precision, steps, 100,
                wall, seed, dict(species=names, air=air.__dict__, jet=jet.__dict__, mech_info=info, setup=setup))

"""Closed-form pins of the oracle's flow routines (oracle/orc_flow.c) on the
CPU.  Each check is a property the mathematics fixes, chosen so that a
plausible mistake (a dropped term, a wrong sign or index, a transposed
operand) fails it:

* AUSM+ (src/InviscidScheme.cuh:70-149): equal left/right states give the
  exact Euler flux (the Mach and pressure splittings sum to M and 1);
  supersonic flow is fully upwinded; with linearly varying primitives the
  MUSCL/minmod reconstruction is exact at the face; at an extremum it drops
  to first order.
* Inviscid residual of a uniform flow on a Cartesian block is exactly zero.
* Viscous flux (src/ViscousScheme.cuh): u = a y^2 gives the x-momentum
  residual 2 mu a * cell volume.
* Local time step (src/TimeAdvanceFunc.cuh:21-71): closed form on a
  Cartesian block.
* DPLUR Jacobian product (src/DPLURHPP.cuh:78-125): equals a central finite
  difference of the exact Euler flux in conservative variables.
* SST source (src/SST.cu:60-178): simple shear u = a y with uniform k, omega:
  production mu_t a^2 and the blended destruction terms (Menter's constants).
* Symmetry BC (src/BoundCond.cuh): tangential velocity at the face node,
  mirrored ghosts with the normal component reversed.
* update (src/FieldOperation.cuh:139-183): primitive recovery from cv + dq.
* limit_flow (src/TimeAdvanceFunc.cu:34-224): a bad cell takes the average of
  its valid neighbours, with the reference's i-range bounded by mz."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_ffi as O
from openstc_amd import _lib as L
from openstc_amd import cases

GAMMA = 1.4
RU = 8314.462618
MW = 28.9647  # air, kg/kmol (src/Constants.h:26)


@pytest.fixture(scope="module")
def lib():
    return O.load()


def air(rans=False, viscous=True, cfl=0.5):
    fs = cases.air_inflow(2.0, 101325.0, 300.0, rans=rans)
    p = cases.air_param(rans=rans, viscous=viscous, implicit=True, dplur_steps=2, cfl=cfl, freestream=fs)
    return O.copy_struct(O.OrcParam, p)


def cartesian(param, mx=8, my=7, mz=6, dx=0.1, dy=0.2, dz=0.3):
    ob = O.OracleBlock(mx, my, mz, 2, param)
    k, j, i = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    ob.geometry(i * dx, j * dy, k * dz)
    return ob


def coords(ob):
    """(x, y, z) of every padded point."""
    return ob.arr["x"][0], ob.arr["y"][0], ob.arr["z"][0]


def set_state(ob, rho, u, v, w, p):
    t = p * MW / (rho * RU)
    shape = ob.shape
    for l, val in enumerate((rho, u, v, w, p, t)):
        ob.arr["bv"][l] = np.broadcast_to(val, shape)
    ob.arr["vel"][0] = np.sqrt(ob.arr["bv"][1] ** 2 + ob.arr["bv"][2] ** 2 + ob.arr["bv"][3] ** 2)


def euler_flux(q, k):
    """Exact Euler flux k.F(Q) in conservative variables (air)."""
    rho, mu_, mv, mw, e = q
    u, v, w = mu_ / rho, mv / rho, mw / rho
    p = (GAMMA - 1) * (e - 0.5 * rho * (u * u + v * v + w * w))
    U = k[0] * u + k[1] * v + k[2] * w
    return np.array([rho * U, mu_ * U + k[0] * p, mv * U + k[1] * p, mw * U + k[2] * p, (e + p) * U])


def cons(rho, u, v, w, p):
    return np.array([rho, rho * u, rho * v, rho * w, p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v + w * w)])


def ausm(lib, param, pv4, kvec, jac=1.0):
    """orc_ausm_face on 4 points of primitives (rho,u,v,w,p); metric row k."""
    pv = np.ascontiguousarray(np.array(pv4, dtype=np.float64).ravel())
    ml = np.ascontiguousarray(np.array(kvec, dtype=np.float64))
    mr = ml.copy()
    fc = np.zeros(8)
    lib.orc_ausm_face(C.byref(param), O.dp(pv), O.dp(ml), O.dp(mr), jac, jac, O.dp(fc))
    return fc[:5]


@pytest.mark.parametrize("mach", [0.3, -0.7, 0.999, 1.6, -2.4])
def test_ausm_equal_states_is_exact_flux(lib, mach):
    param = air(viscous=False)
    rho, p = 1.2, 9.0e4
    c = math.sqrt(GAMMA * p / rho)
    kvec = np.array([0.6, -0.3, 0.2])
    n = kvec / np.linalg.norm(kvec)
    vel = mach * c * n + np.array([0.0, 30.0, -10.0]) - np.dot([0.0, 30.0, -10.0], n) * n  # + tangential
    state = [rho, *vel, p]
    jac = 0.37
    got = ausm(lib, param, [state] * 4, kvec, jac)
    ref = euler_flux(cons(*state), jac * kvec)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_ausm_supersonic_upwind(lib):
    param = air(viscous=False)
    kvec = np.array([1.0, 0.0, 0.0])
    left = [1.1, 900.0, 20.0, 0.0, 1.0e5]   # both supersonic to +x
    right = [0.9, 1000.0, -5.0, 3.0, 0.8e5]
    # reconstruction stencil: two points each side; minmod of the jump
    # leaves the face states at the cell values when data are piecewise
    # constant
    got = ausm(lib, param, [left, left, right, right], kvec)
    ref = euler_flux(cons(*left), kvec)
    assert np.allclose(got, ref, rtol=1e-12)
    got = ausm(lib, param, [right, right, left, left], -kvec)  # reversed: flow to -n, upwind = right side
    ref = euler_flux(cons(*left), -kvec)
    assert np.allclose(got, ref, rtol=1e-12)


def test_muscl_linear_is_exact_and_extremum_first_order(lib):
    param = air(viscous=False)   # reconstruction 2 = MUSCL, minmod
    kvec = np.array([0.8, 0.1, 0.0])
    base = np.array([1.0, 150.0, 10.0, -5.0, 9.0e4])
    slope = np.array([0.01, 3.0, -0.5, 0.25, 600.0])
    pts = [base + q * slope for q in range(4)]
    got = ausm(lib, param, pts, kvec)
    mid = base + 1.5 * slope
    ref = euler_flux(cons(*mid), kvec)
    assert np.allclose(got, ref, rtol=1e-12)
    # extremum at the two middle points -> first order (cell values)
    pts = [base, base + slope, base + slope, base]
    got = ausm(lib, param, pts, kvec)
    ref = euler_flux(cons(*(base + slope)), kvec)
    assert np.allclose(got, ref, rtol=1e-12)
    # monotone, unequal jumps 1, 2, 4 (x slope): minmod takes the smaller one
    # on each side, so the face states are p1 + 0.5 and p2 - 1.0 (in units of
    # slope); first-order AUSM+ of those two states must agree
    pts = [base, base + slope, base + 3 * slope, base + 7 * slope]
    got = ausm(lib, param, pts, kvec)
    first = air(viscous=False)
    first.reconstruction = 1
    l_, r_ = base + 1.5 * slope, base + 2.0 * slope
    ref = ausm(lib, first, [l_, l_, r_, r_], kvec)
    assert np.allclose(got, ref, rtol=1e-13)


def test_uniform_flow_has_zero_inviscid_residual(lib):
    param = air(viscous=False)
    ob = cartesian(param, dx=0.125, dy=0.25, dz=0.5)  # dyadic: exact metrics
    set_state(ob, 1.17, 410.0, -35.0, 12.0, 1.01e5)
    ob.arr["dq"][:] = 0
    lib.orc_inviscid_flux(ob.ptr, C.byref(param))
    assert np.abs(ob.arr["dq"][:5, 2:-2, 2:-2, 2:-2]).max() == 0.0


def test_viscous_quadratic_shear(lib):
    param = air(viscous=True)
    dx, dy, dz = 0.1, 0.2, 0.3
    ob = cartesian(param, dx=dx, dy=dy, dz=dz)
    set_state(ob, 1.2, 0.0, 0.0, 0.0, 1.0e5)
    a = 3.0
    y = coords(ob)[1]
    ob.arr["bv"][1] = a * y * y
    ob.arr["vel"][0] = np.abs(ob.arr["bv"][1])
    lib.orc_compute_cv_from_bv(ob.ptr, C.byref(param))
    lib.orc_update_physical_properties(ob.ptr, C.byref(param))
    mu = ob.arr["mul"][0, 2:-2, 2:-2, 2:-2]
    ob.arr["dq"][:] = 0
    lib.orc_viscous_flux(ob.ptr, C.byref(param))
    got = ob.arr["dq"][1, 2:-2, 2:-2, 2:-2]
    ref = 2 * mu * a * dx * dy * dz
    assert np.allclose(got, ref, rtol=1e-10)
    assert np.abs(ob.arr["dq"][0]).max() == 0


def test_local_time_step_closed_form(lib):
    param = air(viscous=True, cfl=0.8)
    dx, dy, dz = 0.1, 0.2, 0.3
    ob = cartesian(param, dx=dx, dy=dy, dz=dz)
    u, v, w, rho, p = 300.0, -40.0, 25.0, 1.1, 0.9e5
    set_state(ob, rho, u, v, w, p)
    lib.orc_compute_cv_from_bv(ob.ptr, C.byref(param))
    lib.orc_update_physical_properties(ob.ptr, C.byref(param))
    lib.orc_local_time_step(ob.ptr, C.byref(param))
    c = math.sqrt(GAMMA * p / rho)
    mu = ob.arr["mul"][0, 3, 3, 3]
    sri = (abs(u) + c) / dx + (abs(v) + c) / dy + (abs(w) + c) / dz
    srv = (1 / dx ** 2 + 1 / dy ** 2 + 1 / dz ** 2) * max(GAMMA, 4 / 3) / rho * mu / param.Pr
    ref = param.cfl / (sri + srv)
    got = ob.arr["dt_local"][0, 2:-2, 2:-2, 2:-2]
    assert np.allclose(got, ref, rtol=1e-12)


@pytest.mark.parametrize("dirn", [0, 1, 2])
def test_dplur_jacobian_matches_finite_difference(lib, dirn):
    param = air(viscous=False)
    ob = cartesian(param, dx=0.1, dy=0.2, dz=0.3)
    state = (1.15, 320.0, -60.0, 45.0, 9.5e4)
    set_state(ob, *state)
    rng = np.random.default_rng(dirn)
    dq = rng.normal(size=5) * np.array([0.01, 3.0, 3.0, 3.0, 2.0e3])
    ob.arr["dq"][:5] = dq[:, None, None, None]
    out = np.zeros(8)
    lib.orc_jac_times_dq_point(ob.ptr, C.byref(param), 3, 3, 3, dirn, 0.0, O.dp(out))
    kvec = np.zeros(3)
    kvec[dirn] = 1.0 / (0.1, 0.2, 0.3)[dirn]
    q0 = cons(*state)
    eps = 1e-6
    fd = (euler_flux(q0 + eps * dq, kvec) - euler_flux(q0 - eps * dq, kvec)) / (2 * eps)
    assert np.allclose(out[:5], fd, rtol=1e-7, atol=1e-9 * np.abs(fd).max())
    # the spectral-radius shift is added to the diagonal
    out2 = np.zeros(8)
    lib.orc_jac_times_dq_point(ob.ptr, C.byref(param), 3, 3, 3, dirn, 7.0, O.dp(out2))
    assert np.allclose(out2[:5] - out[:5], 7.0 * dq, rtol=1e-12)


def test_sst_source_simple_shear(lib):
    param = air(rans=True, viscous=True)
    dx, dy, dz = 0.1, 0.2, 0.3
    ob = cartesian(param, dx=dx, dy=dy, dz=dz)
    rho, tke, om, a, dwall = 1.2, 2.0, 400.0, 150.0, 0.05
    set_state(ob, rho, 0.0, 0.0, 0.0, 1.0e5)
    y = coords(ob)[1]
    ob.arr["bv"][1] = a * y
    ob.arr["sv"][0] = tke
    ob.arr["sv"][1] = om
    ob.arr["wall_distance"][0] = dwall
    lib.orc_compute_cv_from_bv(ob.ptr, C.byref(param))
    lib.orc_update_physical_properties(ob.ptr, C.byref(param))
    ob.arr["dq"][:] = 0
    lib.orc_compute_source(ob.ptr, C.byref(param))
    mu = ob.arr["mul"][0, 3, 3, 3]
    # Menter SST constants
    beta_star, a1, sig_w2 = 0.09, 0.31, 0.856
    beta1, beta2, kappa, sig_w1 = 0.075, 0.0828, 0.41, 0.5
    g1 = beta1 / beta_star - sig_w1 * kappa ** 2 / math.sqrt(beta_star)
    g2 = beta2 / beta_star - sig_w2 * kappa ** 2 / math.sqrt(beta_star)
    p1 = math.sqrt(tke) / (0.09 * om * dwall)
    p2 = 500 * mu / (rho * dwall ** 2 * om)
    f2 = math.tanh(max(2 * p1, p2) ** 2)
    p3 = 4 * rho * sig_w2 * tke / (1e-20 * dwall ** 2)   # no k-omega cross diffusion
    f1 = math.tanh(min(max(p1, p2), p3) ** 4)
    mut = a1 * rho * tke / max(a1 * om, abs(a) * f2)
    prod_k = mut * a * a
    jac = dx * dy * dz
    ref_k = jac * (prod_k - beta_star * rho * tke * om)
    beta = beta2 + (beta1 - beta2) * f1
    gam = g2 + (g1 - g2) * f1
    ref_w = jac * (gam * rho / mut * prod_k - beta * rho * om * om)
    assert math.isclose(ob.arr["mut"][0, 3, 3, 3], mut, rel_tol=1e-12)
    assert math.isclose(ob.arr["dq"][5, 3, 3, 3], ref_k, rel_tol=1e-10)
    assert math.isclose(ob.arr["dq"][6, 3, 3, 3], ref_w, rel_tol=1e-10)


def test_symmetry_reflection(lib):
    param = air(viscous=False)
    ob = cartesian(param)
    rng = np.random.default_rng(5)
    shape = ob.shape
    for l, (lo, hi) in enumerate(((1.0, 1.3), (100, 300), (-50, 50), (-80, 80), (9e4, 1.1e5))):
        ob.arr["bv"][l] = rng.uniform(lo, hi, shape)
    ob.arr["bv"][5] = ob.arr["bv"][4] * MW / (ob.arr["bv"][0] * RU)
    lib.orc_compute_cv_from_bv(ob.ptr, C.byref(param))
    before = ob.arr["bv"].copy()
    bc = L.BC()
    bc.n_face = 1
    f = bc.face[0]
    mx, my, mz, g = ob.mx, ob.my, ob.mz, 2
    f.range_start[:] = [-g, -g, 0]
    f.range_end[:] = [mx - 1 + g, my - 1 + g, 0]
    f.face, f.direction, f.kind, f.slot = 2, -1, 3, 0
    ob_bc = O.oracle_bc(bc)
    lib.orc_apply_bc(ob.ptr, C.byref(param), C.byref(ob_bc))
    bv = ob.arr["bv"]
    K = lambda k: k + g  # padded index
    # face node: the interior neighbour's state without its normal (w) velocity
    assert np.array_equal(bv[3, K(0)], np.zeros_like(bv[3, K(0)]))
    for l in (0, 1, 2, 4, 5):
        assert np.array_equal(bv[l, K(0)], before[l, K(1)])
    for q in (1, 2):
        assert np.array_equal(bv[3, K(-q)], -before[3, K(q)])
        for l in (0, 1, 2, 4, 5):
            assert np.array_equal(bv[l, K(-q)], before[l, K(q)])


def test_update_recovers_primitives(lib):
    param = air(viscous=False)
    ob = cartesian(param)
    set_state(ob, 1.2, 300.0, 20.0, -10.0, 1.0e5)
    lib.orc_compute_cv_from_bv(ob.ptr, C.byref(param))
    cv0 = ob.arr["cv"][:5, 3, 3, 3].copy()
    dq = np.array([0.01, 2.0, -1.0, 0.5, 3000.0])
    ob.arr["dq"][:5] = dq[:, None, None, None]
    ob.arr["dt_local"][:] = 3.0e-3
    jac = ob.arr["jac"][0, 3, 3, 3]
    assert np.isclose(jac, 0.1 * 0.2 * 0.3, rtol=1e-12)
    lib.orc_update_cv_and_bv(ob.ptr, C.byref(param))
    q = cv0 + dq * 3.0e-3 / jac
    rho = q[0]
    u, v, w = q[1:4] / rho
    p = (GAMMA - 1) * (q[4] - 0.5 * rho * (u * u + v * v + w * w))
    got = ob.arr["bv"][:6, 3, 3, 3]
    ref = np.array([rho, u, v, w, p, p * MW / (rho * RU)])
    assert np.allclose(got, ref, rtol=1e-13)


def test_limit_flow_neighbour_average_and_mz_bound(lib):
    param = air(viscous=False)
    ob = cartesian(param, mx=9, my=6, mz=5)
    rng = np.random.default_rng(6)
    shape = ob.shape
    for l, (lo, hi) in enumerate(((1.0, 1.3), (100, 300), (-50, 50), (-80, 80), (9e4, 1.1e5))):
        ob.arr["bv"][l] = rng.uniform(lo, hi, shape)
    ob.arr["bv"][5] = ob.arr["bv"][4] * MW / (ob.arr["bv"][0] * RU)
    lib.orc_compute_cv_from_bv(ob.ptr, C.byref(param))
    g = 2
    cells = [(2, 2, 2), (4, 3, 2)]  # the second has i = mz-1: neighbours with i >= mz are skipped
    for (i, j, k) in cells:
        ob.arr["bv"][4, k + g, j + g, i + g] = -1.0
    before = ob.arr["bv"].copy()
    n = lib.orc_limit_flow(ob.ptr, C.byref(param))
    assert n == 2
    for (i, j, k) in cells:
        acc, cnt = np.zeros(5), 0
        for dk in (-1, 0, 1):
            for dj in (-1, 0, 1):
                for di in (-1, 0, 1):
                    i1, j1, k1 = i + di, j + dj, k + dk
                    if not (0 <= k1 < ob.mz and 0 <= j1 < ob.my and 0 <= i1 < ob.mz):
                        continue
                    if (i1, j1, k1) in cells:
                        continue
                    acc += before[:5, k1 + g, j1 + g, i1 + g]
                    cnt += 1
        ref = acc / cnt
        assert np.allclose(ob.arr["bv"][:5, k + g, j + g, i + g], ref, rtol=1e-13)


@pytest.mark.parametrize("recon", [1, 3])
def test_first_order_and_nnd2_reconstruction(lib, recon):
    param = air(viscous=False)
    param.reconstruction = recon
    kvec = np.array([0.8, 0.1, 0.0])
    base = np.array([1.0, 150.0, 10.0, -5.0, 9.0e4])
    slope = np.array([0.01, 3.0, -0.5, 0.25, 600.0])
    pts = [base + q * slope for q in range(4)]
    got = ausm(lib, param, pts, kvec)
    if recon == 3:  # NND-2 is exact for linear data: face states at the midpoint
        ref = euler_flux(cons(*(base + 1.5 * slope)), kvec)
        assert np.allclose(got, ref, rtol=1e-12)
    else:           # first order: AUSM+ of the two cell values
        alt = ausm(lib, param, [pts[1], pts[1], pts[2], pts[2]], kvec)
        assert np.array_equal(got, alt)
        assert not np.allclose(got, euler_flux(cons(*(base + 1.5 * slope)), kvec), rtol=1e-6)


def test_sutherland_known_values(lib):
    # Sutherland's law with mu0 = 1.716e-5 Pa s at 273 K, S = 111 K (src/Transport.cu:7)
    assert lib.orc_sutherland(273.0) == pytest.approx(1.716e-5, rel=1e-15)
    t = 500.0
    assert lib.orc_sutherland(t) == pytest.approx(1.716e-5 * (t / 273) ** 1.5 * 384 / (t + 111), rel=1e-14)


def boundary(kind, face, direction, ob, slot=0):
    bc = L.BC()
    bc.n_face = 1
    f = bc.face[0]
    g = 2
    lo = [-g, -g, -g]
    hi = [ob.mx - 1 + g, ob.my - 1 + g, ob.mz - 1 + g]
    n = (ob.mx, ob.my, ob.mz)[face]
    lo[face] = hi[face] = 0 if direction < 0 else n - 1
    f.range_start[:] = lo
    f.range_end[:] = hi
    f.face, f.direction, f.kind, f.slot = face, direction, kind, slot
    return bc


def random_state(ob, seed, rans=False):
    rng = np.random.default_rng(seed)
    shape = ob.shape
    for l, (lo, hi) in enumerate(((1.0, 1.3), (100, 300), (-50, 50), (-80, 80), (9e4, 1.1e5))):
        ob.arr["bv"][l] = rng.uniform(lo, hi, shape)
    ob.arr["bv"][5] = ob.arr["bv"][4] * MW / (ob.arr["bv"][0] * RU)
    if rans:
        ob.arr["sv"][0] = rng.uniform(1.0, 5.0, shape)
        ob.arr["sv"][1] = rng.uniform(100.0, 900.0, shape)
        ob.arr["mut"][0] = rng.uniform(1e-5, 1e-4, shape)
    return rng


def test_isothermal_wall_sst(lib):
    param = air(rans=True, viscous=True)
    ob = cartesian(param)
    random_state(ob, 8, rans=True)
    d = 0.2
    ob.arr["wall_distance"][0] = d
    lib.orc_compute_cv_from_bv(ob.ptr, C.byref(param))
    before = {n: ob.arr[n].copy() for n in ("bv", "sv", "mut")}
    bc = boundary(2, 1, -1, ob)   # j = 0 wall
    tw = 350.0
    bc.wall[0].isothermal = 1
    bc.wall[0].temperature = tw
    lib.orc_apply_bc(ob.ptr, C.byref(param), C.byref(O.oracle_bc(bc)))
    g = 2
    J = lambda j: j + g
    bv, sv = ob.arr["bv"], ob.arr["sv"]
    # face node: no slip, pressure from the first interior point, wall temperature
    p1 = before["bv"][4, :, J(1)]
    rho_w = p1 * MW / (tw * RU)
    assert np.allclose(bv[0, :, J(0)], rho_w, rtol=1e-14)
    assert np.all(bv[1:4, :, J(0)] == 0)
    assert np.array_equal(bv[4, :, J(0)], p1)
    assert np.all(bv[5, :, J(0)] == tw)
    mu_w = lib.orc_sutherland(tw)
    assert np.all(sv[0, :, J(0)] == 0)
    assert np.allclose(sv[1, :, J(0)], 60 * mu_w / (rho_w * 0.075 * d * d), rtol=1e-14)
    # ghosts: velocity reversed, T mirrored about the wall temperature, same pressure
    for q in (1, 2):
        src = before["bv"][:, :, J(q)]
        t_g = 2 * tw - src[5]
        t_g = np.where(t_g <= 0.1 * tw, tw, t_g)
        assert np.array_equal(bv[1:4, :, J(-q)], -src[1:4])
        assert np.array_equal(bv[4, :, J(-q)], src[4])
        assert np.allclose(bv[5, :, J(-q)], t_g, rtol=1e-15)
        assert np.allclose(bv[0, :, J(-q)], src[4] * MW / (t_g * RU), rtol=1e-14)
        assert np.all(sv[0, :, J(-q)] == 0) and np.all(ob.arr["mut"][0, :, J(-q)] == 0)
        assert np.array_equal(sv[1, :, J(-q)], sv[1, :, J(0)])


def test_adiabatic_wall_takes_interior_temperature(lib):
    param = air(viscous=True)
    ob = cartesian(param)
    random_state(ob, 9)
    lib.orc_compute_cv_from_bv(ob.ptr, C.byref(param))
    before = ob.arr["bv"].copy()
    bc = boundary(2, 0, 1, ob)   # i = mx-1 wall
    bc.wall[0].isothermal = 0
    lib.orc_apply_bc(ob.ptr, C.byref(param), C.byref(O.oracle_bc(bc)))
    g, n = 2, ob.mx - 1
    I = lambda i: i + g
    bv = ob.arr["bv"]
    assert np.array_equal(bv[5, :, :, I(n)], before[5, :, :, I(n - 1)])
    for q in (1, 2):
        assert np.array_equal(bv[5, :, :, I(n + q)], before[5, :, :, I(n - q)])
        assert np.array_equal(bv[1:4, :, :, I(n + q)], -before[1:4, :, :, I(n - q)])


def test_inflow_and_outflow_ghosts(lib):
    param = air(rans=True, viscous=True)
    ob = cartesian(param)
    random_state(ob, 10, rans=True)
    before = {n: ob.arr[n].copy() for n in ("bv", "sv", "mut")}
    bc = boundary(5, 0, -1, ob)  # i = 0 inflow
    inf = bc.inflow[0]
    vals = dict(density=1.05, u=600.0, v=-3.0, w=2.0, pressure=8.0e4, temperature=265.0, mut=3.3e-5)
    for k_, v_ in vals.items():
        setattr(inf, k_, v_)
    inf.sv[0], inf.sv[1] = 0.7, 1234.0
    lib.orc_apply_bc(ob.ptr, C.byref(param), C.byref(O.oracle_bc(bc)))
    g = 2
    for q in (1, 2):
        col = ob.arr["bv"][:, :, :, g - q]
        for l, key in enumerate(("density", "u", "v", "w", "pressure", "temperature")):
            assert np.all(col[l] == vals[key])
        assert np.all(ob.arr["sv"][0, :, :, g - q] == 0.7) and np.all(ob.arr["sv"][1, :, :, g - q] == 1234.0)
        assert np.all(ob.arr["mut"][0, :, :, g - q] == 3.3e-5)
    # the boundary points themselves are untouched by an inflow face
    assert np.array_equal(ob.arr["bv"][:, :, :, g], before["bv"][:, :, :, g])
    bc = boundary(6, 2, 1, ob)   # k = mz-1 outflow: ghosts copy the boundary point
    lib.orc_apply_bc(ob.ptr, C.byref(param), C.byref(O.oracle_bc(bc)))
    kb = g + ob.mz - 1
    for q in (1, 2):
        assert np.array_equal(ob.arr["bv"][:, kb + q], ob.arr["bv"][:, kb])
        assert np.array_equal(ob.arr["sv"][:, kb + q], ob.arr["sv"][:, kb])
        assert np.array_equal(ob.arr["mut"][0, kb + q], ob.arr["mut"][0, kb])

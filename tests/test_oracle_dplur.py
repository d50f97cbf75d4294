"""Pin of the oracle's DPLUR sweeps (oracle/orc_flow.c, restating
src/DPLURHPP.cuh:13-273) against their textbook meaning, on the CPU.

DPLUR is point-Jacobi on the implicit-Euler linearisation of a first-order
flux split with the spectral radius rho: the flux through the face between
points p and p+e_d is linearised as A+_p dQ_p + A-_{p+e_d} dQ_{p+e_d} with
A+- = (A +- rho I)/2, A = d(grad xi_d . F)/dQ at that point (unscaled
metric).  In the residual units the update works in (x = J dQ / dt), point p
then satisfies

    (1 + dt_p sum_d rho_{p,d}) x_p
        - dt_p sum_d [ A+_{p-e_d} x_{p-e_d} - A-_{p+e_d} x_{p+e_d} ] = R_p,

with neighbours outside the block dropped (dq is not exchanged).  The test
builds the A's by central finite differences of the exact Euler flux in
conservative variables -- for air, and for the 9-species H2/air mixture with
T solved from the internal energy by Newton on the NASA enthalpies -- and
checks: the oracle's first two sweeps equal the Jacobi iterates of this
system, and 300 sweeps reach its numpy solution.  The same mixture flux pins
the mixture branch of AUSM+ (equal states give the exact flux)."""
import math

import numpy as np
import pytest

import oracle_ffi as O
from openstc_amd import cases

RU = 8314.462618
MW_AIR = 28.9647


def air_setup():
    fs = cases.air_inflow(2.0, 101325.0, 300.0, rans=False)
    p = cases.air_param(rans=False, viscous=False, implicit=True, dplur_steps=2, cfl=2.0, freestream=fs)
    return O.copy_struct(O.OrcParam, p)


def mix_setup():
    mech, _ = cases.load_mechanism(with_reactions=False)
    fs = cases.air_inflow(2.0, 101325.0, 300.0, rans=False)
    p = cases.mixture_param(mech, rans=False, implicit=True, dplur_steps=2, cfl=2.0, chem_method=0,
                            freestream=fs)
    p.viscous_scheme = 0
    return O.copy_struct(O.OrcParam, p)


class Thermo:
    """Per-species h(T), cp(T) from the oracle's NASA evaluation (pinned
    separately against the closed forms) and the exact mixture Euler flux."""

    def __init__(self, lib, prm):
        self.lib, self.prm = lib, prm
        self.ns = prm.n_spec
        self.w = np.array(prm.mw[:self.ns]) if self.ns else None

    def h_cp(self, t):
        h, cp = np.zeros(O.MAX_SPEC), np.zeros(O.MAX_SPEC)
        self.lib.orc_enthalpy_and_cp(t, O.dp(h), O.dp(cp), self.prm)
        return h[:self.ns], cp[:self.ns]

    def temperature(self, rho_k, rho_e, t0):
        r = RU / self.w
        t = t0
        for _ in range(60):
            h, cp = self.h_cp(t)
            f = np.dot(rho_k, h - r * t) - rho_e
            t_new = t - f / np.dot(rho_k, cp - r)
            if abs(t_new - t) <= 1e-14 * t:
                return t_new
            t = t_new
        return t

    def flux(self, q, k, t0):
        rho, m, rho_E = q[0], q[1:4], q[4]
        u = m / rho
        U = float(np.dot(k, u))
        if self.ns == 0:
            p = 0.4 * (rho_E - 0.5 * rho * np.dot(u, u))
            extra = []
        else:
            rho_k = q[5:5 + self.ns]
            rho_e = rho_E - 0.5 * np.dot(m, m) / rho
            t = self.temperature(rho_k, rho_e, t0)
            p = t * RU * np.sum(rho_k / self.w)
            extra = list(rho_k * U)
        return np.array([rho * U, *(m * U + k * p), (rho_E + p) * U, *extra])


def build_block(lib, prm, seed, mx=4, my=3, mz=3):
    rng = np.random.default_rng(seed)
    ob = O.OracleBlock(mx, my, mz, 2, prm)
    k, j, i = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    x = 0.1 * i + 0.012 * np.sin(1.3 * j + 0.7 * k)
    y = 0.08 * j + 0.01 * np.cos(0.9 * i)
    z = 0.07 * k + 0.008 * np.sin(i + j)
    ob.geometry(x, y, z)
    shape = ob.shape
    bv, sv = ob.arr["bv"], ob.arr["sv"]
    bv[0] = rng.uniform(0.6, 1.2, shape)
    bv[1] = rng.uniform(-200, 500, shape)
    bv[2] = rng.uniform(-150, 150, shape)
    bv[3] = rng.uniform(-100, 100, shape)
    t = rng.uniform(600.0, 1800.0, shape)
    ns = prm.n_spec
    if ns:
        y_ = rng.dirichlet(np.ones(ns), size=shape)           # (..., ns)
        sv[:ns] = np.moveaxis(y_, -1, 0)
        w = np.array(prm.mw[:ns])
        mw = 1.0 / np.tensordot(1.0 / w, sv[:ns], axes=1)
    else:
        mw = MW_AIR
    bv[4] = bv[0] * RU * t / mw
    bv[5] = t
    ob.arr["vel"][0] = np.sqrt(bv[1] ** 2 + bv[2] ** 2 + bv[3] ** 2)
    lib.orc_compute_cv_from_bv(ob.ptr, prm)
    lib.orc_update_physical_properties(ob.ptr, prm)
    lib.orc_local_time_step(ob.ptr, prm)
    nv = prm.n_var
    ob.arr["dq"][:] = 0
    scale = np.array([0.5, 200.0, 200.0, 200.0, 4.0e5] + [0.05] * (nv - 5))
    r = rng.normal(size=(nv, mz, my, mx)) * scale[:, None, None, None]
    g = 2
    ob.arr["dq"][:nv, g:g + mz, g:g + my, g:g + mx] = r
    return ob, r


def point_data(ob, prm, th, p):
    """State, conservative vector, metric rows, sound speed and spectral radii
    at padded point p = (k, j, i)."""
    k, j, i = p
    a = ob.arr
    rho = a["bv"][0][p]
    u = np.array([a["bv"][l][p] for l in (1, 2, 3)])
    t = a["bv"][5][p]
    pres = a["bv"][4][p]
    ns = prm.n_spec
    if ns:
        y = np.array([a["sv"][l][p] for l in range(ns)])
        h, cp = th.h_cp(t)
        gam = np.dot(y, cp) / np.dot(y, cp - RU / th.w)
        e = np.dot(y, h) - pres / rho
        q = np.concatenate([[rho], rho * u, [rho * (e + 0.5 * np.dot(u, u))], rho * y])
    else:
        gam = 1.4
        q = np.concatenate([[rho], rho * u, [pres / 0.4 + 0.5 * rho * np.dot(u, u)]])
    c = math.sqrt(gam * pres / rho)
    rows = [np.array([a["metric"][3 * d + cc][p] for cc in range(3)]) for d in range(3)]
    sr = [abs(np.dot(rows[d], u)) + c * np.linalg.norm(rows[d]) for d in range(3)]
    return q, rows, sr, t


def fd_jacobian(th, q, kvec, t0):
    n = len(q)
    A = np.empty((n, n))
    for c in range(n):
        eps = 1e-6 * max(abs(q[c]), 1e-3 * abs(q[0]))
        dq = np.zeros(n)
        dq[c] = eps
        A[:, c] = (th.flux(q + dq, kvec, t0) - th.flux(q - dq, kvec, t0)) / (2 * eps)
    return A


def linear_system(ob, prm, th):
    mx, my, mz, g = ob.mx, ob.my, ob.mz, ob.ngg
    nv = prm.n_var
    pts = [(k, j, i) for k in range(mz) for j in range(my) for i in range(mx)]
    index = {p: n for n, p in enumerate(pts)}
    data = {}
    for (k, j, i) in pts:
        pp = (k + g, j + g, i + g)
        q, rows, sr, t = point_data(ob, prm, th, pp)
        data[(k, j, i)] = (q, rows, sr, t, [fd_jacobian(th, q, rows[d], t) for d in range(3)])
    N = len(pts)
    D = np.zeros(N)
    off = np.zeros((N * nv, N * nv))   # dt_p * sum of neighbour blocks
    I = np.eye(nv)
    for p in pts:
        n = index[p]
        dt = ob.arr["dt_local"][0][(p[0] + g, p[1] + g, p[2] + g)]
        D[n] = 1 + dt * sum(data[p][2])
        for d in range(3):
            step = [0, 0, 0]
            step[2 - d] = 1            # d = 0 -> i, 1 -> j, 2 -> k; p is (k, j, i)
            lo = tuple(p[c] - step[c] for c in range(3))
            hi = tuple(p[c] + step[c] for c in range(3))
            if lo in index:
                A, sr = data[lo][4][d], data[lo][2][d]
                off[n * nv:(n + 1) * nv, index[lo] * nv:(index[lo] + 1) * nv] += dt * 0.5 * (A + sr * I)
            if hi in index:
                A, sr = data[hi][4][d], data[hi][2][d]
                off[n * nv:(n + 1) * nv, index[hi] * nv:(index[hi] + 1) * nv] -= dt * 0.5 * (A - sr * I)
    Mtx = np.diag(np.repeat(D, nv)) - off
    return Mtx, np.repeat(D, nv), off, pts


@pytest.mark.parametrize("kind", ["air", "mixture"])
def test_dplur_is_jacobi_on_the_split_linearisation(kind):
    lib = O.load()
    prm = air_setup() if kind == "air" else mix_setup()
    th = Thermo(lib, prm)
    ob, r = build_block(lib, prm, seed=11 if kind == "air" else 12)
    nv, g = prm.n_var, ob.ngg
    Mtx, Dv, off, pts = linear_system(ob, prm, th)
    rhs = np.concatenate([r[:, k, j, i] for (k, j, i) in pts])
    sl = (slice(0, nv), slice(g, g + ob.mz), slice(g, g + ob.my), slice(g, g + ob.mx))

    def oracle(steps):
        ob.arr["dq"][sl] = r
        prm.dplur_inner_step = steps
        lib.orc_implicit_treatment(ob.ptr, prm)
        out = ob.arr["dq"][sl]
        return np.concatenate([out[:, k, j, i] for (k, j, i) in pts])

    x = rhs / Dv
    for steps in (1, 2):
        x = (rhs + off @ x) / Dv
        got = oracle(steps)
        err = np.abs(got - x).max() / np.abs(x).max()
        assert err < 2e-7, (steps, err)
    exact = np.linalg.solve(Mtx, rhs)
    got = oracle(300)
    err = np.abs(got - exact).max() / np.abs(exact).max()
    assert err < 2e-7, err
    # and the sweeps do matter: two sweeps are far from the DQ0 start
    assert np.abs(exact - rhs / Dv).max() > 1e-3 * np.abs(exact).max()


@pytest.mark.parametrize("mach", [0.4, -1.7])
def test_mixture_ausm_equal_states_is_exact_flux(mach):
    """AUSM+ with mixture reconstruction (E and gamma from the reconstructed
    T, src/InviscidScheme.cuh:33-61): equal states give the exact mixture
    Euler flux, species fluxes rho Y_k U included."""
    lib = O.load()
    prm = mix_setup()
    th = Thermo(lib, prm)
    ns = prm.n_spec
    rng = np.random.default_rng(3)
    y = rng.dirichlet(np.ones(ns))
    rho, t = 0.7, 1400.0
    pres = rho * RU * t * np.sum(y / th.w)
    h, cp = th.h_cp(t)
    gam = np.dot(y, cp) / np.dot(y, cp - RU / th.w)
    c = math.sqrt(gam * pres / rho)
    kvec = np.array([0.5, 0.2, -0.4])
    u = mach * c * kvec / np.linalg.norm(kvec) + np.array([0.0, 15.0, 20.0])
    state = np.concatenate([[rho], u, [pres], y])
    pv = np.ascontiguousarray(np.tile(state, 4))
    ml = np.ascontiguousarray(kvec)
    fc = np.zeros(prm.n_var + 2)
    lib.orc_ausm_face(prm, O.dp(pv), O.dp(ml), O.dp(ml.copy()), 1.0, 1.0, O.dp(fc))
    e = np.dot(y, h) - pres / rho
    q = np.concatenate([[rho], rho * u, [rho * (e + 0.5 * np.dot(u, u))], rho * y])
    ref = th.flux(q, kvec, t)
    assert np.allclose(fc[:prm.n_var], ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())

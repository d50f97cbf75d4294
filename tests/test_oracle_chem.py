"""CPU pins of the chemistry side: the CHEMKIN reader against the reference's
input files, and the oracle's thermo / kinetics / point-implicit routines
against closed forms, invariants and textbook formulas."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import oracle_ffi as O
from openstc_amd import _lib as L
from openstc_amd import cases

RU = 8314.462618
RC = 1.98720425864083
PATM = 101325.0


@pytest.fixture(scope="module")
def mech():
    p, info = cases.load_mechanism(True)
    return p, info


def orc_param(p):
    return O.copy_struct(O.OrcParam, p)


def test_parser_species_and_reactions(mech):
    """Species order, reaction count and types of the sample mechanism."""
    p, info = mech
    assert cases.species_names(info) == ["H2", "H", "O2", "O", "OH", "HO2", "H2O2", "H2O", "N2"]
    assert p.n_spec == 9 and p.n_reac == 19
    types = list(p.reac_type[:19])
    assert [i for i, t in enumerate(types) if t == 4] == [4, 5, 6, 7]
    assert [i for i, t in enumerate(types) if t == 6] == [8, 14]
    assert [i for i, t in enumerate(types) if t == 3] == [13, 18]
    assert all(t == 1 for i, t in enumerate(types) if i not in (4, 5, 6, 7, 8, 13, 14, 18))
    # reaction orders: sum(nu_b) - sum(nu_f)
    assert list(p.reac_order[:19]) == [0, 0, 0, 0, 1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0]


def test_parser_rates_and_quirks(mech):
    p, _ = mech
    # main-line Ea converted from cal/mol to K
    assert p.A[0] == 3.55e15 and p.b[0] == -0.40 and p.Ea[0] == pytest.approx(1.66e4 / RC, rel=1e-15)
    # DUPLICATE pair merged into the first slot
    assert p.A[13] == 4.20e14 and p.A2[13] == 1.30e11 and p.Ea2[13] == pytest.approx(-1.63e3 / RC, rel=1e-15)
    assert p.A[18] == 1.00e12 and p.A2[18] == 5.80e14 and p.Ea2[18] == pytest.approx(9.56e3 / RC, rel=1e-15)
    # LOW activation energies are stored as read (the reference's units quirk)
    assert (p.A2[8], p.b2[8], p.Ea2[8]) == (6.37e20, -1.72, 525.0)
    assert (p.A2[14], p.b2[14], p.Ea2[14]) == (1.2e17, 0.0, 45500.0)
    assert (p.troe_alpha[8], p.troe_t3[8], p.troe_t1[8]) == (0.8, 1e-30, 1e30)
    # third-body efficiencies (default 1)
    tb = lambda r: [p.third_body_coeff[r * 20 + s] for s in range(9)]
    assert tb(4) == [2.5, 1, 1, 1, 1, 1, 1, 12.0, 1]
    assert tb(8) == [2.0, 1, 0.78, 1, 1, 1, 1, 11.0, 1]
    assert tb(0) == [1.0] * 9


def test_parser_thermo_columns(mech):
    """NASA-7 coefficients and molecular weights vs an independent fixed-column read."""
    p, info = mech
    names = cases.species_names(info)
    lines = open(os.path.join(cases.DATA, "therm.dat")).read().split("\n")
    aw = {"H": 1.008, "O": 15.999, "N": 14.007}
    seen = 0
    for q, line in enumerate(lines):
        name = line[:18].split()[0] if line[:18].strip() else ""
        if name not in names or len(line) < 79 or line[79:80] != "1":
            continue
        s = names.index(name)
        block = "".join(l[:75] for l in lines[q + 1:q + 4])
        nums = [float(block[15 * m:15 * m + 15]) for m in range(14)]
        assert [p.high_temp_coeff[s * 7 + m] for m in range(7)] == nums[:7]
        assert [p.low_temp_coeff[s * 7 + m] for m in range(7)] == nums[7:14]
        comp = {}
        for m in range(4):
            f = line[24 + 5 * m:29 + 5 * m].strip()
            if not f or f[0] == "0":
                break
            el, n = f.split()
            comp[el] = int(n)
        mw = sum(aw[e] * comp.get(e, 0) for e in ("H", "N", "O"))
        assert p.mw[s] == mw
        assert p.t_low[s] == float(line[45:55]) and p.t_high[s] == float(line[55:65])
        seen += 1
    assert seen == 9
    assert p.mw[names.index("H2O")] == 2 * 1.008 + 15.999


def nasa(p, s, t):
    c = p.low_temp_coeff if t < p.t_mid[s] else p.high_temp_coeff
    return [c[s * 7 + m] for m in range(7)]


def test_thermo_closed_forms(mech):
    """h, cp, g/RT against the NASA-7 closed forms, incl. the T<T_low branches."""
    p, _ = mech
    op = orc_param(p)
    lib = O.load()
    h = np.zeros(20)
    cp = np.zeros(20)
    g = np.zeros(20)
    for t in (250.0, 700.0, 999.0, 1000.0, 1800.0, 3200.0):
        lib.orc_enthalpy_and_cp(t, O.dp(h), O.dp(cp), C.byref(op))
        lib.orc_gibbs_div_rt(t, C.byref(op), O.dp(g))
        for s in range(9):
            a = nasa(p, s, t)
            if t >= p.t_low[s]:
                h_rt = a[0] + a[1] * t / 2 + a[2] * t ** 2 / 3 + a[3] * t ** 3 / 4 + a[4] * t ** 4 / 5 + a[5] / t
                cp_r = a[0] + a[1] * t + a[2] * t ** 2 + a[3] * t ** 3 + a[4] * t ** 4
                s_r = a[0] * math.log(t) + a[1] * t + a[2] * t ** 2 / 2 + a[3] * t ** 3 / 3 + a[4] * t ** 4 / 4 + a[6]
                assert h[s] == pytest.approx(h_rt * RU * t / p.mw[s], rel=1e-12, abs=1e-6)
                assert cp[s] == pytest.approx(cp_r * RU / p.mw[s], rel=1e-13)
                assert g[s] == pytest.approx(h_rt - s_r, rel=1e-11, abs=1e-9)
    # below T_low: h linear with slope cp(T_low), cp constant, g frozen at T_low
    s = 8  # N2, T_low = 300
    lib.orc_enthalpy_and_cp(300.0, O.dp(h), O.dp(cp), C.byref(op))
    h300, cp300 = h[s], cp[s]
    lib.orc_gibbs_div_rt(300.0, C.byref(op), O.dp(g))
    g300 = g[s]
    lib.orc_enthalpy_and_cp(250.0, O.dp(h), O.dp(cp), C.byref(op))
    assert cp[s] == pytest.approx(cp300, rel=1e-15)
    assert h[s] == pytest.approx(h300 - 50 * cp300, rel=1e-13)
    lib.orc_gibbs_div_rt(250.0, C.byref(op), O.dp(g))
    assert g[s] == pytest.approx(g300, rel=1e-15)


def test_thermo_continuity_at_tmid(mech):
    p, _ = mech
    op = orc_param(p)
    lib = O.load()
    lo, hi = np.zeros(20), np.zeros(20)
    clo, chi = np.zeros(20), np.zeros(20)
    lib.orc_enthalpy_and_cp(1000.0 - 1e-9, O.dp(lo), O.dp(clo), C.byref(op))
    lib.orc_enthalpy_and_cp(1000.0, O.dp(hi), O.dp(chi), C.byref(op))
    for s in range(9):
        assert hi[s] == pytest.approx(lo[s], rel=2e-3, abs=2e4)  # fit tolerance of the tables
        assert chi[s] == pytest.approx(clo[s], rel=2e-3)


def random_states(n, seed=0):
    rng = np.random.default_rng(seed)
    t = rng.uniform(600, 2800, n)
    y = rng.dirichlet(np.ones(9) * 0.6, n)
    y = np.maximum(y, 1e-8)
    y /= y.sum(1, keepdims=True)
    rho = rng.uniform(0.05, 2.0, n)
    return rho, t, y


def chem_point(op, rho, t, y):
    lib = O.load()
    om, od, jac, jd = np.zeros(20), np.zeros(20), np.zeros(400), np.zeros(20)
    kf, kb = np.zeros(50), np.zeros(50)
    yy = np.ascontiguousarray(np.pad(y, (0, 20 - len(y))))
    lib.orc_chem_source_point(C.byref(op), rho, t, O.dp(yy), O.dp(om), O.dp(od), O.dp(jac), O.dp(jd), O.dp(kf),
                              O.dp(kb))
    return om[:9], od[:9], jac[:81].reshape(9, 9), jd[:9], kf[:19], kb[:19]


def test_element_conservation(mech):
    """sum_k a_ek * omega_k / W_k = 0 for H, O, N."""
    p, info = mech
    op = orc_param(p)
    comp = np.array([[info.elem_comp[s * L.MAX_ELEM + e] for e in range(info.n_elem)] for s in range(9)])
    mw = np.array(p.mw[:9])
    rho, t, y = random_states(40, 1)
    for i in range(40):
        om, od, *_ = chem_point(op, rho[i], t[i], y[i])
        flux = comp.T @ (om / mw)
        scale = comp.T @ (od / mw) + 1e-300
        assert np.all(np.abs(flux) <= 1e-12 * scale), (flux, scale)


def test_rates_arrhenius_troe_and_detailed_balance(mech):
    """kf by Arrhenius / third-body / Troe (Gilbert et al. 1983) and kb = kf/Kc
    with Kc from NASA-7 Gibbs energies, computed independently in numpy."""
    p, _ = mech
    op = orc_param(p)
    rho, t, y = random_states(10, 2)
    mw = np.array(p.mw[:9])
    nf = np.array(p.stoi_f[:19 * 20]).reshape(19, 20)[:, :9]
    nb = np.array(p.stoi_b[:19 * 20]).reshape(19, 20)[:, :9]
    tb = np.array(p.third_body_coeff[:19 * 20]).reshape(19, 20)[:, :9]
    for i in range(10):
        T = t[i]
        c = rho[i] * y[i] / mw * 1e-3
        _, _, _, _, kf, kb = chem_point(op, rho[i], T, y[i])
        for r in range(19):
            k = p.A[r] * T ** p.b[r] * math.exp(-p.Ea[r] / T)
            typ = p.reac_type[r]
            if typ == 3:
                k += p.A2[r] * T ** p.b2[r] * math.exp(-p.Ea2[r] / T)
            if typ == 4:
                k *= tb[r] @ c
            if typ in (5, 6, 7):
                k0 = p.A2[r] * T ** p.b2[r] * math.exp(-p.Ea2[r] / T)
                pr = k0 * (tb[r] @ c) / k
                F = 1.0
                if typ >= 6:
                    a = p.troe_alpha[r]
                    fc = (1 - a) * math.exp(-T / p.troe_t3[r]) + a * math.exp(-T / p.troe_t1[r])
                    if typ == 7:
                        fc += math.exp(-p.troe_t2[r] / T)
                    cc = -0.4 - 0.67 * math.log10(fc)
                    n = 0.75 - 1.27 * math.log10(fc)
                    f1 = (math.log10(pr) + cc) / (n - 0.14 * (math.log10(pr) + cc))
                    F = 10 ** (math.log10(fc) / (1 + f1 * f1))
                k = k * pr / (1 + pr) * F
            assert kf[r] == pytest.approx(k, rel=1e-12)
            # detailed balance with an independent Gibbs evaluation
            g = []
            for s in range(9):
                a = nasa(p, s, max(T, p.t_low[s]))
                tt = max(T, p.t_low[s])
                h_rt = a[0] + a[1] * tt / 2 + a[2] * tt ** 2 / 3 + a[3] * tt ** 3 / 4 + a[4] * tt ** 4 / 5 + a[5] / tt
                s_r = a[0] * math.log(tt) + a[1] * tt + a[2] * tt ** 2 / 2 + a[3] * tt ** 3 / 3 + a[4] * tt ** 4 / 4 + a[6]
                g.append(h_rt - s_r)
            dg = sum(g[s] * (nb[r, s] - nf[r, s]) for s in range(9))
            dn = int(nb[r].sum() - nf[r].sum())
            kc = (PATM / RU * 1e-3 / T) ** dn * math.exp(-dg)
            assert kb[r] == pytest.approx(k / kc, rel=1e-10)


def test_epi_jacobian_matches_frozen_rate_finite_differences(mech):
    """J_mn = d omega_m / d(rho Y_n) with kf, kb (and [M]) frozen."""
    p, _ = mech
    op = orc_param(p)
    mw = np.array(p.mw[:9])
    nf = np.array(p.stoi_f[:19 * 20]).reshape(19, 20)[:, :9].astype(float)
    nb = np.array(p.stoi_b[:19 * 20]).reshape(19, 20)[:, :9].astype(float)
    rho, t, y = random_states(5, 3)
    for i in range(5):
        om, _, jac, _, kf, kb = chem_point(op, rho[i], t[i], y[i])

        def omega(ry):
            c = ry / mw * 1e-3
            q = kf * np.prod(c ** nf, axis=1) - kb * np.prod(c ** nb, axis=1)
            return 1e3 * mw * ((nb - nf).T @ q)

        ry = rho[i] * y[i]
        assert np.allclose(omega(ry), om, rtol=1e-10, atol=1e-10 * np.abs(om).max())
        fd = np.zeros((9, 9))
        for n in range(9):
            h = 1e-6 * ry[n]
            e = np.zeros(9)
            e[n] = h
            fd[:, n] = (omega(ry + e) - omega(ry - e)) / (2 * h)
        scale = np.abs(fd).max()
        assert np.allclose(jac, fd, rtol=1e-5, atol=1e-7 * scale)


def test_lu_against_numpy():
    lib = O.load()
    rng = np.random.default_rng(5)
    for _ in range(20):
        a = rng.normal(size=(9, 9)) + np.diag(rng.uniform(5, 20, 9)) * rng.choice([-1, 1], 9)
        b = rng.normal(size=9)
        aa, bb = np.ascontiguousarray(a.copy()), np.ascontiguousarray(b.copy())
        lib.orc_solve_lu(O.dp(aa), O.dp(bb), 9)
        assert np.allclose(bb, np.linalg.solve(a, b), rtol=1e-11, atol=1e-12)


def test_newton_temperature_round_trip(mech):
    p, _ = mech
    op = orc_param(p)
    lib = O.load()
    rho, t, y = random_states(20, 6)
    h = np.zeros(20)
    for i in range(20):
        lib.orc_enthalpy_and_cp(t[i], O.dp(h), O.dp(np.zeros(20)), C.byref(op))
        R = RU * np.sum(y[i] / np.array(p.mw[:9]))
        e = np.sum(h[:9] * y[i]) - R * t[i]
        yy = np.ascontiguousarray(np.pad(y[i], (0, 11)))
        out = C.c_double()
        it = lib.orc_newton_temperature(C.byref(op), O.dp(yy), e, t[i] * 1.2, C.byref(out))
        assert 1 < it < 20
        assert out.value == pytest.approx(t[i], rel=1e-6)


def test_plain_chem_step_solves_the_point_implicit_system(mech):
    p, _ = mech
    op = orc_param(p)
    lib = O.load()
    rho, t, y = random_states(10, 7)
    mw = np.array(p.mw[:9])
    for i in range(10):
        pres = 101325.0 * (0.5 + i / 4)
        mwm = 1 / np.sum(y[i] / mw)
        r = pres * mwm / (RU * t[i])
        for dt in (1e-9, 1e-6):
            yy = np.ascontiguousarray(np.pad(y[i], (0, 11)))
            dy = np.zeros(20)
            lib.orc_plain_chem_step(C.byref(op), t[i], pres, O.dp(yy), dt, O.dp(dy))
            om, _, jac, *_ = chem_point(op, r, t[i], y[i])
            lhs = np.eye(9) - dt * jac
            assert np.allclose(lhs @ (dy[:9] * r), dt * om, rtol=1e-9, atol=1e-12 * dt * np.abs(om).max())
        # small-dt limit (dt*|J| = 1e-8): explicit Euler
        om, _, jac, *_ = chem_point(op, r, t[i], y[i])
        dt = 1e-8 / np.abs(jac).max()
        yy = np.ascontiguousarray(np.pad(y[i], (0, 11)))
        dy = np.zeros(20)
        lib.orc_plain_chem_step(C.byref(op), t[i], pres, O.dp(yy), dt, O.dp(dy))
        assert np.allclose(dy[:9], dt * om / r, rtol=1e-6, atol=1e-9 * dt * np.abs(om).max() / r)

"""Planted mistakes in the oracle, on the CPU: each mutant is a copy of the
oracle's C sources with one plausible mistake (a dropped term, a wrong sign
or index, a transposed operand, a mixed-up constant), built with the
oracle's own flags into a temporary directory.  The pins of
tests/test_oracle_pins.py must fail on every mutant; the unmutated oracle
passes them (that file's own tests).  A mutation whose text is no longer in
the source fails here, so the list follows the oracle as it changes.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess

import pytest

import oracle_ffi as O
import test_oracle_pins as P

SRC = ["orc_geometry.c", "orc_thermo.c", "orc_flow.c", "orc_chem.c", "orc_sur.c"]
HDR = ["ostc_oracle.h", "orc_internal.h"]
CFLAGS = ["-O1", "-std=gnu11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared", "-Wl,-Bsymbolic"]

# name: (file, text, replacement, pin that must fail)
MUTANTS = {
    "jacobian_term_sign": ("orc_geometry.c", "dxd2 * dyd3 * dzd1 + dxd3 * dyd1 * dzd2",
                           "dxd2 * dyd3 * dzd1 - dxd3 * dyd1 * dzd2", P.check_metrics_affine),
    "metric_cross_term": ("orc_geometry.c", "metric[3 * plane + id] = (dyd3 * dzd1 - dyd1 * dzd3) / J;",
                          "metric[3 * plane + id] = (dyd1 * dzd3 - dyd3 * dzd1) / J;", P.check_metrics_warped),
    "ghost_corner_rule": ("orc_geometry.c", "X(i1, j1, k) = X(i1, 0, k) + X(0, j1, k) - X(0, 0, k);",
                          "X(i1, j1, k) = X(i1, 0, k) + X(0, j1, k) - X(0, 1, k);", P.check_metrics_affine),
    "inner_no_bv_update": ("orc_flow.c", "          update_bv_1_point(z, p, idx[0], idx[1], idx[2]);\n",
                           "\n", P.check_inner_face_average_and_ghosts),
    "inner_copy_not_average": ("orc_flow.c", "const double ave = 0.5 * (*b + *a);", "const double ave = *a;",
                               P.check_inner_face_average_and_ghosts),
    "inner_ghost_without_t": ("orc_flow.c", "for (int l = 0; l < 6; ++l)\n            z->bv[IX(z,",
                              "for (int l = 0; l < 5; ++l)\n            z->bv[IX(z,",
                              P.check_inner_face_average_and_ghosts),
    "inner_target_offset": ("orc_flow.c",
                            "it[q] = f->target_start[q] + f->target_loop_dir[q] * d[f->src_tar[q]];",
                            "it[q] = f->target_start[q] + f->target_loop_dir[q] * d[f->src_tar[q]] + (q == 1);",
                            P.check_inner_split_equals_single),
    "residual_u_for_speed": ("orc_flow.c", "d = S(vel, i, j, k) - F(bv_last, i, j, k, 1);",
                             "d = F(bv, i, j, k, 1) - F(bv_last, i, j, k, 1);", P.check_residual_sums),
    "residual_store_u": ("orc_flow.c", "F(bv_last, i, j, k, 1) = S(vel, i, j, k);",
                         "F(bv_last, i, j, k, 1) = F(bv, i, j, k, 1);", P.check_residual_sums),
    "residual_sums_ghosts": ("orc_flow.c", "for (int i = 0; i < z->mx; ++i) s += F(bv_last, i, j, k, l);",
                             "for (int i = -1; i < z->mx; ++i) s += F(bv_last, i, j, k, l);", P.check_residual_sums),
    "epi_transposed": ("orc_chem.c", "lhs[m * ns + n] = (m == n ? d : 0.0) - dt * F(chem_src_jac, i, j, k, m * ns + n);",
                       "lhs[m * ns + n] = (m == n ? d : 0.0) - dt * F(chem_src_jac, i, j, k, n * ns + m);",
                       lambda: P.check_explicit_point_implicit(1, False)),
    "epi_dqk_without_diag": ("orc_chem.c", "lhs[m * ns + n] = m == n ? diag - dt * F(chem_src_jac",
                             "lhs[m * ns + n] = m == n ? 1.0 - dt * F(chem_src_jac",
                             lambda: P.check_dplur_chemistry_block(1)),
    "da_dq0_sign": ("orc_flow.c", "for (int l = 0; l < ns; ++l) F(dq, i, j, k, 5 + l) /= diag - dt * F(chem_src_jac",
                    "for (int l = 0; l < ns; ++l) F(dq, i, j, k, 5 + l) /= diag + dt * F(chem_src_jac",
                    lambda: P.check_dplur_chemistry_block(2)),
    "da_explicit_sign": ("orc_flow.c", "F(dq, i, j, k, 5 + l) /= 1 - dt * F(chem_src_jac, i, j, k, l);",
                         "F(dq, i, j, k, 5 + l) /= 1 + dt * F(chem_src_jac, i, j, k, l);",
                         lambda: P.check_explicit_point_implicit(2, False)),
    "sst_explicit_row": ("orc_flow.c", "F(dq, i, j, k, ns + 6) /= 1 - dt * F(turb_src_jac, i, j, k, 1);",
                         "F(dq, i, j, k, ns + 6) /= 1 - dt * F(turb_src_jac, i, j, k, 0);",
                         lambda: P.check_explicit_point_implicit(1, True)),
    "sigma_k_blend_reversed": ("orc_flow.c", "const double sigma_k = SST_SIGMA_K2 + SST_DELTA_SIGMA_K * f1;",
                               "const double sigma_k = SST_SIGMA_K2 + SST_DELTA_SIGMA_K * (1 - f1);",
                               lambda: P.check_viscous_sst(False)),
    "sigma_omega_outer": ("orc_flow.c", "const double sigma_omega = SST_SIGMA_OMEGA2 + SST_DELTA_SIGMA_OMEGA * f1;",
                          "const double sigma_omega = SST_SIGMA_OMEGA1 + SST_DELTA_SIGMA_OMEGA * f1;",
                          lambda: P.check_viscous_sst(True)),
    "normal_stress_no_average": ("orc_flow.c", "const double tt = -2.0 / 3 * 0.5 * (",
                                 "const double tt = -2.0 / 3 * (", lambda: P.check_viscous_sst(True)),
    "species_correction_sign": ("orc_flow.c",
                                "fv[5 + l] = diffusivity[l] * (kx * y_x[l] + ky * y_y[l] + kz * y_z[l]) -",
                                "fv[5 + l] = diffusivity[l] * (kx * y_x[l] + ky * y_y[l] + kz * y_z[l]) +",
                                lambda: P.check_species_diffusion("linear", False)),
    "enthalpy_flux_species": ("orc_flow.c", "const double DMulH = diffusivity[l] * h[l];",
                              "const double DMulH = diffusivity[l] * h[0];",
                              lambda: P.check_species_diffusion("quadratic", False)),
    "turbulent_schmidt": ("orc_flow.c", "turb_diffusivity = mut / p->Sct;", "turb_diffusivity = mut / p->Prt;",
                          lambda: P.check_species_diffusion("quadratic", True)),
}


def _build(root, name, spec):
    fname, old, new, _ = spec
    d = os.path.join(root, name)
    os.makedirs(d, exist_ok=True)
    for f in SRC + HDR:
        shutil.copy(os.path.join(O.HERE, f), d)
    if fname is not None:
        path = os.path.join(d, fname)
        text = open(path).read()
        n = text.count(old)
        if n != 1:
            return name, None, f"mutation text found {n} times in {fname}"
        open(path, "w").write(text.replace(old, new))
    out = os.path.join(d, "liboracle_mutant.so")
    r = subprocess.run(["gcc", *CFLAGS, "-o", out, *SRC, "-lm"], cwd=d, capture_output=True, text=True)
    if r.returncode != 0:
        return name, None, r.stderr[-2000:]
    return name, out, ""


@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    root = str(tmp_path_factory.mktemp("oracle_mutants"))
    jobs = list(MUTANTS.items()) + [("control", (None, None, None, None))]
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        built = list(ex.map(lambda kv: _build(root, *kv), jobs))
    return {name: (path, msg) for name, path, msg in built}


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_pin_fails_on_planted_mistake(mutants, name):
    path, msg = mutants[name]
    assert path, f"{name}: {msg}"
    check = MUTANTS[name][3]
    with O.use_library(path):
        with pytest.raises(AssertionError):
            check()


def test_unmutated_copy_passes_every_pin(mutants):
    """The same build recipe without a mutation passes all the pins used
    above, so a failure on a mutant comes from its planted mistake."""
    path, msg = mutants["control"]
    assert path, msg
    checks = {id(spec[3]): spec[3] for spec in MUTANTS.values()}
    with O.use_library(path):
        for check in checks.values():
            check()

"""Mechanisms other than the shipped 9-species one: a 7-species / 8-reaction
subset of the H2-air mechanism (H2 H O2 O OH H2O N2 and the reactions among
them, without HO2 and H2O2), read from CHEMKIN files written here from the
shipped ones.  The library runs it on its 9-species kernels with two inert
species appended (zero mass fraction, no reactions); the oracle runs the
7-species mechanism as given (its loops take any species count), so these
tests compare the padded GPU path with an unpadded restatement.

Fields are exchanged in the caller's layout: 7 species planes in SV, 12
conserved variables for laminar flow (14 with SST), 49 Jacobian planes."""
import ctypes as C
import os

import numpy as np
import pytest

from harness import Pair, interior, scaled_err
from openstc_amd import _lib as L
from openstc_amd import cases

pytestmark = pytest.mark.gpu

SPECIES = ["H2", "H", "O2", "O", "OH", "H2O", "N2"]


def say(*a):
    print("[gpu-mech]", *a, flush=True)


def write_subset(tmp, species=SPECIES, name="chemkin7"):
    """The shipped mechanism cut to `species` (in that order) and its first
    eight reactions."""
    src = open(os.path.join(cases.DATA, "combustion_mechanism.inp")).read().splitlines()
    out, in_species, in_reactions = [], False, False
    for line in src:
        u = line.strip().upper()
        if u.startswith("SPECIES"):
            in_species = True
            out.append(line)
            out.append(" " + "  ".join(species))
            continue
        if in_species:
            if u.startswith("END"):
                in_species = False
                out.append(line)
            continue
        if u.startswith("REACTIONS"):
            in_reactions = True
        elif in_reactions and u.startswith("H+O2(+M)"):
            out.append("END")
            break
        out.append(line)
    d = os.path.join(tmp, name)
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "combustion_mechanism.inp"), "w") as f:
        f.write("\n".join(out) + "\n")
    for name in ("therm.dat", "tran.dat"):
        with open(os.path.join(cases.DATA, name)) as a, open(os.path.join(d, name), "w") as b:
            b.write(a.read())
    return d


@pytest.fixture(scope="module")
def mech_dir(tmp_path_factory):
    return write_subset(str(tmp_path_factory.mktemp("mech")))


# the same subset with N2, which takes part in no reaction, second in the
# species list: the EPI block of the species that react has a gap
SPECIES_N2_INSIDE = ["H2", "N2", "H", "O2", "O", "OH", "H2O"]


@pytest.fixture(scope="module")
def mech_dir_n2_inside(tmp_path_factory):
    return write_subset(str(tmp_path_factory.mktemp("mech_n2")), SPECIES_N2_INSIDE, "chemkin7_n2")


def compare(pr: Pair, b, tol_scaled, tol_y, label):
    ns = pr.case.param.n_spec
    errs = {}
    for f in ("BV", "SV", "CV"):
        a = interior(pr.zg(b).download(f))
        o = interior(pr.oarr(f, b))
        if f == "SV":
            errs["Y(abs)"] = float(np.abs(a[:ns] - o[:ns]).max())
            if a.shape[0] > ns:
                errs["SV:turb"] = scaled_err(a[ns:], o[ns:])
        elif f == "CV":
            # flow and turbulence planes scaled by their own maxima; the
            # partial densities by the largest density (trace radicals sit
            # at round-off levels, where a plane's own maximum scales nothing)
            errs["CV"] = max(scaled_err(a[:5], o[:5]), scaled_err(a[5 + ns:], o[5 + ns:]) if a.shape[0] > 5 + ns
                             else 0.0)
            errs["CV:rhoY/rho_max"] = float(np.abs(a[5:5 + ns] - o[5:5 + ns]).max() / np.abs(o[0]).max())
        else:
            errs[f] = scaled_err(a, o)
    say(label, {k: f"{v:.2e}" for k, v in errs.items()})
    bad = {k: v for k, v in errs.items() if not v <= (tol_y if k in ("Y(abs)", "CV:rhoY/rho_max") else tol_scaled)}
    assert not bad, (label, bad)


def test_subset_mechanism_reads(mech_dir):
    mech, info = cases.load_mechanism(True, mech_dir)
    say("species", mech.n_spec, "reactions", mech.n_reac)
    assert mech.n_spec == 7 and mech.n_reac == 8
    assert cases.species_names(info) == SPECIES


def test_mixing_layer_7_species_fp64(mech_dir):
    case = cases.mixing_layer(mx=24, my=16, mz=8, cfl=0.2, mech_dir=mech_dir)
    assert case.param.n_spec == 7
    say("mixing_layer 24x16x8, 7 species / 8 reactions, FP64, source stage + 5 iterations")
    pr = Pair(case, precision=L.FP64)
    pr.init()
    z = pr.zg()
    assert z.planes("SV") == 7 and z.planes("CV") == 12 and z.planes("CHEM_JAC") == 49
    assert z.planes("RHO_D") == 7
    # the source stage alone: omega and the 7x7 EPI Jacobian
    L.check(pr.lib.ostc_compute_source(pr.c, z.h, None), "source")
    pr.orc.lib.orc_compute_source(pr.zo(), pr.p)
    ej = pr.err("CHEM_JAC")
    edq = pr.err("DQ")
    say("after the source: CHEM_JAC", ej, "DQ", edq)
    assert ej < 1e-10 and edq < 1e-10
    pr.close()
    pr = Pair(case, precision=L.FP64)
    pr.init()
    for _ in range(5):
        pr.gpu.step()
        pr.orc.step()
    compare(pr, 0, 1e-9, 1e-12, "7 species FP64 after 5 iterations")
    pr.close()


def test_mixing_layer_7_species_fp32(mech_dir):
    case = cases.mixing_layer(mx=24, my=16, mz=8, cfl=0.2, mech_dir=mech_dir)
    say("mixing_layer 24x16x8, 7 species, FP32 bulk vs FP64 oracle, 10 iterations")
    pr = Pair(case, precision=L.FP32)
    pr.init()
    for _ in range(10):
        pr.gpu.step()
        pr.orc.step()
    compare(pr, 0, 2e-5, 2e-6, "7 species FP32 after 10 iterations")
    pr.close()


def test_jet_crossflow_7_species_sst_fp64(mech_dir):
    case = cases.jet_crossflow(mx=16, my=12, mz=8, nblocks=2, steps=5, precision=L.FP64, cfl=0.2,
                               mech_dir=mech_dir)
    assert case.param.n_spec == 7
    say("jet_crossflow 2 x 16x12x8, 7 species + SST, FP64, 5 iterations")
    pr = Pair(case, precision=L.FP64)
    pr.init()
    assert pr.zg().planes("SV") == 9 and pr.zg().planes("CV") == 14
    for _ in range(5):
        pr.gpu.step()
        pr.orc.step()
    for b in range(2):
        compare(pr, b, 1e-9, 1e-12, f"jet block {b} after 5 iterations")
    pr.close()


def test_mixing_layer_7_species_n2_inside_fp64(mech_dir_n2_inside):
    """N2 second of seven species: the DPLUR chemistry block is factored and
    solved over the six species that react (species 0, 2..6 of the caller,
    with the two appended inert species also outside it), and N2's rows are
    divided by the diagonal; the oracle factors the whole 7x7 block."""
    mech, info = cases.load_mechanism(True, mech_dir_n2_inside)
    assert cases.species_names(info) == SPECIES_N2_INSIDE
    case = cases.mixing_layer(mx=24, my=16, mz=8, cfl=0.2, mech_dir=mech_dir_n2_inside)
    say("mixing_layer 24x16x8, 7 species with N2 second, FP64, 5 iterations")
    pr = Pair(case, precision=L.FP64)
    pr.init()
    for _ in range(5):
        pr.gpu.step()
        pr.orc.step()
    compare(pr, 0, 1e-9, 1e-12, "7 species (N2 second) FP64 after 5 iterations")
    pr.close()


def write_n2_collider(tmp):
    """The shipped 9-species mechanism with a 20th reaction in which N2 is a
    specific collision partner, N2+H2=N2+H+H (the H2 dissociation rate): N2
    takes part in a reaction with no net change, so all nine species enter
    the EPI block -- N2 with a zero row and a nonzero column -- and the
    block's pivots take two packed words."""
    d = os.path.join(tmp, "chemkin9_n2")
    os.makedirs(d, exist_ok=True)
    src = open(os.path.join(cases.DATA, "combustion_mechanism.inp")).read().splitlines()
    end = max(i for i, line in enumerate(src) if line.strip().upper() == "END")
    out = src[:end] + ["N2+H2=N2+H+H    4.58E+19   -1.40    1.04E+05"] + src[end:]
    with open(os.path.join(d, "combustion_mechanism.inp"), "w") as f:
        f.write("\n".join(out) + "\n")
    for name in ("therm.dat", "tran.dat"):
        with open(os.path.join(cases.DATA, name)) as a, open(os.path.join(d, name), "w") as b:
            b.write(a.read())
    return d


@pytest.fixture(scope="module")
def mech_dir_n2_collider(tmp_path_factory):
    return write_n2_collider(str(tmp_path_factory.mktemp("mech_n2c")))


def test_mixing_layer_all_nine_react_fp64(mech_dir_n2_collider):
    mech, info = cases.load_mechanism(True, mech_dir_n2_collider)
    assert mech.n_spec == 9 and mech.n_reac == 20
    case = cases.mixing_layer(mx=24, my=16, mz=8, cfl=0.2, mech_dir=mech_dir_n2_collider)
    say("mixing_layer 24x16x8, 9 species / 20 reactions (N2 a collider), FP64, 5 iterations")
    pr = Pair(case, precision=L.FP64)
    pr.init()
    for _ in range(5):
        pr.gpu.step()
        pr.orc.step()
    compare(pr, 0, 1e-9, 1e-12, "9 species (N2 a collider) FP64 after 5 iterations")
    pr.close()


def test_mixing_layer_all_nine_react_fp32(mech_dir_n2_collider):
    case = cases.mixing_layer(mx=24, my=16, mz=8, cfl=0.2, mech_dir=mech_dir_n2_collider)
    say("mixing_layer 24x16x8, 9 species / 20 reactions (N2 a collider), FP32 bulk vs FP64 oracle, 10 iterations")
    pr = Pair(case, precision=L.FP32)
    pr.init()
    for _ in range(10):
        pr.gpu.step()
        pr.orc.step()
    compare(pr, 0, 2e-5, 2e-6, "9 species (N2 a collider) FP32 after 10 iterations")
    pr.close()

"""Mechanisms of more than nine species run on the 20-species kernels: the
shipped H2-air mechanism with argon and helium added in front of its species
(11 species, the nine reacting ones at positions 2..10; AR and HE take part
in no reaction but count as third bodies with the default efficiency), read
from CHEMKIN files written here.  The library appends nine inert species and
factors the whole 20 x 20 chemistry block with run-time loops; the oracle
runs the 11-species mechanism as given.  The NASA-7 fits of AR and HE are
the published constant-cp ones (cp/R = 5/2)."""
import os

import numpy as np
import pytest

from harness import Pair
from openstc_amd import _lib as L
from openstc_amd import cases
from test_gpu_mechanisms import compare

pytestmark = pytest.mark.gpu

SPECIES = ["AR", "HE", "H2", "H", "O2", "O", "OH", "HO2", "H2O2", "H2O", "N2"]

THERM_AR_HE = """\
AR                120186AR  1               G   300.00   5000.00  1000.00      1
 0.02500000E+02 0.00000000E+00 0.00000000E+00 0.00000000E+00 0.00000000E+00    2
-0.07453750E+04 0.04366000E+02 0.02500000E+02 0.00000000E+00 0.00000000E+00    3
 0.00000000E+00 0.00000000E+00-0.07453750E+04 0.04366000E+02                   4
HE                120186HE  1               G   300.00   5000.00  1000.00      1
 0.02500000E+02 0.00000000E+00 0.00000000E+00 0.00000000E+00 0.00000000E+00    2
-0.07453750E+04 0.09153488E+01 0.02500000E+02 0.00000000E+00 0.00000000E+00    3
 0.00000000E+00 0.00000000E+00-0.07453750E+04 0.09153489E+01                   4
"""


def say(*a):
    print("[gpu-mech20]", *a, flush=True)


def write_extended(tmp):
    src = open(os.path.join(cases.DATA, "combustion_mechanism.inp")).read().splitlines()
    out, section = [], None
    for line in src:
        u = line.strip().upper()
        if u.startswith("ELEMENTS"):
            section = "e"
            out.append(line)
            continue
        if u.startswith("SPECIES"):
            section = "s"
            out.append(line)
            out.append(" " + "  ".join(SPECIES))
            continue
        if section in ("e", "s") and u.startswith("END"):
            if section == "e":
                out.append(" H  O  N  AR  HE")
            section = None
            out.append(line)
            continue
        if section in ("e", "s"):
            continue
        out.append(line)
    d = os.path.join(tmp, "chemkin11")
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "combustion_mechanism.inp"), "w") as f:
        f.write("\n".join(out) + "\n")
    therm = open(os.path.join(cases.DATA, "therm.dat")).read().rstrip().splitlines()
    assert therm[-1].strip().upper() == "END"
    with open(os.path.join(d, "therm.dat"), "w") as f:
        f.write("\n".join(therm[:-1]) + "\n" + THERM_AR_HE + "END\n")
    with open(os.path.join(cases.DATA, "tran.dat")) as a, open(os.path.join(d, "tran.dat"), "w") as b:
        b.write(a.read())
    return d


@pytest.fixture(scope="module")
def mech_dir(tmp_path_factory):
    return write_extended(str(tmp_path_factory.mktemp("mech11")))


def test_extended_mechanism_reads(mech_dir):
    mech, info = cases.load_mechanism(True, mech_dir)
    say("species", mech.n_spec, "reactions", mech.n_reac)
    assert mech.n_spec == 11 and mech.n_reac == 19
    assert cases.species_names(info) == SPECIES
    assert abs(mech.mw[0] - 39.95) < 1e-6 and abs(mech.mw[1] - 4.002602) < 1e-6


def test_mixing_layer_11_species_fp64(mech_dir):
    case = cases.mixing_layer(mx=24, my=16, mz=8, cfl=0.2, mech_dir=mech_dir)
    assert case.param.n_spec == 11
    say("mixing_layer 24x16x8, 11 species / 19 reactions, FP64, source stage + 5 iterations")
    pr = Pair(case, precision=L.FP64)
    pr.init()
    z = pr.zg()
    assert z.planes("SV") == 11 and z.planes("CHEM_JAC") == 121
    L.check(pr.lib.ostc_compute_source(pr.c, z.h, None), "source")
    pr.orc.lib.orc_compute_source(pr.zo(), pr.p)
    ej, edq = pr.err("CHEM_JAC"), pr.err("DQ")
    say("after the source: CHEM_JAC", ej, "DQ", edq)
    assert ej < 1e-10 and edq < 1e-10
    pr.close()
    pr = Pair(case, precision=L.FP64)
    pr.init()
    for _ in range(5):
        pr.gpu.step()
        pr.orc.step()
    compare(pr, 0, 1e-9, 1e-12, "11 species FP64 after 5 iterations")
    pr.close()


def test_mixing_layer_11_species_fp32(mech_dir):
    case = cases.mixing_layer(mx=24, my=16, mz=8, cfl=0.2, mech_dir=mech_dir)
    say("mixing_layer 24x16x8, 11 species, FP32 bulk vs FP64 oracle, 10 iterations")
    pr = Pair(case, precision=L.FP32)
    pr.init()
    for _ in range(10):
        pr.gpu.step()
        pr.orc.step()
    compare(pr, 0, 2e-5, 2e-6, "11 species FP32 after 10 iterations")
    pr.close()


def test_jet_crossflow_11_species_sst_fp64(mech_dir):
    case = cases.jet_crossflow(mx=16, my=12, mz=8, nblocks=2, steps=5, precision=L.FP64, cfl=0.2,
                               mech_dir=mech_dir)
    assert case.param.n_spec == 11
    say("jet_crossflow 2 x 16x12x8, 11 species + SST, FP64, 5 iterations")
    pr = Pair(case, precision=L.FP64)
    pr.init()
    assert pr.zg().planes("SV") == 13 and pr.zg().planes("CV") == 18
    for _ in range(5):
        pr.gpu.step()
        pr.orc.step()
    for b in range(2):
        compare(pr, b, 1e-9, 1e-12, f"11 species jet block {b} after 5 iterations")
    pr.close()

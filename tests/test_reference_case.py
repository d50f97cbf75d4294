"""The reference's own sample case: its setup files (tests/golden/
reference_setup, from the reference's input_files/setup and
boundary_condition) with its grid (tests/golden/reference_case_grid.npz,
written from its input_files/grid by tests/golden/make_reference_case.py):
two 2-D blocks, 100x100x1 (inlet channel) and 120x158x1 (chamber), joined
by an inner face; 9-species / 19-reaction H2-air finite rate with EPI, k-omega
SST, NND-2 + minmod + AUSM+, central viscous flux, DPLUR x3, CFL 0.5.

2-D blocks (mz = 1) take the reference's 2-D branches: the k ghost layers of
the grid repeat x, y and step z by one per layer (src/Mesh.cpp:840-856), the
metric's interior pass covers k = 0 (src/Mesh.cpp:263-268), boundary faces
span the k ghost layers (src/Mesh.cpp:26-29) and inner faces register in
2-D (src/Mesh.cpp:66-79), the update zeroes the z momentum
(src/FieldOperation.cuh:153-155), and eliminate_k_gradient copies the k = 0
plane of bv/sv into the k ghost layers before the first boundary conditions
and after every exchange (src/FieldOperation.cu:42-62, src/Driver.cu:59-65,
259-265).

CPU: the case reads (2-D mesh, registered inner faces and boundary ranges as
the reference's registration produces them); the product's host geometry is
bit-exact with the oracle's restatement, and the 2-D ghost layers hold what
the rule fixes.  GPU: the C++ driver (FP64) against the FP64 oracle over 5
iterations, and the FP32-bulk driver against it.
"""
import os
import shutil

import numpy as np
import pytest

import oracle_ffi as O
from harness import interior, scaled_err
from openstc_amd import _lib as L
from openstc_amd import case_files, cases

HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLE = os.path.join(HERE, "golden", "reference_setup")
GRID = os.path.join(HERE, "golden", "reference_case_grid.npz")


def reference_case_dir(root):
    d = os.path.join(str(root), "reference_case")
    inp = os.path.join(d, "input_files")
    for sub in ("setup", "grid", "boundary_condition", "chemistry"):
        os.makedirs(os.path.join(inp, sub), exist_ok=True)
    for fn in os.listdir(SAMPLE):
        if fn.endswith(".txt") and fn[0].isdigit():
            shutil.copyfile(os.path.join(SAMPLE, fn), os.path.join(inp, "setup", fn))
    shutil.copyfile(os.path.join(SAMPLE, "boundary0.txt"), os.path.join(inp, "boundary_condition", "boundary   0.txt"))
    shutil.copyfile(os.path.join(SAMPLE, "inner0.txt"), os.path.join(inp, "boundary_condition", "inner   0.txt"))
    for fn in ("combustion_mechanism.inp", "therm.dat", "tran.dat"):
        shutil.copyfile(os.path.join(cases.DATA, fn), os.path.join(inp, "chemistry", fn))
    g = np.load(GRID)
    blocks = []
    for b, (mx, my, mz) in enumerate(g["dims"]):
        blocks.append(cases.Block(int(mx), int(my), int(mz), g[f"x{b}"], g[f"y{b}"], g[f"z{b}"]))
    case_files.write_grid(os.path.join(inp, "grid", "grid   0.grd"), blocks)
    return d


def case_from_dir(cd, precision=L.FP64):
    """The case directory as a Case for the oracle: interior coordinates,
    initial fields, boundary programs and inner faces as the C++ reader
    produced them."""
    g = cd.ngg
    p = cd.param()
    blocks = []
    for b in range(cd.info.n_block):
        mx, my, mz = cd.info.mx[b], cd.info.my[b], cd.info.mz[b]
        x, y, z = (interior(cd.field(b, f, 1)[0], g) for f in ("X", "Y", "Z"))
        blk = cases.Block(mx, my, mz, np.ascontiguousarray(x), np.ascontiguousarray(y), np.ascontiguousarray(z))
        blk.bv0 = cd.field(b, "BV", 6)
        blk.sv0 = cd.field(b, "SV", p.n_scalar)
        blk.bc = cd.bc(b)
        blk.inner = cd.inner(b)
        blocks.append(blk)
    return cases.Case("reference_sample_case", p, blocks, g, precision, cd.info.total_step, cd.info.output_screen,
                      cd.wall_points())


def test_reference_case_reads_as_2d(tmp_path):
    cd = case_files.CaseDir(reference_case_dir(tmp_path))
    info = cd.info
    assert info.n_block == 2
    assert [(info.mx[b], info.my[b], info.mz[b]) for b in range(2)] == [(100, 100, 1), (120, 158, 1)]
    p = cd.param()
    assert (p.mix_model, p.turb_method, p.n_spec, p.n_var, p.implicit_method, p.dplur_inner_step) == \
        (L.FR, L.RANS, 9, 16, 1, 3)
    # boundary faces span the k ghost layers; walls are trimmed to k = 0
    bc0 = cd.bc(0)
    for q in range(bc0.n_face):
        f = bc0.face[q]
        if f.kind == 2:
            assert (f.range_start[2], f.range_end[2]) == (0, 0)
        else:
            assert (f.range_start[2], f.range_end[2]) == (-2, 2)
    # inner faces, registered in 2-D (src/Mesh.cpp:44-79, 110-177): block 0's
    # i = 100 face onto block 1's i = 1, j 59..158 (1-based records)
    f0 = cd.inner(0)
    assert len(f0) == 1
    a = f0[0]
    assert (a.face, a.direction, a.target_block, a.target_face, a.target_direction) == (0, 1, 1, 0, -1)
    assert list(a.range_start) == [99, 0, 0] and list(a.range_end) == [101, 99, 0]
    assert list(a.target_start) == [0, 58, 0] and list(a.target_end) == [0, 157, 0]
    assert list(a.src_tar) == [0, 1, 2] and list(a.loop_dir) == [1, 1, 1] and list(a.target_loop_dir) == [1, 1, 1]
    assert list(a.n_point) == [3, 100, 1]
    b1 = cd.inner(1)[0]
    assert (b1.face, b1.direction, b1.target_block, b1.target_direction) == (0, -1, 0, 1)
    assert list(b1.range_start) == [-2, 58, 0] and list(b1.range_end) == [0, 157, 0]
    assert list(b1.target_start) == [97, 0, 0] and list(b1.n_point) == [3, 100, 1]
    # ghost grid of a 2-D block: the k layers repeat x, y and step z by one
    # unit per layer (after the grid scale, as the reference orders it)
    # (interior i, j: the edge points take the parallelogram rule's round-off)
    g = cd.ngg
    x = cd.field(0, "X", 1)[0][:, g:-g, g:-g]
    z = cd.field(0, "Z", 1)[0][:, g:-g, g:-g]
    for k in (1, 2):
        assert np.array_equal(x[g + k], x[g]) and np.array_equal(x[g - k], x[g])
        assert np.array_equal(z[g + k], z[g] + k) and np.array_equal(z[g - k], z[g] - k)
    cd.close()


def test_reference_case_geometry_matches_oracle(tmp_path):
    cd = case_files.CaseDir(reference_case_dir(tmp_path))
    case = case_from_dir(cd)
    orc = O.OracleSimulation(case)
    for b in range(2):
        ob = orc.blocks[b]
        for name, fld in (("x", "X"), ("y", "Y"), ("z", "Z"), ("jac", "JAC")):
            mine = cd.field(b, fld, 1)[0]
            assert np.array_equal(mine, ob.arr[name][0]), (b, name)
        assert np.array_equal(cd.field(b, "METRIC", 9), ob.arr["metric"]), b
        # one k plane with z ghost layers one unit apart: zeta_z = 1 and the
        # other z entries vanish (to the round-off of the corner ghost points'
        # parallelogram rule)
        m = ob.arr["metric"]
        scale = np.abs(m[:2]).max()
        assert np.abs(m[8] - 1.0).max() < 1e-12, b
        for q in (2, 5, 6, 7):
            assert np.abs(m[q]).max() < 1e-12 * scale, (b, q)
    cd.close()


def _driver_vs_oracle(tmp_path, precision, steps, tol_bv, tol_y):
    cd = case_files.CaseDir(reference_case_dir(tmp_path))
    case = case_from_dir(cd, precision)
    drv = case_files.Driver(cd, precision)
    drv.initialize()
    orc = O.OracleSimulation(case)
    orc.initialize_computation()
    ns = case.param.n_spec
    for s in range(steps):
        drv.step()
        orc.step(output_screen=case.output_screen)
    worst = {}
    for b in range(2):
        g = case.ngg
        bv = interior(drv.download(b, "BV", 6), g)
        sv = interior(drv.download(b, "SV", ns + 2), g)
        obv = interior(orc.blocks[b].arr["bv"], g)
        osv = interior(orc.blocks[b].arr["sv"][:ns + 2], g)
        e_bv = scaled_err(bv, obv)
        e_y = float(np.abs(sv[:ns] - osv[:ns]).max())
        e_t = scaled_err(sv[ns:], osv[ns:])
        print(f"[gpu-ref-case] block {b} precision {precision}: BV {e_bv:.2e}, max |dY| {e_y:.2e}, k/omega {e_t:.2e}",
              flush=True)
        worst[b] = (e_bv, e_y, e_t)
        assert e_bv < tol_bv and e_y < tol_y and e_t < tol_bv, (b, worst[b])
    drv.close()
    cd.close()


@pytest.mark.gpu
def test_reference_case_driver_fp64_vs_oracle(tmp_path):
    print("[gpu-ref-case] reference sample case (2-D, 2 blocks) FP64, 5 iterations", flush=True)
    _driver_vs_oracle(tmp_path, L.FP64, 5, 1e-9, 1e-10)


@pytest.mark.gpu
def test_reference_case_driver_fp32_vs_oracle(tmp_path):
    print("[gpu-ref-case] reference sample case (2-D, 2 blocks) FP32 bulk, 5 iterations", flush=True)
    _driver_vs_oracle(tmp_path, L.FP32, 5, 1e-3, 1e-4)

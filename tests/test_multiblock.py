"""Multi-block path: inner-face registration (src/Mesh.cpp:44-178), inner
ghost grids (src/Mesh.cpp:940-985) and the inner-face halo exchange
(src/DataCommunication.cuh:27-98).

CPU: registration against hand-derived records; the product's multi-block
geometry bit-exact against the oracle's two-phase restatement; ghost
coordinates across the inner face equal the neighbour block's points.
GPU: the two-block split of the warped Euler box against the oracle
(initialisation incl. the exchanged ghost layers, then 10 explicit steps),
FP64 scaled error <= 1e-11; the four-block SST + finite-rate jet case
(the jet-in-crossflow case at 16x12x10 per block) after 3 DPLUR iterations: BV and
k/omega scaled error <= 1e-11, species mass fractions within 1e-10
absolute."""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi as O
from harness import Pair, interior
from openstc_amd import _lib as L
from openstc_amd import cases
from openstc_amd.runtime import register_inner_face


def fields(f):
    return (list(f.range_start), list(f.range_end), f.face, f.direction, list(f.target_start), list(f.target_end),
            f.target_block, f.target_face, f.target_direction, list(f.src_tar), list(f.loop_dir),
            list(f.target_loop_dir), list(f.n_point))


def test_register_inner_face_known_answers():
    # block 1 (33 x 64 x 8) face i = max joined to block 2's i = 0, and back
    a = register_inner_face([33, 33, -1, -64, 1, 8, 1, 1, -1, -64, 1, 8, 2], 2)
    assert fields(a) == ([32, 0, 0], [34, 63, 7], 0, 1, [0, 0, 0], [0, 63, 7], 1, 0, -1, [0, 1, 2], [1, 1, 1],
                         [1, 1, 1], [3, 64, 8])
    b = register_inner_face([1, 1, -1, -64, 1, 8, 33, 33, -1, -64, 1, 8, 1], 2)
    assert fields(b) == ([-2, 0, 0], [0, 63, 7], 0, -1, [30, 0, 0], [32, 63, 7], 0, 0, 1, [0, 1, 2], [1, 1, 1],
                         [1, 1, 1], [3, 64, 8])
    # a j-face whose tangential directions run opposite on the target side
    c = register_inner_face([1, 16, 8, 8, -1, -4, 16, 1, 1, 1, -1, -4, 3], 2)
    assert c.face == 1 and c.direction == 1 and c.target_face == 1 and c.target_direction == -1
    assert c.target_block == 2
    # i on this side maps to the target's i with reversed order
    assert list(c.target_loop_dir) == [-1, 1, 1] and list(c.src_tar) == [0, 1, 2]
    # malformed record (no face direction) is rejected
    r = (C.c_int * 13)(*[1, 2, 1, 2, 1, 2, 1, 2, 1, 2, 1, 2, 1])
    assert L.load().ostc_register_inner_face(r, 2, 3, C.byref(L.InnerFace())) != 0


def test_mesh_geometry_matches_oracle_and_neighbours():
    case = cases.warped_box_split(mx=24, my=16, mz=8)
    orc = O.OracleSimulation(case)
    g = case.ngg
    for b, blk in enumerate(case.blocks):
        ob = orc.blocks[b]
        for name, mine in (("x", blk.xg), ("y", blk.yg), ("z", blk.zg), ("jac", blk.jac)):
            assert np.array_equal(mine, ob.arr[name][0] if ob.arr[name].ndim == 4 else ob.arr[name]), (b, name)
        assert np.array_equal(blk.metric, ob.arr["metric"]), b
    # block 1's ghost layers i = mx, mx+1 are block 2's points i = 1, 2
    b0, b1 = case.blocks
    for q in (1, 2):
        assert np.array_equal(b0.xg[g:-g, g:-g, g + b0.mx - 1 + q], b1.x[:, :, q])
        assert np.array_equal(b1.xg[g:-g, g:-g, g - q], b0.x[:, :, b0.mx - 1 - q])
    # the edge ghosts along the inner face come from the parallelogram rule
    # applied after the neighbour's points were copied (src/Mesh.cpp:859-937):
    # x(i2, -1) = x(i2, 0) + x(mx-1, -1) - x(mx-1, 0) with x(i2, 0) block 2's point
    mx0 = b0.mx
    for q in (1, 2):
        for k in range(b0.mz):
            kk = g + k
            got = b0.xg[kk, g - 1, g + mx0 - 1 + q]
            ref = (b1.x[k, 0, q] + b0.xg[kk, g - 1, g + mx0 - 1]) - b0.x[k, 0, mx0 - 1]
            assert got == ref, (q, k)
    # and differ from the single-block reflection (the warp makes it matter)
    from openstc_amd.runtime import block_geometry
    xs = block_geometry(b0.x, b0.y, b0.z, g)[0]
    assert not np.array_equal(xs, b0.xg)


@pytest.mark.gpu
def test_warped_box_split_fp64_vs_oracle():
    case = cases.warped_box_split(mx=32, my=24, mz=8)
    print("[gpu-mb] warped_box split 2 blocks FP64", flush=True)
    pr = Pair(case, precision=L.FP64)
    pr.init()
    for b in range(2):
        for f in ("BV", "CV", "VEL", "SOUND"):
            e = pr.err(f, b=b, halo=2)
            print("[gpu-mb] init block", b, f, e, flush=True)
            assert e < 1e-13, (b, f, e)
    for s in range(10):
        pr.gpu.step()
        pr.orc.step()
    for b in range(2):
        e = pr.err("BV", b=b, halo=2)
        print("[gpu-mb] block", b, "BV after 10 steps", e, flush=True)
        assert e < 1e-11, (b, e)
    pr.close()


@pytest.mark.gpu
def test_jet_crossflow_small_fp64_vs_oracle():
    case = cases.jet_crossflow(mx=16, my=12, mz=10, nblocks=4, steps=3, precision=L.FP64, cfl=0.5)
    print("[gpu-mb] jet_crossflow 4 blocks 16x12x10 FP64, SST + FR, 3 steps", flush=True)
    pr = Pair(case, precision=L.FP64)
    pr.init()
    for b in range(4):
        e = pr.err("BV", b=b, halo=2)
        print("[gpu-mb] init block", b, e, flush=True)
        assert e < 1e-13, (b, e)
    for s in range(3):
        pr.gpu.step()
        pr.orc.step()
    ns = case.param.n_spec
    for b in range(4):
        e = pr.err("BV", b=b, halo=2)
        print("[gpu-mb] block", b, "BV after 3 steps", e, flush=True)
        assert e < 1e-11, (b, e)
        # species: absolute mass-fraction difference.  Trace radicals in the
        # H2-free vitiated air sit near round-off levels, so a difference
        # scaled by such a plane's own maximum measures nothing physical.
        a = interior(pr.zg(b).download("SV"), case.ngg, 2)
        o = interior(pr.oarr("SV", b), case.ngg, 2)
        dy = float(np.abs(a[:ns] - o[:ns]).max())
        turb = [float(np.abs(a[l] - o[l]).max() / np.abs(o[l]).max()) for l in (ns, ns + 1)]
        print("[gpu-mb] block", b, "max |dY|", dy, "k, omega scaled", turb, flush=True)
        assert dy < 1e-10, (b, dy)
        assert max(turb) < 1e-11, (b, turb)
    pr.close()


@pytest.mark.gpu
def test_jet_crossflow_nonreacting_fp64_vs_oracle():
    """The jet-in-crossflow blocks with reactions off (mix_model = MIXTURE):
    SST on a 9-species mixture, whose time step forms the local time step in
    the source kernel and dq0 in the first DPLUR sweep -- 3 steps against the
    oracle on 2 blocks."""
    case = cases.jet_crossflow(mx=16, my=12, mz=10, nblocks=2, steps=3, precision=L.FP64, cfl=0.5)
    case.param.mix_model = L.MIXTURE
    print("[gpu-mb] jet_crossflow 2 blocks 16x12x10 FP64, SST + non-reacting mixture, 3 steps", flush=True)
    pr = Pair(case, precision=L.FP64)
    pr.init()
    for s in range(3):
        pr.gpu.step()
        pr.orc.step()
    ns = case.param.n_spec
    for b in range(2):
        e = pr.err("BV", b=b, halo=2)
        a = interior(pr.zg(b).download("SV"), case.ngg, 2)
        o = interior(pr.oarr("SV", b), case.ngg, 2)
        dy = float(np.abs(a[:ns] - o[:ns]).max())
        turb = [float(np.abs(a[l] - o[l]).max() / np.abs(o[l]).max()) for l in (ns, ns + 1)]
        print("[gpu-mb] block", b, "BV", e, "max |dY|", dy, "k, omega scaled", turb, flush=True)
        assert e < 1e-11, (b, e)
        assert dy < 1e-10, (b, dy)
        assert max(turb) < 1e-11, (b, turb)
    pr.close()

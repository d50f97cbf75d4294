"""CPU checks of the C-ABI boundary: the library loads, exports every symbol
include/ostc.h declares, and its structures match the oracle's layouts."""
import ctypes as C
import os
import re

import pytest

from openstc_amd import _lib as L
import oracle_ffi as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:int|long|const char \*)\s*\*?\s*(ostc_\w+)\s*\(", txt, re.M))
    return names


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    assert declared_symbols() <= set(L.SIGNATURES), sorted(declared_symbols() - set(L.SIGNATURES))


def test_struct_layouts_match_library_and_oracle():
    lib = L.load()
    assert lib.ostc_sizeof_param() == C.sizeof(L.Param)
    assert lib.ostc_sizeof_bc() == C.sizeof(L.BC)
    for a, b in ((L.Param, O.OrcParam), (L.BC, O.OrcBC), (L.InnerFace, O.OrcInnerFace), (L.Face, O.OrcFace)):
        assert C.sizeof(a) == C.sizeof(b)
        assert [f[0] for f in a._fields_] == [f[0] for f in b._fields_]
        for (n, _), (m, _) in zip(a._fields_, b._fields_):
            assert getattr(a, n).offset == getattr(b, m).offset, n


def test_bad_parameters_are_rejected_without_gpu():
    """ostc_ctx_create validates the model before touching the device."""
    lib = L.load()
    p = L.Param()
    p.mix_model = 7
    h = C.c_void_p()
    assert lib.ostc_ctx_create(C.byref(p), C.byref(h)) == 1
    assert b"mix_model" in lib.ostc_last_error()
    p.mix_model = L.AIR
    p.n_spec, p.n_scalar, p.n_var = 0, 0, 6  # inconsistent n_var
    assert lib.ostc_ctx_create(C.byref(p), C.byref(h)) == 1


def test_host_geometry_bit_exact_with_oracle():
    """ostc_block_geometry (host C++) vs the oracle restatement: bit-exact."""
    import numpy as np
    from openstc_amd import cases
    from openstc_amd.runtime import block_geometry

    c = cases.warped_box(mx=20, my=17, mz=9)
    blk = c.blocks[0]
    ob = O.OracleBlock(blk.mx, blk.my, blk.mz, 2, O.copy_struct(O.OrcParam, c.param))
    ob.geometry(blk.x, blk.y, blk.z)
    xo, yo, zo, jac, met, nneg = block_geometry(blk.x, blk.y, blk.z)
    assert np.array_equal(jac, ob.arr["jac"][0])
    assert np.array_equal(met, ob.arr["metric"])
    assert np.array_equal(xo, ob.arr["x"][0]) and np.array_equal(zo, ob.arr["z"][0])

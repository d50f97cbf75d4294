"""Writes tests/golden/reference_case_grid.npz: the coordinates of the
reference's sample case grid (input_files/grid/"grid   0.grd" of the
reference repository: two 2-D blocks, 100x100x1 and 120x158x1, in
millimetres -- the setup's gridScale 0.001 turns them into metres), so that
the tests can rebuild that case directory without the reference present.
The file is read as whitespace-separated numbers in the grid reader's order
(src/Mesh.cpp:597-640): block count, mx my mz per block, then per block all
x, then all y, then all z, i fastest.  Run from the repository root:

    python tests/golden/make_reference_case.py /root/reference
"""
import os
import sys

import numpy as np


def main(ref_root):
    path = os.path.join(ref_root, "input_files", "grid", "grid   0.grd")
    with open(path) as f:
        nums = f.read().split()
    nb = int(nums[0])
    dims = np.array([[int(v) for v in nums[1 + 3 * b:4 + 3 * b]] for b in range(nb)])
    pos = 1 + 3 * nb
    out = {"dims": dims}
    for b, (mx, my, mz) in enumerate(dims):
        n = mx * my * mz
        for name in ("x", "y", "z"):
            out[f"{name}{b}"] = np.array(nums[pos:pos + n], dtype=np.float64).reshape(mz, my, mx)
            pos += n
    assert pos == len(nums), (pos, len(nums))
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "reference_case_grid.npz"), **out)
    print("blocks", dims.tolist(), "numbers", len(nums))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

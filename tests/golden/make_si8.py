"""Convert the reference's Si8 example wavefunctions into a binary fixture.

Input: ``examples/si8/data/psirow{0,1}`` of USTC-Hu-Group/LibISDF (BSD-3-Clause,
Copyright the LibISDF authors).  Each file is ASCII: a ``rows`` line, a ``cols``
line, then the values column-major (reader at ``examples/si8/readpsi.cpp:30-55``).
The two files are the 1-D row blocks of one 4096 x 35 matrix (rank 0 holds grid
rows 0..2047, rank 1 rows 2048..4095, ``README.md:277-299``).

Output: ``si8_psi.npy``, float64, shape (35, 4096): row ``s`` is state ``s`` over
the 16^3 grid in grid order (x fastest), i.e. the column-major Ng x nstate
matrix the ISDF entry point takes.  ``float()`` parses the 20-digit decimals with
correct rounding, as ``operator>>`` does in the reference reader.

Run once in a container that has the reference checkout:
``python tests/golden/make_si8.py /path/to/LibISDF``.
"""
import os
import sys

import numpy as np


def main(ref_root):
    blocks = []
    for rank in range(2):
        path = os.path.join(ref_root, "examples", "si8", "data", "psirow%d" % rank)
        with open(path) as f:
            toks = f.read().split()
        rows, cols = int(toks[0]), int(toks[1])
        vals = np.array([float(t) for t in toks[2:]], dtype=np.float64)
        assert vals.size == rows * cols
        blocks.append(vals.reshape(cols, rows))
    psi = np.ascontiguousarray(np.concatenate(blocks, axis=1))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "si8_psi.npy")
    np.save(out, psi)
    print("wrote", out, psi.shape)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

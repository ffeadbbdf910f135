"""Writes the golden vectors under tests/golden/ (run from the repo root).

reference_input_direct.npy: the direct sum K w (K = 1/r, isinf -> 0 as
test.hpp:18-34, isinf at :29-30) over the reference's example input
(reference_input/field_test.bin, charge_test.bin: copies of PBBFMM3D's
input/*_test.bin), computed with plain numpy, independently of the C oracle.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    d = os.path.join(HERE, "reference_input")
    f = np.fromfile(os.path.join(d, "field_test.bin"), dtype=np.float64)
    N = f.size // 3
    xyz = f.reshape(3, N).T
    q = np.fromfile(os.path.join(d, "charge_test.bin"), dtype=np.float64)[:N]
    out = np.zeros(N)
    for i0 in range(0, N, 500):
        D = np.linalg.norm(xyz[i0:i0 + 500, None, :] - xyz[None, :, :], axis=2)
        with np.errstate(divide="ignore"):
            K = np.where(D > 0, 1.0 / D, 0.0)
        out[i0:i0 + 500] = K @ q
    np.save(os.path.join(HERE, "reference_input_direct.npy"), out)


if __name__ == "__main__":
    main()

"""Pack the reference's CDKLM16 regression outputs into one fixture.

Reads test/timestep50/CDKLM16_*_{eta,u,v}.dat (text, float32 values written as
float64, shape 74x54 including ghost cells; "u"/"v" hold hu/hv) from the
reference checkout and writes tests/golden/timestep50.npz (float32).  These are
data files shipped with the reference's own test suite
(test/schemes/CDKLM16_test.py:112-419, generated by test/Tests2D.ipynb).
Run in the build container only: python tests/golden/make_timestep50.py
"""
import glob
import os

import numpy as np

SRC = "/root/reference/test/timestep50"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "timestep50.npz")


def main():
    arrays = {}
    for path in sorted(glob.glob(os.path.join(SRC, "CDKLM16_*.dat"))):
        key = os.path.basename(path)[len("CDKLM16_"):-len(".dat")]
        arrays[key] = np.loadtxt(path).astype(np.float32)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {len(arrays)} arrays to {OUT}")


if __name__ == "__main__":
    main()

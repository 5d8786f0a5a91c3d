"""Extract a small golden fixture from the reference's stored SOR outputs.

Source (data files held by the reference, read as text, nothing executed):
  scratch/solverTestCases/Single Building/CPU/Final velocity, cell-centered.dat
  scratch/solverTestCases/Single Building/GPU/Final velocity, cell-centered.dat
Each row is `i j k x y z u v w` with cell-centred velocities of the 2018
single-building prototype: 40x40x40 cells, dx = dy = dz = 5 m, a 20 m cube,
double precision, omega 1.78, itermax 500 (CPU: lexicographic, GPU: red-black).

Kept rows: the vertical plane through the building centre (j = 20) and the
horizontal plane k = 3 (inside the building's height), for both files.
Run once in the development container:  python tests/golden/make_legacy_fixture.py
"""
import os
import sys

import numpy as np

REF = "/root/reference/scratch/solverTestCases/Single Building"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "legacy_single_building.npz")


def load(path):
    a = np.loadtxt(path)
    assert a.shape == (65600, 9), a.shape
    return a


def main():
    cpu = load(os.path.join(REF, "CPU", "Final velocity, cell-centered.dat"))
    gpu = load(os.path.join(REF, "GPU", "Final velocity, cell-centered.dat"))
    assert np.array_equal(cpu[:, :3], gpu[:, :3])
    ijk = cpu[:, :3].astype(np.int16)
    keep = (ijk[:, 1] == 20) | (ijk[:, 2] == 3)
    np.savez_compressed(OUT, ijk=ijk[keep], cpu_uvw=cpu[keep, 6:9].astype(np.float64),
                        gpu_uvw=gpu[keep, 6:9].astype(np.float64))
    print(f"wrote {OUT}: {keep.sum()} rows")


if __name__ == "__main__":
    sys.exit(main())

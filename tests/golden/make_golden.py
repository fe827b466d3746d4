"""Writes tests/golden/*.json: results of the CPU oracle on the parity workloads.

Run in the build container (the oracle is a CPU restatement of the reference;
the reference itself is not buildable here).  Each fixture records the sphere
count, the SHA-256 of the sorted FP64 bit patterns of (centre, radius), the
per-run totals of the draw trace and the MT19937 words consumed.
"""
import json
import os
import platform
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import _oracle as O  # noqa: E402
from _common import set_hash  # noqa: E402

CASES = {
    # 3D monodisperse r=0.005 in the periodic unit cube, seed 0 (the bench workload)
    "3d_r0.005_seed0": dict(dim=3, L=1.0, classes=[(0.005, 1.0, 0)], seed=0),
    # 3D monodisperse r=0.002 in the periodic unit cube, seed 0 (~11.5M spheres)
    "3d_r0.002_seed0": dict(dim=3, L=1.0, classes=[(0.002, 1.0, 0)], seed=0),
    # 3D polydisperse: classes r=0.01/0.005/0.0025 at volume fraction 0.10 each,
    # then fill with r=0.0025 (class order as given; equal radii keep their order)
    # 4D monodisperse r=0.02 in the periodic unit hypercube, seed 0 (the reference's 4D voxel
    # path keeps its TabCorner defect: the run ends when its voxel list empties)
    "4d_r0.02_seed0": dict(dim=4, L=1.0, classes=[(0.02, 1.0, 0)], seed=0),
    # test.hxx:104-109 test_bug_2D: throw_spheres<2, 1> in the unit square, r=0.0005 (voxel coordinates
    # past 32 bits in the reference's history)
    "bug2d_2d_r0.0005_seed0": dict(dim=2, L=1.0, classes=[(0.0005, 1.0, 0)], seed=0),
    "3d_poly_seed0": dict(dim=3, L=1.0, classes=[(0.01, 0.10, 0), (0.005, 0.10, 1), (0.0025, 0.10, 2),
                                                     (0.0025, 1.0, 3)], seed=0),
}


def toolchain():
    gcc = subprocess.run(["g++", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    libc = platform.libc_ver()
    return {"g++": gcc, "glibc": "-".join(libc)}


def main(names):
    for name in names:
        c = CASES[name]
        t = time.time()
        o = O.run(c["dim"], [0.0] * c["dim"], [c["L"]] * c["dim"], c["classes"], seed=c["seed"])
        if "classes" in c:
            c = dict(c, classes=[list(x) for x in c["classes"]])
        dt = time.time() - t
        tr = o["trace"]
        out = dict(case=c, n=int(o["n"]), sha256=set_hash(o["pos"], o["rad"]), words=o["words"],
                   draws=int(len(tr)), shots=int(tr[:, 2].sum()), refines=int(tr[:, 5].sum()),
                   rounds=int(tr[:, 4].sum()), nvox_max=int(max(tr[:, 1].max(), tr[:, 6].max())),
                   end=o["msg"], oracle_seconds=round(dt, 2), toolchain=toolchain(),
                   phases={str(int(p)): int((o["phase"] == p).sum()) for p in sorted(set(o["phase"].tolist()))})
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(out, f, indent=1)
        print(name, out)


if __name__ == "__main__":
    main(sys.argv[1:] or list(CASES))

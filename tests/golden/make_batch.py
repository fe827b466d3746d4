"""Writes tests/golden/batch_3d_256.json: the batched workload (256 independent 3D
realizations, box edge L ~ U[0.2, 1.0], radius r ~ U[0.004, 0.008], seeds
0..255, drawn once from numpy's PCG64 with meta-seed 20261203) and the CPU
oracle's results
(uniform_generate<3, 1> on [0, L]^3 with radius r: count and set hash) for
every realization.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

META_SEED = 20261203
N = 256


def realizations():
    g = np.random.Generator(np.random.PCG64(META_SEED))
    L = g.uniform(0.2, 1.0, size=N)
    r = g.uniform(0.004, 0.008, size=N)
    return [dict(L=float(L[k]), r=float(r[k]), seed=k) for k in range(N)]


def main():
    import _oracle as O
    from _common import set_hash
    real = realizations()
    checks = []
    for k in range(N):
        c = real[k]
        t = time.time()
        o = O.run(3, [0.0] * 3, [c["L"]] * 3, [(c["r"], 1.0, 0)], seed=c["seed"])
        checks.append(dict(index=k, n=int(o["n"]), sha256=set_hash(o["pos"], o["rad"]), end=o["msg"],
                           oracle_seconds=round(time.time() - t, 2)))
        print(checks[-1], flush=True)
    out = dict(meta_seed=META_SEED, generator="numpy.random.PCG64", dim=3, nbshots=6000,
               realizations=real, oracle=checks)
    with open(os.path.join(HERE, "batch_3d_256.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()

"""Autotuner of the device's run-time choices, cached per dimension and
box-to-radius ratio.

What is tuned never changes a result, only how the device gets it:
  - the refinement thresholds of a context (rsa_gpu_set_tuning): the voxel-list
    length from which the coverage pass walks voxels in cell-grouped order
    (group_min) and reads the placed spheres from a packed copy (pack_min);
  - the number of realizations batch_run keeps in flight in its device work
    queue (workers).
autotune() and autotune_batch() time the candidates on the workload they are
given, check that every candidate places the same spheres, and store the fastest
in a JSON cache; Domain(..., tuning="auto") and batch_run(workers=None) read it.

The cache key is the dimension and the ratio side / radius in half-octave bins
(the number of cells per dimension, hence the list lengths, follow from it).
The cache file is $RSA_MPI_AMD_TUNING, else ~/.cache/rsa_mpi_amd/tuning.json;
RSA_MPI_AMD_TUNING=off disables it.
"""
import hashlib
import json
import math
import os
import time

import numpy as np

VERSION = 1
# candidate thresholds: 2^18 .. 2^21, and "never" for either choice
NEVER = (1 << 62)
THRESHOLD_GRID = tuple((1 << k, 1 << k) for k in (18, 19, 20, 21)) + ((1 << 20, NEVER), (NEVER, NEVER))
WORKER_GRID = (16, 32, 64, 128)  # realizations in flight in the device work queue


def cache_path():
    p = os.environ.get("RSA_MPI_AMD_TUNING")
    if p == "off":
        return None
    return p or os.path.join(os.path.expanduser("~"), ".cache", "rsa_mpi_amd", "tuning.json")


def ratio_bin(side, radius):
    """Half-octave bin of side / radius (the key's second part)."""
    if not (side > 0 and radius > 0):
        raise ValueError("side and radius must be positive")
    return round(2.0 * math.log2(float(side) / float(radius))) / 2.0


def key(dim, side, radius, kind="context"):
    return "%s:%dd:%g" % (kind, int(dim), ratio_bin(side, radius))


def load(path=None):
    path = path or cache_path()
    if not path or not os.path.exists(path):
        return {"version": VERSION, "entries": {}}
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return {"version": VERSION, "entries": {}}
    if d.get("version") != VERSION or not isinstance(d.get("entries"), dict):
        return {"version": VERSION, "entries": {}}
    return d


def store(k, entry, path=None):
    path = path or cache_path()
    if not path:
        return None
    d = load(path)
    d["entries"][k] = entry
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp%d" % os.getpid()
    with open(tmp, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
    os.replace(tmp, path)  # readers see the old or the new file, never a partial one
    return path


def lookup(dim, side, radius, path=None):
    """The stored thresholds {group_min, pack_min} for this dimension and ratio, or None."""
    if not (path or cache_path()):
        return None
    e = load(path)["entries"].get(key(dim, side, radius))
    if not e:
        return None
    return {"group_min": int(e["group_min"]), "pack_min": int(e["pack_min"])}


def lookup_batch(dim, sides, radii, path=None):
    """The stored worker count for a batch of this dimension and median ratio, or None."""
    if not (path or cache_path()) or len(sides) == 0:
        return None
    ratio = float(np.median(np.asarray(sides, dtype=float) / np.asarray(radii, dtype=float)))
    e = load(path)["entries"].get(key(dim, ratio, 1.0, "batch"))
    return int(e["workers"]) if e else None


def _set_digest(dom):
    pos, rad, ph = dom.extract_arrays()
    h = hashlib.sha256()
    for a in (pos, rad, ph):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest(), len(rad)


def autotune(dim, side, radius, seed=0, nbshots=6000, grid=THRESHOLD_GRID, repeats=1, save=True, path=None):
    """Times proceed(seed) of the monodisperse periodic box [0, side]^dim with
    `radius` (volume-fraction class, run to maximal packing) for each (group_min,
    pack_min) of `grid`, the best of `repeats` runs each.  Every candidate must place
    the same spheres in the same order (else ValueError).  Returns {"best": {...},
    "seconds": {candidate: s}, "key": ..., "path": ...} and stores the best."""
    from . import Algo, Domain, RadiusGenerator
    times, ref = {}, None
    for gm, pm in grid:
        best = math.inf
        for _ in range(max(1, int(repeats))):
            dom = Domain([0.0] * dim, [float(side)] * dim, float(radius), tuning={"group_min": gm, "pack_min": pm})
            gen = RadiusGenerator([(float(radius), 1.0, 0)], dom.get_total_volume(), 0.0)
            algo = Algo(dom, gen, nbshots)
            t0 = time.perf_counter()
            algo.proceed(seed)
            best = min(best, time.perf_counter() - t0)
            dig = _set_digest(dom)
            if ref is None:
                ref = dig
            elif dig != ref:
                raise ValueError("tuning (%d, %d) changed the placed set" % (gm, pm))
            del dom
        times[(gm, pm)] = best
    gm, pm = min(times, key=times.get)
    entry = {"group_min": gm, "pack_min": pm, "seconds": times[(gm, pm)], "spheres": ref[1],
             "workload": "%dD r=%g side %g seed %d" % (dim, radius, side, seed),
             "measured": {"%d,%d" % c: round(s, 6) for c, s in times.items()}}
    k = key(dim, side, radius)
    where = store(k, entry, path) if save else None
    return {"best": {"group_min": gm, "pack_min": pm}, "seconds": times, "key": k, "path": where}


def autotune_batch(dim, sides, radii, seeds, nbshots=6000, workers=WORKER_GRID, save=True, path=None):
    """Times batch_run over the given realizations for each worker count; every count
    must give the same sphere counts (else ValueError).  Stores and returns the
    fastest count for the batch's dimension and median ratio."""
    from . import batch_run
    times, ref = {}, None
    for w in workers:
        t0 = time.perf_counter()
        out = batch_run(dim, sides, radii, seeds, nbshots=nbshots, workers=int(w))
        times[int(w)] = time.perf_counter() - t0
        counts = out["counts"].tolist()
        if ref is None:
            ref = counts
        elif counts != ref:
            raise ValueError("worker count %d changed the sphere counts" % w)
    best = min(times, key=times.get)
    ratio = float(np.median(np.asarray(sides, dtype=float) / np.asarray(radii, dtype=float)))
    k = key(dim, ratio, 1.0, "batch")
    entry = {"workers": best, "seconds": times[best], "realizations": len(sides),
             "measured": {str(w): round(s, 6) for w, s in times.items()}}
    where = store(k, entry, path) if save else None
    return {"best": best, "seconds": times, "key": k, "path": where}

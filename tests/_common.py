"""Shared helpers of the tests: set hashes and geometric properties of packings."""
import hashlib

import numpy as np


def canonical_bytes(pos, rad):
    """Spheres as FP64 bit patterns [x.., r], rows sorted lexicographically."""
    a = np.concatenate([np.asarray(pos, dtype=np.float64), np.asarray(rad, dtype=np.float64)[:, None]], axis=1)
    bits = np.ascontiguousarray(a).view(np.uint64)
    order = np.lexsort(bits.T[::-1])
    return np.ascontiguousarray(bits[order]).tobytes()


def set_hash(pos, rad):
    return hashlib.sha256(canonical_bytes(pos, rad)).hexdigest()


def periodic_overlaps(pos, rad, L, rel_tol=1e-12):
    """Pairs overlapping by more than rel_tol of the contact distance (periodic box [0, L))."""
    from scipy.spatial import cKDTree
    pos = np.asarray(pos, dtype=np.float64)
    L = np.asarray(L, dtype=np.float64)
    p = np.mod(pos, L)
    tree = cKDTree(p, boxsize=L)
    pairs = tree.query_pairs(2 * float(np.max(rad)) * (1 + 1e-9), output_type="ndarray")
    if len(pairs) == 0:
        return 0
    d = p[pairs[:, 0]] - p[pairs[:, 1]]
    d -= L * np.round(d / L)
    d2 = np.sum(d * d, axis=1)
    s = rad[pairs[:, 0]] + rad[pairs[:, 1]]
    return int(np.sum(d2 < s * s * (1 - rel_tol)))


def uncovered_samples(pos, rad, L, rmin, n=200000, seed=1):
    """Random points of the box farther than r_i + rmin from every centre (0 for a maximal packing)."""
    from scipy.spatial import cKDTree
    L = np.asarray(L, dtype=np.float64)
    rng = np.random.default_rng(seed)
    q = rng.random((n, len(L))) * L
    p = np.mod(np.asarray(pos, dtype=np.float64), L)
    tree = cKDTree(p, boxsize=L)
    reach = float(np.max(rad)) + rmin
    lists = tree.query_ball_point(q, reach * (1 + 1e-9))
    bad = 0
    for i, lst in enumerate(lists):
        if not lst:
            bad += 1
            continue
        d = p[lst] - q[i]
        d -= L * np.round(d / L)
        if not np.any(np.sum(d * d, axis=1) <= (rad[lst] + rmin) ** 2 * (1 + 1e-9)):
            bad += 1
    return bad


def volume_fraction(rad, dim, volume):
    from math import gamma, pi
    vb = pi ** (dim / 2) / gamma(dim / 2 + 1)
    return float(np.sum(vb * np.asarray(rad) ** dim) / volume)

"""ctypes access to the CPU oracle (oracle/librsa_oracle.so) -- test infrastructure.

Only tests/, the smoke() entry and bench.py's cpu_baseline leg use this module.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "librsa_oracle.so")

_L = None
_D = ctypes.POINTER(ctypes.c_double)


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_end_message.restype = ctypes.c_char_p
        L.orc_num_spheres.restype = ctypes.c_int64
        L.orc_num_draws.restype = ctypes.c_int64
        L.orc_words_used.restype = ctypes.c_uint64
        L.orc_num_voxel_values.restype = ctypes.c_int64
        L.orc_num_dumps.restype = ctypes.c_int64
        L.orc_dump_size.restype = ctypes.c_int64
        L.orc_dump_size.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.orc_dist_samples.restype = ctypes.c_uint64
        L.orc_dist_samples.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.c_int64, _D]
        L.orc_mt_words.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                   ctypes.POINTER(ctypes.c_uint32)]
        L.orc_mt_default_words.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint32)]
        for name in ("orc_num_spheres", "orc_num_draws", "orc_words_used", "orc_end_message", "orc_free",
                     "orc_num_voxel_values", "orc_num_dumps"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.orc_spheres.argtypes = [ctypes.c_void_p, _D, _D, ctypes.POINTER(ctypes.c_uint64)]
        L.orc_trace.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), _D]
        L.orc_voxels.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), _D]
        L.orc_dump.argtypes = [ctypes.c_void_p, ctypes.c_int64, _D, _D, ctypes.POINTER(ctypes.c_uint64),
                               ctypes.POINTER(ctypes.c_uint8)]
        _L = L
    return _L


def _d(a):
    return a.ctypes.data_as(_D)


TRACE_COLS = ("draw", "nvox_before", "shots", "accepted", "rounds", "refined", "nvox_after", "words_before",
              "words_after_poisson", "words_after_draw")


def run(dim, inf, sup, classes, seed=0, mode=1, nbshots=6000, max_draws=-1, drad=None, exclusion=0.0,
        compat=True, reiter=(), dump=0):
    """Full proceed<1>(seed).  classes: [(radius, amount, phase)], amount is a
    volume fraction (mode 1), a count (mode 0); mode 2 is a uniform radius law
    with classes[0] = (r_min, count, phase), classes[1] = (r_max, ., .).  reiter:
    re-iterations after the first proceed (same seed), each a radius (one class at
    volume fraction 1) or a tuple ("law", r_min, r_max, phase, count)."""
    L = lib()
    L.orc_set_reiter_law.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                     ctypes.c_uint64]
    L.orc_set_reiter_law(-1, 0.0, 0.0, 0, 0)
    plain = []
    for k, e in enumerate(reiter):
        if isinstance(e, tuple):
            L.orc_set_reiter_law(k, e[1], e[2], e[3], e[4])
            plain.append(0.0)
        else:
            plain.append(float(e))
    reiter = plain
    inf = np.ascontiguousarray(inf, dtype=np.float64)
    sup = np.ascontiguousarray(sup, dtype=np.float64)
    rad = np.ascontiguousarray([c[0] for c in classes], dtype=np.float64)
    amt = np.ascontiguousarray([c[1] for c in classes], dtype=np.float64)
    ph = np.ascontiguousarray([c[2] for c in classes], dtype=np.int32)
    if drad is None:
        drad = float(max(rad))
    rr = np.ascontiguousarray(list(reiter) + [0.0], dtype=np.float64)
    h = ctypes.c_void_p()
    L.orc_set_dump(int(dump))
    rc = L.orc_run(dim, _d(inf), _d(sup), 1, ctypes.c_double(drad), len(classes), _d(rad), _d(amt),
                   ph.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), mode, ctypes.c_double(exclusion), nbshots,
                   ctypes.c_uint64(seed), ctypes.c_int64(max_draws), 1 if compat else 0, len(reiter), _d(rr),
                   ctypes.byref(h))
    L.orc_set_dump(0)
    L.orc_set_reiter_law(-1, 0.0, 0.0, 0, 0)
    if rc:
        raise RuntimeError(L.orc_last_error().decode())
    try:
        n = L.orc_num_spheres(h)
        pos = np.zeros(n * dim)
        r = np.zeros(n)
        phs = np.zeros(n, dtype=np.uint64)
        L.orc_spheres(h, _d(pos), _d(r), phs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        nd = L.orc_num_draws(h)
        tr = np.zeros((nd, 10), dtype=np.int64)
        miss = np.zeros(nd)
        L.orc_trace(h, tr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _d(miss))
        dumps = []
        for k in range(L.orc_num_dumps(h)):
            m = L.orc_dump_size(h, k)
            dp = np.zeros(m * dim)
            dr = np.zeros(m)
            dpr = np.zeros(m, dtype=np.uint64)
            da = np.zeros(m, dtype=np.uint8)
            outr = L.orc_dump(h, k, _d(dp), _d(dr), dpr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                              da.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
            dumps.append(dict(pos=dp.reshape(m, dim), rad=dr, prio=dpr, accepted=da.astype(bool), outreach=outr))
        nvv = L.orc_num_voxel_values(h)
        vox = np.zeros(nvv, dtype=np.int64)
        vlen = np.zeros(dim)
        L.orc_voxels(h, vox.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _d(vlen))
        return dict(n=n, pos=pos.reshape(n, dim), rad=r, phase=phs, trace=tr, miss=miss,
                    msg=L.orc_end_message(h).decode(), words=int(L.orc_words_used(h)), dumps=dumps,
                    vox=vox.reshape(-1, dim), vox_len=vlen)
    finally:
        L.orc_free(h)


def voxel_step(dim, inf, sup, rmax, pos, rad, coords, length, rmin, mode, drad=None, compat=True):
    """remove_covered (0), subdivide_uncovered (1), both (2) or coverage flags (3)
    on a frozen sphere set (validated, with periodic images) and voxel list."""
    L = lib()
    inf = np.ascontiguousarray(inf, dtype=np.float64)
    sup = np.ascontiguousarray(sup, dtype=np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    rad = np.ascontiguousarray(rad, dtype=np.float64)
    coords = np.ascontiguousarray(coords, dtype=np.int64).reshape(-1)
    length = np.ascontiguousarray(length, dtype=np.float64)
    if drad is None:
        drad = rmax
    h = ctypes.c_void_p()
    rc = L.orc_voxel_step(dim, _d(inf), _d(sup), 1, ctypes.c_double(drad), ctypes.c_double(rmax),
                          ctypes.c_int64(len(rad)), _d(pos), _d(rad), ctypes.c_int64(len(coords) // dim),
                          coords.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _d(length), ctypes.c_double(rmin),
                          mode, 1 if compat else 0, ctypes.byref(h))
    if rc:
        raise RuntimeError(L.orc_last_error().decode())
    try:
        nv = L.orc_num_voxel_values(h)
        out = np.zeros(nv, dtype=np.int64)
        ln = np.zeros(dim)
        L.orc_voxels(h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _d(ln))
        if mode == 3:
            return out.astype(bool)
        return out.reshape(-1, dim), ln
    finally:
        L.orc_free(h)


def mt_words(seed, n, rank=0):
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_mt_words(seed, rank, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out


def mt_default_words(n):
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_mt_default_words(n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out


def dist_samples(seed, kind, param, n):
    out = np.zeros(n)
    used = lib().orc_dist_samples(seed, kind, param, n, _d(out))
    return out, int(used)


def mom_pivot(values, nb_first):
    """The oracle's find_pivot_for_first_elements (raises with the reference's message)."""
    L = lib()
    v = np.ascontiguousarray(values, dtype=np.uint64)
    out = ctypes.c_double()
    rc = L.orc_mom_pivot(v.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.c_int64(len(v)),
                         ctypes.c_int64(nb_first), ctypes.byref(out))
    if rc:
        raise RuntimeError(L.orc_last_error().decode())
    return out.value


def neighbour_table(dim, n):
    out = np.zeros(3 ** dim, dtype=np.int32)
    nn = np.ascontiguousarray(n, dtype=np.int32)
    lib().orc_neighbour_table(dim, nn.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return out

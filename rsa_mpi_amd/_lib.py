"""ctypes binding of librsa_mpi_amd.so (the C-ABI declared in include/rsa_mpi_amd.h).

The library is built in-tree by ``rsa_mpi_amd/build.sh``; loading fails loudly when
it is missing -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "librsa_mpi_amd.so")

# every symbol include/rsa_mpi_amd.h declares
EXPORTED = (
    "rsa_gpu_last_error", "rsa_gpu_device_info", "rsa_gpu_create", "rsa_gpu_destroy",
    "rsa_gpu_algo_init", "rsa_gpu_algo_init_fractions", "rsa_gpu_reset_classes",
    "rsa_gpu_set_miss_rate", "rsa_gpu_proceed", "rsa_gpu_add_spheres", "rsa_gpu_count",
    "rsa_gpu_extract", "rsa_gpu_volume_of_spheres", "rsa_gpu_voxels", "rsa_gpu_mt_words",
    "rsa_gpu_profile_enable", "rsa_gpu_kernel_profile", "rsa_gpu_batch_run", "rsa_gpu_algo_init_law",
    "rsa_gpu_reset_law", "rsa_gpu_extract_noalloc", "rsa_gpu_set_tuning", "rsa_gpu_get_tuning",
    "rsa_gpu_learn_reset", "rsa_gpu_learn_set_mode", "rsa_gpu_learn_get_stats", "rsa_gpu_learn_samples",
    "rsa_gpu_learn_train", "rsa_gpu_learn_params", "rsa_gpu_learn_gradcheck", "rsa_gpu_learn_forward",
    "rsa_gpu_batch_queue",
)

# rsa_gpu_radius_law: void (*)(const double* u, double* radius, int64_t n, void* user)
RADIUS_LAW = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), ctypes.c_int64,
                              ctypes.c_void_p)


class Stats(ctypes.Structure):
    """rsa_gpu_stats"""
    _fields_ = [(n, ctypes.c_int64) for n in (
        "loop_iterations", "draws", "shots", "accepted", "refines", "rounds", "rounds_max",
        "nvox_max", "nvox_final", "words", "ghost_cell_drops", "image_cell_mismatch")] + [
        ("end_reason", ctypes.c_int), ("seconds", ctypes.c_double), ("voxel_seconds", ctypes.c_double)] + [
        (n, ctypes.c_int64) for n in ("overlap_pairs", "cover_samples", "cover_free")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(ctypes.Structure):
    """rsa_gpu_tuning"""
    _fields_ = [("group_min", ctypes.c_int64), ("pack_min", ctypes.c_int64)]


class LearnStats(ctypes.Structure):
    """rsa_gpu_learn_stats"""
    _fields_ = [(n, ctypes.c_int64) for n in (
        "voxels_predicted", "voxels_mispredicted", "voxels_covered", "voxels_predicted_covered",
        "voxels_covered_predicted", "waves",
        "waves_uniform", "draws_predicted", "draws_mispredicted", "refines", "refines_predicted",
        "speculation_skipped", "traced_voxels", "traced_draws")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchStats(ctypes.Structure):
    """rsa_gpu_batch_stats"""
    _fields_ = [("seconds", ctypes.c_double)] + [(n, ctypes.c_int64) for n in (
        "ticks", "draw_launch_groups", "refine_launch_groups", "max_active", "draws", "refines")] + [
        (n, ctypes.c_double) for n in ("t_admit", "t_plan", "t_wait", "t_refine", "t_finish", "t_prep", "t_words",
                                        "t_refine_wait")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


NPAR = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2  # the MLP's FP32 parameters

_lib = None

_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_U64 = ctypes.POINTER(ctypes.c_uint64)
_I64 = ctypes.POINTER(ctypes.c_int64)
_I = ctypes.POINTER(ctypes.c_int)
_U32 = ctypes.POINTER(ctypes.c_uint32)


def lib():
    """Load (once) and return the library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run rsa_mpi_amd/build.sh (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    L.rsa_gpu_last_error.restype = ctypes.c_char_p
    L.rsa_gpu_device_info.restype = ctypes.c_char_p
    L.rsa_gpu_create.argtypes = [ctypes.c_int, _D, _D, ctypes.c_int, ctypes.c_double, ctypes.POINTER(_P)]
    L.rsa_gpu_destroy.argtypes = [_P]
    L.rsa_gpu_destroy.restype = None
    L.rsa_gpu_algo_init.argtypes = [_P, ctypes.c_int, _D, _U64, _I, ctypes.c_double, ctypes.c_int]
    L.rsa_gpu_algo_init_fractions.argtypes = [_P, ctypes.c_int, _D, _D, _I, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_int]
    L.rsa_gpu_reset_classes.argtypes = [_P, ctypes.c_int, _D, _U64, _I, ctypes.c_double]
    L.rsa_gpu_set_miss_rate.argtypes = [_P, ctypes.c_double]
    L.rsa_gpu_proceed.argtypes = [_P, ctypes.c_uint64, ctypes.c_int64, ctypes.POINTER(Stats)]
    L.rsa_gpu_add_spheres.argtypes = [_P, ctypes.c_int64, _D, _D, _U64]
    L.rsa_gpu_count.argtypes = [_P, _U64]
    L.rsa_gpu_extract.argtypes = [_P, _D, _D, _U64, ctypes.c_uint64, _U64]
    L.rsa_gpu_extract_noalloc.argtypes = [_P, _P, _P, _P, ctypes.c_uint64, _U64]
    L.rsa_gpu_volume_of_spheres.argtypes = [_P, _D]
    L.rsa_gpu_voxels.argtypes = [_P, _I64, _D, ctypes.c_uint64, _U64]
    L.rsa_gpu_mt_words.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, _U32]
    L.rsa_gpu_profile_enable.argtypes = [_P, ctypes.c_int]
    L.rsa_gpu_kernel_profile.argtypes = [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), _I64, _D, _D]
    L.rsa_gpu_algo_init_law.argtypes = [_P, ctypes.c_double, ctypes.c_double, RADIUS_LAW, ctypes.c_void_p,
                                        ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double, ctypes.c_int]
    L.rsa_gpu_reset_law.argtypes = [_P, ctypes.c_double, ctypes.c_double, RADIUS_LAW, ctypes.c_void_p,
                                    ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double]
    _F = ctypes.POINTER(ctypes.c_float)
    _U8 = ctypes.POINTER(ctypes.c_uint8)
    L.rsa_gpu_learn_reset.argtypes = [ctypes.c_uint64]
    L.rsa_gpu_set_tuning.argtypes = [_P, ctypes.POINTER(Tuning)]
    L.rsa_gpu_get_tuning.argtypes = [_P, ctypes.POINTER(Tuning)]
    L.rsa_gpu_learn_set_mode.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.rsa_gpu_learn_get_stats.argtypes = [_P, ctypes.POINTER(LearnStats)]
    L.rsa_gpu_learn_samples.argtypes = [_I64]
    L.rsa_gpu_learn_train.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_uint64, _D, _D, _I64]
    L.rsa_gpu_learn_params.argtypes = [_F, _F]
    L.rsa_gpu_learn_gradcheck.argtypes = [ctypes.c_int, _F, _U8, _F, _F, _D]
    L.rsa_gpu_learn_forward.argtypes = [ctypes.c_int, _F, _F, _F]
    L.rsa_gpu_batch_run.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _U64, ctypes.c_int, ctypes.c_int, _U64, _D,
                                    _D, _D, ctypes.c_uint64, _D]
    L.rsa_gpu_batch_queue.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _U64, ctypes.c_int, ctypes.c_int, _U64, _D,
                                      _D, _D, ctypes.c_uint64, ctypes.POINTER(BatchStats)]
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise RuntimeError(lib().rsa_gpu_last_error().decode())


def dptr(a):
    return a.ctypes.data_as(_D)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def mt_words(seed, n, rank=0):
    """Raw MT19937 words of seed_seq{rank, seed} from the library's host generator."""
    out = np.zeros(n, dtype=np.uint32)
    check(lib().rsa_gpu_mt_words(seed, rank, n, out.ctypes.data_as(_U32)))
    return out

"""rsa_gpu_extract_noalloc: the extraction into caller-owned device memory that
allocates nothing (its block counts live in a static device array).  Checked
against rsa_gpu_extract on two calls in a row, the second on a smaller packing,
through the C-ABI; device buffers come from the HIP runtime the library uses."""
import ctypes

import numpy as np
import pytest

import rsa_mpi_amd as R
from rsa_mpi_amd import _lib

pytestmark = pytest.mark.gpu


def hip():
    h = ctypes.CDLL("libamdhip64.so.7")  # the runtime librsa_mpi_amd.so is linked with (already loaded)
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return h


D2H = 2


def noalloc_extract(dom, poison=0x7f):
    L = R.lib()
    L.rsa_gpu_extract_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    n = ctypes.c_uint64()
    _lib.check(L.rsa_gpu_extract_noalloc(dom._ctx, None, None, None, 0, ctypes.byref(n)))
    cnt = n.value
    h = hip()
    bufs = []
    for nbytes in (cnt * dom.dim * 8, cnt * 8, cnt * 8):
        p = ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(p), max(nbytes, 8)) == 0
        assert h.hipMemset(p, poison, max(nbytes, 8)) == 0
        bufs.append(p)
    try:
        _lib.check(L.rsa_gpu_extract_noalloc(dom._ctx, bufs[0], bufs[1], bufs[2], cnt, ctypes.byref(n)))
        assert n.value == cnt
        pos = np.zeros(cnt * dom.dim)
        rad = np.zeros(cnt)
        ph = np.zeros(cnt, dtype=np.uint64)
        for host, dev in ((pos, bufs[0]), (rad, bufs[1]), (ph, bufs[2])):
            if host.nbytes:
                assert h.hipMemcpy(host.ctypes.data, dev, host.nbytes, D2H) == 0
        return pos.reshape(cnt, dom.dim), rad, ph
    finally:
        for p in bufs:
            h.hipFree(p)


def packing(dim, L, r, seed):
    dom = R.Domain([0.0] * dim, [L] * dim, r)
    algo = R.Algo(dom, R.RadiusGenerator([(r, 1.0, 0)], dom.get_total_volume(), 0.0), 6000)
    algo.proceed(seed)
    return dom


def test_noalloc_extract_matches_extract_twice():
    """A larger 3D packing, then a smaller 2D one: the scratch keeps the first call's block
    counts, and the second call must not depend on them."""
    for dim, L, r, seed in ((3, 0.3, 0.005, 1), (2, 0.2, 0.01, 2)):
        dom = packing(dim, L, r, seed)
        want_p, want_r, want_ph = dom.extract_arrays()
        got_p, got_r, got_ph = noalloc_extract(dom)
        assert len(got_r) == len(want_r) > 50
        assert np.array_equal(got_p.view(np.uint64), want_p.view(np.uint64))
        assert np.array_equal(got_r.view(np.uint64), want_r.view(np.uint64))
        assert np.array_equal(got_ph, want_ph)


def test_noalloc_extract_capacity_and_empty():
    dom = packing(2, 0.2, 0.01, 0)
    L = R.lib()
    n = ctypes.c_uint64()
    L.rsa_gpu_extract_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    _lib.check(L.rsa_gpu_extract_noalloc(dom._ctx, None, None, None, 0, ctypes.byref(n)))
    assert n.value > 0
    rc = L.rsa_gpu_extract_noalloc(dom._ctx, ctypes.c_void_p(8), ctypes.c_void_p(8), None, n.value - 1,
                                   ctypes.byref(n))
    assert rc != 0 and b"capacity" in L.rsa_gpu_last_error()


def test_noalloc_extract_rejects_grids_beyond_the_scratch(monkeypatch):
    """Cell grids with more blocks than the scratch holds are refused with a clear error
    (the limit lowered to one block of 256 cells through RSA_GPU_NOALLOC_BLOCKS)."""
    dom = packing(2, 0.2, 0.01, 0)  # 12 x 12 cells with the ghost layers: 144 cells, 1 block
    big = packing(2, 0.5, 0.01, 0)  # 27 x 27 = 729 cells, 3 blocks
    monkeypatch.setenv("RSA_GPU_NOALLOC_BLOCKS", "1")
    L = R.lib()
    L.rsa_gpu_extract_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    n = ctypes.c_uint64()
    assert L.rsa_gpu_extract_noalloc(dom._ctx, None, None, None, 0, ctypes.byref(n)) == 0 and n.value > 0
    assert L.rsa_gpu_extract_noalloc(big._ctx, None, None, None, 0, ctypes.byref(n)) != 0
    assert b"scratch holds 256" in L.rsa_gpu_last_error()

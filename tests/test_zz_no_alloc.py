"""The allocation-free dense solve (bh_zgesv_noalloc: device buffers, the
factorisation's work area in the library's fixed device array) against the
allocating entry point bh_zgesv on the same systems: the same LAPACK pivot
order, so equal factors and solutions to rounding (1e-12), on two calls in a
row, the second smaller; sizes beyond the fixed work area are refused
before anything is enqueued."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def _dev(h, a):
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
    assert h.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
    return p


def _get(h, p, like):
    out = np.empty_like(like)
    assert h.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
    return out


def _system(n, nrhs, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    return A, B


def test_noalloc_solve_matches_zgesv():
    import bem3_emf_b2_amd as bem
    bem.init(0)
    L = bem.lib()
    h = _hip()
    for n, nrhs, seed in ((700, 3, 1), (300, 2, 2)):   # the second call smaller than the first
        A, B = _system(n, nrhs, seed)
        # the allocating entry point
        A1, B1 = A.copy(), B.copy()
        bem.check(L.bh_zgesv(n, nrhs, bem.dp(A1.view(np.float64)), bem.dp(B1.view(np.float64))))
        # the allocation-free one on device buffers of the caller
        dA, dB = _dev(h, A.copy()), _dev(h, B.copy())
        dpiv, dinfo = _dev(h, np.zeros(n, np.int32)), _dev(h, np.full(1, 7, np.int32))
        bem.check(L.bh_zgesv_noalloc(n, nrhs, dA, dB, dpiv, dinfo, None))
        assert h.hipDeviceSynchronize() == 0
        A2, B2 = _get(h, dA, A), _get(h, dB, B)
        piv, info = _get(h, dpiv, np.zeros(n, np.int32)), _get(h, dinfo, np.zeros(1, np.int32))
        for p in (dA, dB, dpiv, dinfo):
            h.hipFree(p)
        assert info[0] == 0
        assert np.abs(A2 - A1).max() < 1e-12 * np.abs(A1).max(), n
        assert np.abs(B2 - B1).max() < 1e-12 * np.abs(B1).max(), n
        # and the solution solves the system
        assert np.abs(A @ B2 - B).max() < 1e-10 * np.abs(B).max(), n
        # the pivots are LAPACK's: each is at or below its column
        assert (piv >= np.arange(n)).all() and (piv < n).all()


def test_noalloc_solve_refuses_sizes_beyond_its_work_area():
    import bem3_emf_b2_amd as bem
    bem.init(0)
    L = bem.lib()
    rc = L.bh_zgesv_noalloc(8193, 1, C.c_void_p(1), C.c_void_p(1), C.c_void_p(1), C.c_void_p(1), None)
    assert rc == -3 and b"8192" in L.bh_last_error()

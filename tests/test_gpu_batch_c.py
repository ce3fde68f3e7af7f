"""The drop-in layer's C entry points for fields at many points,
mo_EH_t_batch and mo_EH_s_batch (no counterpart in the reference: one device
call for a whole grid), called through the C library with ctypes on the
reference's stored verification1 state (mo_dat_read of its v1.dat: two metal
cones of triangles and bi-linear quadrangles): against the library's scalar
mo_EH_t / mo_EH_s (mo_field.c:25-100) point by point, 1e-13, with the same
object ids, and against the oracle loaded with the same state, 1e-10."""
import ctypes as C
import gzip
import os

import numpy as np
import pytest

import oracle as orc
from fixtures import GOLDEN, v1_oracle

pytestmark = pytest.mark.gpu


class MOBJ(C.Structure):
    # include/bem3_emf_b2.h: int N; char **mod_fn; double **rs; DOMD *md
    _fields_ = [("N", C.c_int), ("mod_fn", C.c_void_p), ("rs", C.c_void_p), ("md", C.c_void_p)]


def _lib():
    import bem3_emf_b2_amd as bem
    lib = bem.lib()
    D, I = C.POINTER(C.c_double), C.POINTER(C.c_int)
    lib.mo_dat_read.argtypes = [C.c_char_p, C.POINTER(MOBJ)]
    lib.mo_finalize.argtypes = [C.POINTER(MOBJ)]
    for f in (lib.mo_EH_t_batch, lib.mo_EH_s_batch):
        f.argtypes = [D, D, D, C.c_int, C.c_int, C.POINTER(MOBJ), I]
        f.restype = C.c_int
    for f in (lib.mo_EH_t, lib.mo_EH_s):
        f.argtypes = [D, D, D, C.c_int, C.POINTER(MOBJ)]
        f.restype = C.c_int
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_batched_field_entry_points(tmp_path):
    lib = _lib()
    path = str(tmp_path / "v1.dat")
    with open(path, "wb") as fp:
        fp.write(gzip.decompress(open(os.path.join(GOLDEN, "v1.dat.gz"), "rb").read()))
    mo = MOBJ()
    lib.mo_dat_read(path.encode(), C.byref(mo))
    assert mo.N == 2
    rng = np.random.default_rng(17)
    # half inside the two cones' bounding boxes (some points inside a cone),
    # half around them
    box = [rng.uniform([-0.14, -0.14, z0], [0.14, 0.14, z0 + 0.25], (24, 3)) for z0 in (0.075, -0.325)]
    pts = np.ascontiguousarray(np.concatenate(box + [rng.uniform(-0.6, 0.6, (48, 3))]))
    O = v1_oracle()
    worst = 0.0
    for name, mode in (("t", 1), ("s", 0)):
        batch = getattr(lib, "mo_EH_%s_batch" % name)
        scalar = getattr(lib, "mo_EH_%s" % name)
        E = np.zeros((len(pts), 3), np.complex128)
        H = np.zeros((len(pts), 3), np.complex128)
        oid = np.zeros(len(pts), np.int32)
        assert batch(_dp(E.view(np.float64)), _dp(H.view(np.float64)), _dp(pts), len(pts), 1, C.byref(mo),
                     oid.ctypes.data_as(C.POINTER(C.c_int))) == 0
        assert (oid >= 0).any() and (oid < 0).any()   # points inside the cones and outside
        for p in range(len(pts)):
            e = np.zeros(3, np.complex128)
            h = np.zeros(3, np.complex128)
            r = np.ascontiguousarray(pts[p])
            assert scalar(_dp(e.view(np.float64)), _dp(h.view(np.float64)), _dp(r), 1, C.byref(mo)) == oid[p]
            scale = max(np.abs(e).max(), np.abs(h).max(), 1e-300)
            assert max(np.abs(E[p] - e).max(), np.abs(H[p] - h).max()) < 1e-13 * scale, (name, p)
        ref, oid_ref = orc.mo_field(O, pts, "EH", name, 1)
        assert (oid_ref == oid).all()
        got = np.concatenate([E, H], axis=1)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        worst = max(worst, err)
        assert err < 1e-10, (name, err)
    lib.mo_finalize(C.byref(mo))
    print("batched field entry points vs the oracle: %.2e" % worst, flush=True)

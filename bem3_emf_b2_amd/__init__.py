"""bem3_emf_b2_amd -- MI355X (gfx950) implementation of the bem3_emf_b2 hot
path behind a C-ABI (include/bem3_hip.h).

This module is a thin ctypes front end of ``libbem3hip.so`` for tests,
benchmarks and Python users; the C host library (drop-in for the reference's
bem3_emf_b2.h) links the same shared object.  There is no CPU fallback: if
the library or the GPU is missing every call raises.
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libbem3hip.so")
_lib = None

D = C.POINTER(C.c_double)
Fp = C.POINTER(C.c_float)
Ip = C.POINTER(C.c_int)
VP = C.c_void_p


class BemError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BemError("libbem3hip.so is not built (run __graft_entry__.build() or python -m bem3_emf_b2_amd.build)")
        L = C.CDLL(LIB_PATH)
        sig = {
            "bh_last_error": (C.c_char_p, []),
            "bh_init": (C.c_int, [C.c_int]),
            "bh_sync": (C.c_int, []),
            "bh_object_create": (VP, [C.c_int, D, C.c_int, Ip, Ip, Ip, C.c_int, D, C.c_double]),
            "bh_object_destroy": (None, [VP]),
            "bh_object_info": (C.c_int, [VP, Ip]),
            "bh_object_geometry": (C.c_int, [VP, D, D, D, D, Ip]),
            "bh_object_assemble": (C.c_int, [VP, C.c_int]),
            "bh_object_gh_count": (C.c_int, [VP, C.c_int]),
            "bh_object_matrix": (C.c_int, [VP, C.c_int, C.c_int, D]),
            "bh_object_matrix_rows": (C.c_int, [VP, C.c_int, C.c_int, C.c_int, Ip, D]),
            "bh_zgesv": (C.c_int, [C.c_int, C.c_int, D, D]),
            "bh_zgesv_noalloc": (C.c_int, [C.c_int, C.c_int, VP, VP, VP, VP, VP]),
            "bh_object_set_beams": (C.c_int, [VP, C.c_int, D]),
            "bh_object_set_mfb": (C.c_int, [VP, C.c_int, D]),
            "bh_object_factor": (C.c_int, [VP, C.c_int]),
            "bh_object_solve": (C.c_int, [VP]),
            "bh_object_solve_EH": (C.c_int, [VP]),
            "bh_object_state": (C.c_int, [VP, C.c_int, C.c_int, D, C.c_int]),
            "bh_object_incident": (C.c_int, [VP, C.c_int, D, C.c_int]),
            "bh_mo_solve": (C.c_int, [C.POINTER(VP), C.c_int, C.c_double, C.c_int, C.c_int, D, Ip]),
            "bh_mo_field": (C.c_int, [C.POINTER(VP), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, D, D, Ip]),
            "bh_mo_force": (C.c_int, [C.POINTER(VP), C.c_int, C.c_int, C.c_int, D, D, D]),
            "bh_mo_absorb": (C.c_int, [C.POINTER(VP), C.c_int, C.c_int, C.c_int, D]),
            "bh_zgemm_host": (C.c_int, [C.c_int, C.c_int, C.c_int, D, D, D, D, D]),
            "bh_object_domain": (C.c_int, [VP, C.c_int, D, Ip]),
            "bh_mo_EH_bd": (C.c_int, [C.POINTER(VP), C.c_int, C.c_int, C.c_int, C.c_int, Ip, D, C.c_int, C.c_int, D]),
            "bh_prof_enable": (C.c_int, [C.c_int]),
            "bh_prof_read": (C.c_int, [C.c_int, D, D, Ip]),
            "bh_prof_read_bytes": (C.c_int, [C.c_int, D, D, D, Ip]),
            "bh_stream": (VP, []),
            "bh_object_solve_stats": (C.c_int, [VP, Ip, D, Ip]),
            "bh_object_set_precision": (C.c_int, [VP, C.c_int]),
            "bh_cgesv": (C.c_int, [C.c_int, C.c_int, Fp, Fp]),
            "bh_cgemm_host": (C.c_int, [C.c_int, C.c_int, C.c_int, Fp, Fp, Fp, Fp, Fp]),
            "bh_finalize": (C.c_int, []),
            "bh_mo_set_coupling": (C.c_int, [C.c_int]),
            "bh_mo_plan_stats": (C.c_int, [Ip, Ip, Ip]),
            "bh_mo_plan_split": (C.c_int, [Ip]),
            "bh_mo_prepare": (C.c_int, [C.POINTER(VP), C.c_int, C.c_int, C.c_int]),
            "bh_mo_coupling_sums": (C.c_int, [C.POINTER(VP), C.c_int, C.c_int, D]),
            "bh_mo_solve_beams": (C.c_int, [C.POINTER(VP), C.c_int, C.c_int, D, C.c_double, C.c_int, C.c_int, Ip]),
            "bh_object_beam_state": (C.c_int, [VP, C.c_int, C.c_int, C.c_int, D]),
            "bh_object_select_beam": (C.c_int, [VP, C.c_int]),
            "bh_mo_set_lanes": (C.c_int, [C.c_int]),
            "bh_object_set_block_rows": (C.c_int, [VP, C.c_int]),
            "bh_elem_coef": (C.c_int, [VP, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, Ip, Ip, D, D, D, D, D]),
            "bh_mo_force_beams": (C.c_int, [C.POINTER(VP), C.c_int, C.c_int, D, D, D, Ip]),
            "bh_mo_far_field_beams": (C.c_int, [C.POINTER(VP), C.c_int, C.c_int, C.c_int, D, C.c_int, D]),
            "bh_tune_set": (C.c_int, [C.c_int, C.c_int]),
            "bh_tune_get": (C.c_int, [C.c_int, Ip]),
            "bh_tune_load": (C.c_int, [C.c_char_p]),
            "bh_tune_name": (C.c_char_p, [C.c_int]),
            "bh_tune_source": (C.c_char_p, []),
            "bh_gemm_time": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise BemError("bem3_hip error %d: %s" % (rc, lib().bh_last_error().decode()))
    return rc


def dp(a):
    return a.ctypes.data_as(D)


def ip(a):
    return a.ctypes.data_as(Ip)


def init(device=0):
    check(lib().bh_init(device))


def sync():
    check(lib().bh_sync())


# live objects: closed, and the library's device state released, before the
# interpreter shuts down (no device frees from finalizers at exit)
_live = weakref.WeakSet()


def finalize():
    """destroy every live Object, then release the library's device state"""
    for o in list(_live):
        o.close()
    if _lib is not None:
        check(_lib.bh_finalize())


atexit.register(finalize)


def set_coupling(mode):
    """how the coupling loop sums the field of one object at another's
    slots: "auto" (a multipole series where the mixed form allows it) or
    "direct" (the element sums everywhere)"""
    check(lib().bh_mo_set_coupling({"auto": 0, "direct": 1}[mode]))


def plan_stats():
    """(object pairs by series, by element sums, largest series order) of
    the last coupling loop"""
    a = np.zeros(3, np.int32)
    check(lib().bh_mo_plan_stats(ip(a[0:1]), ip(a[1:2]), ip(a[2:3])))
    return int(a[0]), int(a[1]), int(a[2])


def plan_split():
    """object pairs of the last one-beam coupling loop whose series takes only
    some 64-slot tiles of the destination (the rest by element sums)"""
    a = np.zeros(1, np.int32)
    check(lib().bh_mo_plan_split(ip(a)))
    return int(a[0])


# kernel ids of the timing registry (bem3_hip.h)
KERNELS = {"asm_gl": 0, "asm_gh": 1, "asm_self": 2, "zgemm": 3, "lu_panel": 4, "trsm": 5, "field": 6, "scatter": 7,
           "zgemm_small": 8, "field_bd": 9, "cgemm": 10, "cgemm_small": 11, "lu_panel_c": 12, "trsm_c": 13,
           "asm_glr": 14, "field_bd32": 15, "zgemv": 16, "cgemv": 17, "scatter32": 18, "scatter_far": 19,
           "scatter_far32": 20, "mp_basis": 21, "mp_moments": 22, "mp_eval": 23, "field_rows": 24}
# factorisation forms of bh_object_factor
FACTOR_MODES = {"auto": 0, "dense": 1, "block": 2, "mixed": 3}


def prof_enable(on=True):
    check(lib().bh_prof_enable(1 if on else 0))


def prof_read(name):
    """(kernel ms summed over launches, algorithmic work summed, launches)"""
    ms, work, n = np.zeros(1), np.zeros(1), np.zeros(1, np.int32)
    check(lib().bh_prof_read(KERNELS[name], dp(ms), dp(work), ip(n)))
    return float(ms[0]), float(work[0]), int(n[0])


def prof_read_bytes(name):
    """(kernel ms, algorithmic work, algorithmic bytes, launches) summed"""
    ms, work, by, n = np.zeros(1), np.zeros(1), np.zeros(1), np.zeros(1, np.int32)
    check(lib().bh_prof_read_bytes(KERNELS[name], dp(ms), dp(work), dp(by), ip(n)))
    return float(ms[0]), float(work[0]), float(by[0]), int(n[0])


class Object:
    """Device-resident object (one paired triangle mesh and its media)."""

    def __init__(self, rn, ed, md, sd, n_med, lambda0):
        rn = np.ascontiguousarray(rn, dtype=np.float64)
        ed = np.ascontiguousarray(ed, dtype=np.int32)
        md = np.ascontiguousarray(md, dtype=np.int32)
        sd = np.ascontiguousarray(sd, dtype=np.int32)
        n = np.ascontiguousarray(n_med, dtype=np.complex128)
        h = lib().bh_object_create(rn.shape[0] - 1, dp(rn), ed.shape[0] - 1, ip(ed), ip(md), ip(sd), len(n) - 1,
                                   dp(n.view(np.float64)), float(lambda0))
        if not h:
            raise BemError("bh_object_create: " + lib().bh_last_error().decode())
        self.h = C.c_void_p(h)
        info = np.zeros(20, np.int32)
        lib().bh_object_info(self.h, ip(info))
        self.Nn, self.Ne, self.NN, self.MN = (int(x) for x in info[:4])
        self.dNe = [int(x) for x in info[4:4 + self.MN + 1]]
        self.nd = [int(x) for x in info[12:12 + self.MN + 1]]
        _live.add(self)

    def assemble(self, type_=2):
        check(lib().bh_object_assemble(self.h, type_))

    def set_precision(self, prec):
        """"fp64" (default) or "mixed": FP32 far-element quadrature and
        boundary field sums, factor mode 0 -> the FP32 block elimination
        refined in FP64; "mixed_lm": the same with the matrices stored only
        in FP32 and their FP64 rows assembled again where needed (objects
        whose FP64 matrices do not fit the device)"""
        check(lib().bh_object_set_precision(self.h, {"fp64": 0, "mixed": 1, "mixed_lm": 2}[prec]))

    def set_block_rows(self, rows):
        """rows per block of the "mixed_lm" form's on-demand assembly (0: by
        the memory budget)"""
        check(lib().bh_object_set_block_rows(self.h, rows))

    def gh_count(self, did):
        return lib().bh_object_gh_count(self.h, did)

    def matrix(self, did, which):
        nd = self.nd[did]
        k = {"G": 0, "H": 1, "dG": 2, "dH": 3, "dF": 4}[which]
        if k < 2:
            out = np.zeros(nd * nd, np.complex128)
        elif k < 4:
            out = np.zeros(2 * nd * nd, np.complex128)
        else:
            out = np.zeros(3 * nd)
        check(lib().bh_object_matrix(self.h, did, k, dp(out.view(np.float64))))
        if k < 2:
            return out.reshape(nd, nd)
        if k < 4:
            return out.reshape(nd, 2, nd)
        return out.reshape(nd, 3)

    def matrix_rows(self, did, key, rows):
        """selected rows of one domain matrix: G, H (nd columns), dG, dH
        (2 x nd), dF (3)"""
        which = ["G", "H", "dG", "dH", "dF"].index(key)
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        nd = self.nd[did]
        if which < 2:
            out = np.zeros((len(rows), nd), np.complex128)
        elif which < 4:
            out = np.zeros((len(rows), 2, nd), np.complex128)
        else:
            out = np.zeros((len(rows), 3))
        check(lib().bh_object_matrix_rows(self.h, did, which, len(rows), ip(rows), dp(out.view(np.float64))))
        return out

    def geometry(self):
        E1 = self.Ne + 1
        g = dict(cr=np.zeros((E1, 3, 4)), cw=np.zeros((E1, 3, 3)), ren=np.zeros((E1, 4, 3)), wen=np.zeros((E1, 4, 3)),
                 eni=np.zeros((E1, 4), np.int32))
        check(lib().bh_object_geometry(self.h, dp(g["cr"]), dp(g["cw"]), dp(g["ren"]), dp(g["wen"]), ip(g["eni"])))
        return g

    def set_beams(self, beams):
        """rows of 12 (plane waves: ipw.txt columns) or of 16 (any
        multi_fbeam type: type, lambda0, ni, NA or deflection angle, power,
        re/im e0x, re/im e0y, fx, fy, fz, theta, phi, spiral mode, points)"""
        b = np.ascontiguousarray(beams, dtype=np.float64)
        b = b.reshape(-1, b.shape[-1] if b.ndim > 1 else 12)
        if b.shape[1] == 16:
            check(lib().bh_object_set_mfb(self.h, b.shape[0], dp(b)))
        else:
            check(lib().bh_object_set_beams(self.h, b.shape[0], dp(b)))

    def factor(self, mode=0):
        """mode: 0/"auto", 1/"dense", 2/"block", 3/"mixed" (complex-FP32
        block elimination refined in FP64)"""
        if isinstance(mode, str):
            mode = FACTOR_MODES[mode]
        check(lib().bh_object_factor(self.h, mode))

    def solve(self):
        check(lib().bh_object_solve(self.h))

    def solve_stats(self, gmres=False):
        """(refinement steps, final relative residual) of the last solve,
        with gmres=True also the GMRES steps spent on the Schur complement"""
        it, res, gm = np.zeros(1, np.int32), np.zeros(1), np.zeros(1, np.int32)
        check(lib().bh_object_solve_stats(self.h, ip(it), dp(res), ip(gm)))
        if gmres:
            return int(it[0]), float(res[0]), int(gm[0])
        return int(it[0]), float(res[0])

    def solve_EH(self):
        check(lib().bh_object_solve_EH(self.h))

    def state(self, did, which, value=None):
        k = ["U", "dU", "E", "dE", "H", "dH"].index(which)
        nc = 4 if k < 2 else 3
        n = (self.dNe[did] + 1) * 4 * nc
        if value is None:
            buf = np.zeros(n, np.complex128)
            check(lib().bh_object_state(self.h, did, k, dp(buf.view(np.float64)), 0))
            return buf.reshape(self.dNe[did] + 1, 4, nc)
        buf = np.ascontiguousarray(value, dtype=np.complex128).reshape(-1)
        assert buf.size == n
        check(lib().bh_object_state(self.h, did, k, dp(buf.view(np.float64)), 1))

    def beam_state(self, b, did, which):
        """beam b's values after mo_solve_beams: U dU E dE H dH of domain
        did, or Ui dUi Ui0 dUi0"""
        keys = ["U", "dU", "E", "dE", "H", "dH", "Ui", "dUi", "Ui0", "dUi0"]
        k = keys.index(which)
        if k >= 6:
            out = np.zeros((self.Ne + 1, 4, 4), np.complex128)
        else:
            out = np.zeros((self.dNe[did] + 1, 4, 4 if k < 2 else 3), np.complex128)
        check(lib().bh_object_beam_state(self.h, b, did, k, dp(out.view(np.float64))))
        return out

    def nbm(self):
        """beams of the last mo_solve_beams (0: none)"""
        return getattr(self, "_nbm", 0)

    def elem_coef(self, nd, type_, k, s, t=None, ze=None, rt=None, v0=None, v1=None):
        """element coefficients (bh_elem_coef): n requests of signed source
        elements s; t == s -> the self element at local coordinates ze,
        otherwise the points rt; nd derivative directions (v0, v1).  Rows of
        27 complex values: CC, dC/dv0, dC/dv1 (G 0-3, H 4-7, F 8)"""
        s = np.ascontiguousarray(s, dtype=np.int32).reshape(-1)
        n = s.size

        def arr(a, w, dt=np.float64):
            return None if a is None else np.ascontiguousarray(a, dtype=dt).reshape(n, w)
        t, ze, rt, v0, v1 = arr(t, 1, np.int32), arr(ze, 2), arr(rt, 3), arr(v0, 3), arr(v1, 3)
        out = np.zeros((n, 27), np.complex128)
        k = complex(k)
        check(lib().bh_elem_coef(self.h, n, nd, type_, k.real, k.imag, ip(s), None if t is None else ip(t),
                                 None if ze is None else dp(ze), None if rt is None else dp(rt),
                                 None if v0 is None else dp(v0), None if v1 is None else dp(v1),
                                 dp(out.view(np.float64))))
        return out

    def select_beam(self, b):
        """make beam b of mo_solve_beams the object's one-beam state (the
        field, force and power functions then report beam b)"""
        check(lib().bh_object_select_beam(self.h, b))

    def incident(self, which, value=None):
        k = ["Ui", "dUi", "Ui0", "dUi0"].index(which)
        n = (self.Ne + 1) * 16
        if value is None:
            buf = np.zeros(n, np.complex128)
            check(lib().bh_object_incident(self.h, k, dp(buf.view(np.float64)), 0))
            return buf.reshape(self.Ne + 1, 4, 4)
        buf = np.ascontiguousarray(value, dtype=np.complex128).reshape(-1)
        check(lib().bh_object_incident(self.h, k, dp(buf.view(np.float64)), 1))

    def domain(self, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        did = np.zeros(pts.shape[0], np.int32)
        check(lib().bh_object_domain(self.h, pts.shape[0], dp(pts), ip(did)))
        return did

    def close(self):
        if self.h:
            lib().bh_object_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def zgesv(A, B, return_lu=False):
    """X = A^{-1} B on the GPU (LU with partial pivoting); with return_lu the
    packed LU factors (LAPACK zgetrf layout, row-major) come back too."""
    A = np.ascontiguousarray(A, dtype=np.complex128).copy()
    B = np.ascontiguousarray(B, dtype=np.complex128).copy()
    nrhs = 1 if B.ndim == 1 else B.shape[1]
    check(lib().bh_zgesv(A.shape[0], nrhs, dp(A.view(np.float64)), dp(B.view(np.float64))))
    return (B, A) if return_lu else B


def cgesv(A, B):
    """complex-FP32 LU solve on the device (the kernels of the mixed form)"""
    A = np.array(A, dtype=np.complex64, order="C")
    B = np.array(B, dtype=np.complex64, order="C").reshape(A.shape[0], -1)
    check(lib().bh_cgesv(A.shape[0], B.shape[1], A.view(np.float32).ctypes.data_as(Fp),
                         B.view(np.float32).ctypes.data_as(Fp)))
    return B


def cgemm(A, B, Cm=None, alpha=1.0, beta=0.0):
    """C = alpha A B + beta C in complex FP32 on the device"""
    A = np.ascontiguousarray(A, dtype=np.complex64)
    B = np.ascontiguousarray(B, dtype=np.complex64)
    m, k = A.shape
    n = B.shape[1]
    Cm = np.zeros((m, n), np.complex64) if Cm is None else np.array(Cm, dtype=np.complex64, order="C")
    al = np.array([complex(alpha).real, complex(alpha).imag], np.float32)
    be = np.array([complex(beta).real, complex(beta).imag], np.float32)
    f = lambda a: a.ctypes.data_as(Fp)
    check(lib().bh_cgemm_host(m, n, k, f(al), f(A.view(np.float32)), f(B.view(np.float32)), f(be), f(Cm.view(np.float32))))
    return Cm


def zgemm(A, B, Cm=None, alpha=1.0, beta=0.0):
    A = np.ascontiguousarray(A, dtype=np.complex128)
    B = np.ascontiguousarray(B, dtype=np.complex128)
    m, k = A.shape
    n = B.shape[1]
    Cm = np.zeros((m, n), np.complex128) if Cm is None else np.ascontiguousarray(Cm, dtype=np.complex128).copy()
    a = np.array([complex(alpha).real, complex(alpha).imag])
    b = np.array([complex(beta).real, complex(beta).imag])
    check(lib().bh_zgemm_host(m, n, k, dp(a), dp(A.view(np.float64)), dp(B.view(np.float64)), dp(b),
                              dp(Cm.view(np.float64))))
    return Cm


# entries of the tuning table (include/bem3_hip.h, BH_TUNE_*)
TUNE_KEYS = {"lanes": 0, "asm_batch": 1, "gemm_tiles_fp64": 2, "gemm_tiles_fp32": 3, "gemm_longk_tiles": 4,
             "gemm_longk_k": 5, "gemm_bf6": 6}


def tune_get(name=None):
    """one entry of the tuning table, or all of them as a dict"""
    if name is None:
        return {k: tune_get(k) for k in TUNE_KEYS}
    v = C.c_int(0)
    check(lib().bh_tune_get(TUNE_KEYS[name], C.byref(v)))
    return v.value


def tune_set(name, value):
    check(lib().bh_tune_set(TUNE_KEYS[name], int(value)))


def tune_load(path):
    check(lib().bh_tune_load(os.fsencode(path)))


def tune_source():
    return lib().bh_tune_source().decode()


GEMM_KERNELS = {"64x64": 0, "128x128": 1, "bf16split": 2}


def gemm_time(prec, kernel, m, n, k, reps=5):
    """mean ms of one product kernel on the device (prec "fp64" or "fp32";
    kernel "64x64", "128x128" or "bf16split")"""
    ms = C.c_double(0.0)
    check(lib().bh_gemm_time({"fp64": 0, "fp32": 1}[prec], GEMM_KERNELS[kernel], m, n, k, reps, C.byref(ms)))
    return ms.value


def _handles(objs):
    return (C.c_void_p * len(objs))(*[o.h.value for o in objs])


def mo_prepare(objs, type_=2, mode="auto"):
    """assembly of every object as one batch, then the factorisations spread
    over several device streams"""
    check(lib().bh_mo_prepare(_handles(objs), len(objs), type_, FACTOR_MODES[mode]))


def mo_coupling_sums(objs, dst):
    """the loop's coupling sums at object dst's node slots from the current
    boundary values: (U, dU) arrays of shape (Ne+1, 4, 4)"""
    n = objs[dst].Ne + 1
    out = np.zeros(2 * n * 16, np.complex128)
    check(lib().bh_mo_coupling_sums(_handles(objs), len(objs), dst, dp(out.view(np.float64))))
    return out[:n * 16].reshape(n, 4, 4), out[n * 16:].reshape(n, 4, 4)


def mo_solve_beams(objs, beams, ccd=1e-6, iter_min=10, iter_max=100):
    """every beam (rows of 12 doubles) as a separate right-hand side of the
    coupling loop; returns the iteration count"""
    par = np.ascontiguousarray(beams, dtype=np.float64).reshape(-1, 12)
    it = np.zeros(1, np.int32)
    check(lib().bh_mo_solve_beams(_handles(objs), len(objs), par.shape[0], dp(par), ccd, iter_min, iter_max, ip(it)))
    for o in objs:
        o._nbm = par.shape[0]
    return int(it[0])


def set_lanes(n=0):
    """device streams of the per-object work (0: default)"""
    check(lib().bh_mo_set_lanes(n))


def select_beam(objs, b):
    """beam b of mo_solve_beams as every object's one-beam state"""
    for o in objs:
        o.select_beam(b)


def mo_force_beams(objs, type_, rc=None):
    """force and torque of every object under every beam of mo_solve_beams:
    (F, N, flag) of shapes (nb, N, 3), (nb, N, 3), (nb, N); rc: torque
    centres (N, 3) or None (origin)"""
    n = len(objs)
    nb = objs[0].nbm()
    F = np.zeros((nb, n, 3))
    Nq = np.zeros((nb, n, 3))
    flag = np.zeros((nb, n), np.int32)
    rcp = None
    if rc is not None:
        rca = np.ascontiguousarray(rc, dtype=np.float64).reshape(n, 3)
        rcp = dp(rca)
    check(lib().bh_mo_force_beams(_handles(objs), n, type_, rcp, dp(F), dp(Nq), ip(flag)))
    return F, Nq, flag


def mo_far_field_beams(objs, pts, type_=1, what="I"):
    """scattered field of every beam of mo_solve_beams at far points by the
    objects' series: what "I": |E|^2, |H|^2 of shape (nb, np, 2); "EH":
    complex E, H of shape (nb, np, 6)"""
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    nb = objs[0].nbm()
    w = {"EH": 0, "I": 1}[what]
    if w == 0:
        out = np.zeros((nb, pts.shape[0], 6), np.complex128)
        buf = out.view(np.float64)
    else:
        out = np.zeros((nb, pts.shape[0], 2))
        buf = out
    check(lib().bh_mo_far_field_beams(_handles(objs), len(objs), type_, pts.shape[0], dp(pts), w, dp(buf)))
    return out


def far_field_grid(lambda0, mf=2000.0, n_polar=180, n_azimuth=360):
    """points of a far-field grid at radius mf * lambda0 (example4.c's 3-d
    map): polar angles (i + 1/2) pi / n_polar, azimuths j 2 pi / n_azimuth;
    shape (n_polar * n_azimuth, 3), polar-major"""
    ra = mf * lambda0
    ph = (np.arange(n_polar) + 0.5) * np.pi / n_polar
    th = np.arange(n_azimuth) * 2.0 * np.pi / n_azimuth
    P, T = np.meshgrid(ph, th, indexing="ij")
    return np.stack([ra * np.sin(P) * np.cos(T), ra * np.sin(P) * np.sin(T), ra * np.cos(P)], -1).reshape(-1, 3)


def mo_solve(objs, ccd=1e-6, iter_min=10, iter_max=100):
    trace = np.zeros(iter_max + 1)
    it = np.zeros(1, np.int32)
    check(lib().bh_mo_solve(_handles(objs), len(objs), ccd, iter_min, iter_max, dp(trace), ip(it)))
    return int(it[0]), trace


def mo_field(objs, pts, what="EH", mode="t", type_=0):
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    w = {"EH": 0, "mEMP": 1, "EH_mEMP": 2}[what]
    m = {"s": 0, "t": 1, "i": 2}[mode]
    nc = 4 if w == 1 else 6
    out = np.zeros(pts.shape[0] * nc, np.complex128)
    oid = np.zeros(pts.shape[0], np.int32)
    check(lib().bh_mo_field(_handles(objs), len(objs), w, m, type_, pts.shape[0], dp(pts), dp(out.view(np.float64)),
                            ip(oid)))
    return out.reshape(-1, nc), oid


def mo_EH_bd(objs, oid, did, t, ze, mode="t", type_=2):
    """E, H (6 complex per point) at element t[p] of domain did, local
    coordinates ze[p] = (zeta, eta) on object oid's surface"""
    t = np.ascontiguousarray(t, dtype=np.int32).reshape(-1)
    ze = np.ascontiguousarray(ze, dtype=np.float64).reshape(-1, 2)
    m = {"s": 0, "t": 1, "i": 2, "a": 3}[mode]
    out = np.zeros(t.size * 6, np.complex128)
    check(lib().bh_mo_EH_bd(_handles(objs), len(objs), oid, did, t.size, ip(t), dp(ze), type_, m, dp(out.view(np.float64))))
    return out.reshape(-1, 6)


def mo_force(objs, oid, type_, rc=(0.0, 0.0, 0.0)):
    F, Nq = np.zeros(3), np.zeros(3)
    ret = lib().bh_mo_force(_handles(objs), len(objs), oid, type_, dp(np.asarray(rc, np.float64)), dp(F), dp(Nq))
    if ret not in (0, -1):
        check(ret)
    return ret, F, Nq


def mo_absorb(objs, oid, type_):
    P = np.zeros(1)
    ret = lib().bh_mo_absorb(_handles(objs), len(objs), oid, type_, dp(P))
    if ret not in (0, -1, -2):
        check(ret)
    return ret, float(P[0])

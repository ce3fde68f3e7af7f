"""ctypes front end of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the cpu_baseline leg of ``bench.py`` -- never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_DIR, "liborc.so")
_lib = None

D = C.POINTER(C.c_double)
Ip = C.POINTER(C.c_int)
VP = C.c_void_p


def build():
    subprocess.check_call(["make", "-s", "-C", _DIR])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        sig = {
            "orc_init": (None, []),
            "orc_api_quad": (None, [D, D, D, D]),
            "orc_api_from_msh": (VP, [C.c_char_p, C.c_int, D]),
            "orc_api_from_arrays": (VP, [C.c_int, D, C.c_int, Ip, Ip, Ip, Ip, C.c_int, D]),
            "orc_api_free": (None, [VP]),
            "orc_api_setup": (None, [VP, C.c_int, D, D, C.c_int]),
            "orc_api_setup_mfb": (None, [VP, C.c_int, D, D, C.c_int]),
            "orc_api_beam_EH": (None, [VP, C.c_int, D, D, D]),
            "orc_api_override": (None, [VP, D, D, D]),
            "orc_api_refresh_incident": (None, [VP]),
            "orc_api_info": (C.c_int, [VP, Ip]),
            "orc_api_geometry": (None, [VP, D, Ip, Ip, Ip, Ip, D, D, D, D]),
            "orc_api_sid": (None, [VP, C.c_int, Ip]),
            "orc_api_state": (None, [VP, C.c_int, C.c_int, D, C.c_int]),
            "orc_api_incident": (None, [VP, C.c_int, D, C.c_int]),
            "orc_api_coef_rt": (None, [VP, D, C.c_int, D, C.c_int, D]),
            "orc_api_coef_bd": (None, [VP, D, C.c_int, C.c_double, C.c_double, C.c_int, D, C.c_int, D]),
            "orc_api_dcoef_rt": (None, [VP, D, D, C.c_int, D, C.c_int, D, D]),
            "orc_api_dcoef_node_t2": (C.c_int, [VP, C.c_int, C.c_int, C.c_int, D, C.c_int, D, D, D]),
            "orc_api_dcoef_bd_t2": (C.c_int, [VP, D, D, D, C.c_int, C.c_double, C.c_double, C.c_int, D, C.c_int, D, D,
                                              D]),
            "orc_api_cmatrix": (C.c_int, [VP, C.c_int, C.c_int, D, D, D, D, D, Ip]),
            "orc_api_cmatrix_rows": (C.c_int, [VP, C.c_int, C.c_int, C.c_int, C.c_int, D, D, D, D, D, Ip]),
            "orc_api_create_matrix": (VP, [VP, C.c_int, Ip]),
            "orc_api_sys_free": (None, [VP]),
            "orc_api_assemble": (VP, [VP, C.c_int, Ip]),
            "orc_api_build_A": (None, [VP, VP, D]),
            "orc_api_build_B": (None, [VP, VP, D]),
            "orc_api_store_X": (None, [VP, D]),
            "orc_api_solve_EH": (None, [VP, VP]),
            "orc_api_solve_dEH": (C.c_int, [VP, VP]),
            "orc_api_mo_solve": (C.c_int, [C.POINTER(VP), C.POINTER(VP), C.c_int, C.c_double, C.c_int, D]),
            "orc_api_mEMP2_s": (C.c_int, [VP, D, D, C.c_int, D, D]),
            "orc_api_add_scattered": (None, [VP, VP]),
            "orc_api_reset_incident": (None, [VP]),
            "orc_api_ccd": (C.c_double, [C.POINTER(VP), C.c_int]),
            "orc_api_domain_id": (C.c_int, [VP, D]),
            "orc_api_mo_field": (None, [C.POINTER(VP), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, D, D, Ip]),
            "orc_api_mo_EH_bd": (None, [C.POINTER(VP), C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, D]),
            "orc_api_mo_force": (C.c_int, [C.POINTER(VP), C.c_int, C.c_int, C.c_int, D, D, D]),
            "orc_api_mo_absorb": (C.c_int, [C.POINTER(VP), C.c_int, C.c_int, C.c_int, D]),
            "orc_api_zgesv": (C.c_int, [C.c_int, C.c_int, D, D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L.orc_init()
        _lib = L
    return _lib


def dp(a):
    return a.ctypes.data_as(D)


def ip(a):
    return a.ctypes.data_as(Ip)


def cbuf(n):
    return np.zeros(n, dtype=np.complex128)


def quad_tables():
    xl, wl, xh, wh = (np.zeros(12), np.zeros(12), np.zeros(64), np.zeros(64))
    lib().orc_api_quad(dp(xl), dp(wl), dp(xh), dp(wh))
    return xl, wl, xh, wh


class Obj:
    """One object (one mesh with its media) inside the oracle."""

    def __init__(self, handle):
        if not handle:
            raise ValueError("oracle: object construction failed")
        self.h = C.c_void_p(handle)
        info = np.zeros(12, dtype=np.int32)
        lib().orc_api_info(self.h, ip(info))
        self.Nn, self.Ne, self.NN, self.MN = (int(x) for x in info[:4])
        self.dNe = [int(x) for x in info[4:4 + self.MN + 1]]

    @classmethod
    def from_msh(cls, path, n_med):
        n = np.asarray([0j] + list(n_med), dtype=np.complex128)
        return cls(lib().orc_api_from_msh(path.encode(), len(n_med), dp(n.view(np.float64))))

    @classmethod
    def from_arrays(cls, rn, ed, md, sd, gd, n_med):
        rn = np.ascontiguousarray(rn, dtype=np.float64)
        ed = np.ascontiguousarray(ed, dtype=np.int32)
        md, sd = np.ascontiguousarray(md, dtype=np.int32), np.ascontiguousarray(sd, dtype=np.int32)
        gd = np.ascontiguousarray(gd, dtype=np.int32)
        n = np.asarray(n_med, dtype=np.complex128)
        return cls(lib().orc_api_from_arrays(rn.shape[0] - 1, dp(rn), ed.shape[0] - 1, ip(ed), ip(md), ip(sd), ip(gd),
                                             len(n) - 1, dp(n.view(np.float64))))

    def setup(self, beams, tv=None):
        """beams: rows of 12 (plane waves, ipw.txt columns) or of 16 (any
        multi_fbeam type: type, lambda0, ni, NA or deflection angle, power,
        re/im e0x, re/im e0y, fx, fy, fz, theta, phi, spiral mode, points)"""
        b = np.ascontiguousarray(beams, dtype=np.float64)
        b = b.reshape(-1, b.shape[-1] if b.ndim > 1 else 12)
        t = np.zeros(3) if tv is None else np.asarray(tv, dtype=np.float64)
        fn = lib().orc_api_setup_mfb if b.shape[1] == 16 else lib().orc_api_setup
        fn(self.h, b.shape[0], dp(b), dp(t), 0 if tv is None else 1)

    def beam_EH(self, x, v=None):
        """incident E, H (rows of 6; with directions v, rows of 12 with
        dE/dv, dH/dv) of the object's beams at points x"""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        vv = None if v is None else np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
        out = np.zeros((x.shape[0], 6 if v is None else 12), np.complex128)
        lib().orc_api_beam_EH(self.h, x.shape[0], dp(x), dp(vv) if vv is not None else None, dp(out.view(np.float64)))
        return out

    def override(self, ren=None, wen=None, kn=None):
        a = np.ascontiguousarray(ren, dtype=np.float64) if ren is not None else None
        b = np.ascontiguousarray(wen, dtype=np.float64) if wen is not None else None
        k = np.ascontiguousarray(kn, dtype=np.complex128) if kn is not None else None
        lib().orc_api_override(self.h, dp(a) if a is not None else None, dp(b) if b is not None else None,
                               dp(k.view(np.float64)) if k is not None else None)

    def geometry(self):
        E1 = self.Ne + 1
        g = dict(rn=np.zeros((self.Nn + 1, 3)), ed=np.zeros((E1, 4), np.int32), eni=np.zeros((E1, 4), np.int32),
                 md=np.zeros(E1, np.int32), sd=np.zeros(E1, np.int32), cr=np.zeros((E1, 3, 4)), cw=np.zeros((E1, 3, 3)),
                 ren=np.zeros((E1, 4, 3)), wen=np.zeros((E1, 4, 3)))
        lib().orc_api_geometry(self.h, dp(g["rn"]), ip(g["ed"]), ip(g["eni"]), ip(g["md"]), ip(g["sd"]), dp(g["cr"]),
                               dp(g["cw"]), dp(g["ren"]), dp(g["wen"]))
        return g

    def sid(self, d):
        s = np.zeros(self.dNe[d] + 1, np.int32)
        lib().orc_api_sid(self.h, d, ip(s))
        return s

    def state(self, d, which, value=None):
        """which: U dU E dE H dH -> array [dNe+1, 4, 4 or 3]"""
        k = ["U", "dU", "E", "dE", "H", "dH"].index(which)
        nc = 4 if k < 2 else 3
        if value is None:
            buf = cbuf((self.dNe[d] + 1) * 4 * nc)
            lib().orc_api_state(self.h, d, k, dp(buf.view(np.float64)), 0)
            return buf.reshape(self.dNe[d] + 1, 4, nc)
        buf = np.ascontiguousarray(value, dtype=np.complex128).reshape(-1)
        assert buf.size == (self.dNe[d] + 1) * 4 * nc
        lib().orc_api_state(self.h, d, k, dp(buf.view(np.float64)), 1)

    def incident(self, which, value=None):
        k = ["Ui", "dUi", "Ui0", "dUi0"].index(which)
        if value is None:
            buf = cbuf((self.Ne + 1) * 16)
            lib().orc_api_incident(self.h, k, dp(buf.view(np.float64)), 0)
            return buf.reshape(self.Ne + 1, 4, 4)
        buf = np.ascontiguousarray(value, dtype=np.complex128).reshape(-1)
        lib().orc_api_incident(self.h, k, dp(buf.view(np.float64)), 1)

    def cmatrix(self, did, type_=2, want_d=True):
        N = 4 * self.dNe[did]
        G, H = cbuf(N * N), cbuf(N * N)
        dG = cbuf(2 * N * N) if want_d else None
        dH = cbuf(2 * N * N) if want_d else None
        dF = np.zeros(3 * N)
        ngh = np.zeros(1, np.int32)
        bad = lib().orc_api_cmatrix(self.h, did, type_, dp(G.view(np.float64)), dp(H.view(np.float64)),
                                    dp(dG.view(np.float64)) if want_d else None,
                                    dp(dH.view(np.float64)) if want_d else None, dp(dF), ip(ngh))
        out = dict(G=G.reshape(N, N), H=H.reshape(N, N), dF=dF.reshape(N, 3), ngh=int(ngh[0]), nbad=bad)
        if want_d:
            out["dG"] = dG.reshape(N, 2, N)
            out["dH"] = dH.reshape(N, 2, N)
        return out

    def cmatrix_rows(self, did, t0, t1, type_=2):
        """rows of elements t0 <= t < t1 (1-based) of the padded matrices"""
        N = 4 * self.dNe[did]
        R = 4 * (t1 - t0)
        G, H = cbuf(R * N), cbuf(R * N)
        dG, dH = cbuf(2 * R * N), cbuf(2 * R * N)
        dF = np.zeros(3 * R)
        ngh = np.zeros(1, np.int32)
        bad = lib().orc_api_cmatrix_rows(self.h, did, type_, t0, t1, dp(G.view(np.float64)), dp(H.view(np.float64)),
                                         dp(dG.view(np.float64)), dp(dH.view(np.float64)), dp(dF), ip(ngh))
        return dict(G=G.reshape(R, N), H=H.reshape(R, N), dG=dG.reshape(R, 2, N), dH=dH.reshape(R, 2, N),
                    dF=dF.reshape(R, 3), ngh=int(ngh[0]), nbad=bad)

    def domain_id(self, rt):
        return lib().orc_api_domain_id(self.h, dp(np.asarray(rt, dtype=np.float64)))

    def coef_rt(self, rt, s, k, type_):
        CC = cbuf(9)
        kk = np.array([complex(k).real, complex(k).imag])
        lib().orc_api_coef_rt(self.h, dp(np.asarray(rt, np.float64)), s, dp(kk), type_, dp(CC.view(np.float64)))
        return CC

    def dcoef_rt(self, rt, v, s, k, type_):
        CC, dCC = cbuf(9), cbuf(9)
        kk = np.array([complex(k).real, complex(k).imag])
        lib().orc_api_dcoef_rt(self.h, dp(np.asarray(rt, np.float64)), dp(np.asarray(v, np.float64)), s, dp(kk), type_,
                               dp(CC.view(np.float64)), dp(dCC.view(np.float64)))
        return CC, dCC

    def dcoef_node_t2(self, t, nid, s, k, type_=2):
        CC, dz, de = cbuf(9), cbuf(9), cbuf(9)
        kk = np.array([complex(k).real, complex(k).imag])
        rc = lib().orc_api_dcoef_node_t2(self.h, t, nid, s, dp(kk), type_, dp(CC.view(np.float64)),
                                         dp(dz.view(np.float64)), dp(de.view(np.float64)))
        return rc, CC, dz, de

    def coef_bd(self, rt, t, zt, et, s, k, type_):
        CC = cbuf(9)
        kk = np.array([complex(k).real, complex(k).imag])
        lib().orc_api_coef_bd(self.h, dp(np.asarray(rt, np.float64)), t, zt, et, s, dp(kk), type_,
                              dp(CC.view(np.float64)))
        return CC

    def dcoef_bd_t2(self, rt, vtz, vte, t, zt, et, s, k, type_=2):
        CC, dz, de = cbuf(9), cbuf(9), cbuf(9)
        kk = np.array([complex(k).real, complex(k).imag])
        rc = lib().orc_api_dcoef_bd_t2(self.h, dp(np.asarray(rt, np.float64)), dp(np.asarray(vtz, np.float64)),
                                       dp(np.asarray(vte, np.float64)), t, zt, et, s, dp(kk), type_,
                                       dp(CC.view(np.float64)), dp(dz.view(np.float64)), dp(de.view(np.float64)))
        return rc, CC, dz, de

    def mEMP2_s(self, rt, v, type_=2):
        U, dU = cbuf(4), cbuf(4)
        did = lib().orc_api_mEMP2_s(self.h, dp(np.asarray(rt, np.float64)), dp(np.asarray(v, np.float64)), type_,
                                    dp(U.view(np.float64)), dp(dU.view(np.float64)))
        return did, U, dU

    def free(self):
        if self.h:
            lib().orc_api_free(self.h)
            self.h = None


class System:
    """Factored 8NN system of one object (reference create_matrix); with
    factor=False only the coefficient matrices are built."""

    def __init__(self, obj, type_=2, factor=True):
        st = np.zeros(1, np.int32)
        fn = lib().orc_api_create_matrix if factor else lib().orc_api_assemble
        self.h = C.c_void_p(fn(obj.h, type_, ip(st)))
        self.status = int(st[0])
        self.obj = obj
        if factor and self.status != 0:
            raise RuntimeError("oracle create_matrix failed (status %d)" % self.status)

    def solve_EH(self):
        lib().orc_api_solve_EH(self.obj.h, self.h)

    def solve_dEH(self):
        return lib().orc_api_solve_dEH(self.obj.h, self.h)

    def free(self):
        if self.h:
            lib().orc_api_sys_free(self.h)
            self.h = None


def _handles(objs):
    arr = (C.c_void_p * len(objs))(*[o.h.value for o in objs])
    return arr


def mo_solve(objs, systems, ccd=1e-6, iter_max=100):
    trace = np.zeros(iter_max + 1)
    it = lib().orc_api_mo_solve(_handles(objs), (C.c_void_p * len(systems))(*[s.h.value for s in systems]),
                                len(objs), ccd, iter_max, dp(trace))
    return it, trace


def mo_field(objs, pts, what="EH", mode="t", type_=0):
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    w = {"EH": 0, "mEMP": 1, "EH_mEMP": 2}[what]
    m = {"s": 0, "t": 1, "i": 2}[mode]
    nc = 4 if w == 1 else 6
    out = cbuf(pts.shape[0] * nc)
    oid = np.zeros(pts.shape[0], np.int32)
    lib().orc_api_mo_field(_handles(objs), len(objs), w, m, type_, pts.shape[0], dp(pts), dp(out.view(np.float64)), ip(oid))
    return out.reshape(-1, nc), oid


def mo_EH_bd(objs, oid, did, t, zt, et, type_=2, mode="t"):
    out = cbuf(6)
    lib().orc_api_mo_EH_bd(_handles(objs), len(objs), oid, did, t, zt, et, type_, {"s": 0, "t": 1, "i": 2, "a": 3}[mode],
                           dp(out.view(np.float64)))
    return out


def mo_force(objs, oid, type_, rc=(0.0, 0.0, 0.0)):
    F, Nq = np.zeros(3), np.zeros(3)
    ret = lib().orc_api_mo_force(_handles(objs), len(objs), oid, type_, dp(np.asarray(rc, np.float64)), dp(F), dp(Nq))
    return ret, F, Nq


def mo_absorb(objs, oid, type_):
    P = np.zeros(1)
    ret = lib().orc_api_mo_absorb(_handles(objs), len(objs), oid, type_, dp(P))
    return ret, float(P[0])


def add_scattered(dst, src):
    lib().orc_api_add_scattered(dst.h, src.h)


def zgesv(A, B):
    A = np.ascontiguousarray(A, dtype=np.complex128).copy()
    B = np.ascontiguousarray(B, dtype=np.complex128).copy()
    n = A.shape[0]
    nrhs = B.shape[1] if B.ndim == 2 else 1
    info = lib().orc_api_zgesv(n, nrhs, dp(A.view(np.float64)), dp(B.view(np.float64)))
    return info, B

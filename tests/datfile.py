"""Reader for the reference's solved-state file (``mo_dat_write`` layout,
mo_setup.c:147-203; read back by mo_dat_read, mo_setup.c:205-436).

Binary layout per file: int N, then per object: mod_fn[256], rs[3],
med_fn[128], msh_fn[128], rv[3], th, tv[3], int MN, complex n[MN+1],
complex kn[MN+1], a raw Bobj (264 bytes, multi_fbeam.h:28-41) followed by
the raw beam records of each type in turn (Ipw 240 bytes, ipw.h:79-99; Fpw
272, fpw.h:14-39; LGb 280, lgb.h:14-41; Bsb 248, bsb.h:13-37; BsLGb 264,
bslgb.h:13-39; RAb 272, rab.h:13-44), the boundary tables and, per domain,
sid and the U/dU/E/dE/H/dH node values.
"""
import gzip

import numpy as np

BOBJ_SIZE = 264
IPW_SIZE = 240


class _Rd:
    def __init__(self, buf):
        self.b = buf
        self.p = 0

    def take(self, dtype, count):
        dt = np.dtype(dtype)
        n = dt.itemsize * count
        a = np.frombuffer(self.b, dtype=dt, count=count, offset=self.p)
        self.p += n
        return a.copy()

    def raw(self, n):
        r = self.b[self.p:self.p + n]
        self.p += n
        return r


def _ipw_params(rec):
    """user parameters of one Ipw record as the 12-vector used by the oracle"""
    d = np.frombuffer(rec, dtype=np.float64)
    lam, ni, pw = d[0], d[1], d[2]
    e0x = (d[3], d[4])
    e0y = (d[5], d[6])
    fx, fy, fz, th, ph = d[7], d[8], d[9], d[10], d[11]
    return np.array([lam, ni, pw, e0x[0], e0x[1], e0y[0], e0y[1], fx, fy, fz, th, ph])


# record size and the byte offset of the trailing int members (spiral mode,
# quadrature points) of the other beam types; the doubles lambda0, ni,
# NA / deflection angle, power, e0x, e0y, fx, fy, fz, theta, phi come first
_FB = {1: (272, None, 104), 2: (280, 104, 108), 3: (248, None, None), 4: (264, 104, None), 5: (272, None, 104)}


def _fb_params(typ, rec):
    """one beam record as the 16-vector (type, lambda0, ni, NA or deflection
    angle, power, re/im e0x, re/im e0y, fx, fy, fz, theta, phi, mode, points)"""
    d = np.frombuffer(rec, dtype=np.float64, count=13)
    _, om, on = _FB[typ]
    m = int(np.frombuffer(rec, np.int32, 1, om)[0]) if om is not None else 0
    nn = int(np.frombuffer(rec, np.int32, 1, on)[0]) if on is not None else 0
    return np.concatenate([[typ], d, [m, nn]])


def read_dat(path):
    raw = open(path, "rb").read()
    if raw[:2] == b"\x1f\x8b":
        raw = gzip.decompress(raw)
    r = _Rd(raw)
    N = int(r.take(np.int32, 1)[0])
    objs = []
    for _ in range(N):
        o = {}
        o["mod_fn"] = r.raw(256).split(b"\0")[0].decode()
        o["rs"] = r.take(np.float64, 3)
        o["med_fn"] = r.raw(128).split(b"\0")[0].decode()
        o["msh_fn"] = r.raw(128).split(b"\0")[0].decode()
        o["rv"] = r.take(np.float64, 3)
        o["th"] = float(r.take(np.float64, 1)[0])
        o["tv"] = r.take(np.float64, 3)
        MN = int(r.take(np.int32, 1)[0])
        o["MN"] = MN
        o["n"] = r.take(np.complex128, MN + 1)
        o["kn"] = r.take(np.complex128, MN + 1)
        bobj = r.raw(BOBJ_SIZE)
        counts = [int(np.frombuffer(bobj, np.int32, 1, 32 * i)[0]) for i in range(6)]
        o["n0"], o["lambda0"], o["omega"] = np.frombuffer(bobj, np.float64, 3, 192)
        o["beams"] = np.array([_ipw_params(r.raw(IPW_SIZE)) for _ in range(counts[0])]).reshape(-1, 12)
        mfb = [np.concatenate([[0], b[:2], [0.0], b[2:], [0, 0]]) for b in o["beams"]]
        for typ in range(1, 6):
            mfb += [_fb_params(typ, r.raw(_FB[typ][0])) for _ in range(counts[typ])]
        o["mfb"] = np.array(mfb).reshape(-1, 16)
        Nn = int(r.take(np.int32, 1)[0])
        Ne = int(r.take(np.int32, 1)[0])
        o["rn"] = r.take(np.float64, 3 * (Nn + 1)).reshape(Nn + 1, 3)
        o["ed"] = r.take(np.int32, 4 * (Ne + 1)).reshape(Ne + 1, 4)
        o["NN"] = int(r.take(np.int32, 1)[0])
        o["eni"] = r.take(np.int32, 4 * (Ne + 1)).reshape(Ne + 1, 4)
        o["md"] = r.take(np.int32, Ne + 1)
        o["sd"] = r.take(np.int32, Ne + 1)
        o["gd"] = r.take(np.int32, Ne + 1)
        o["ren"] = r.take(np.float64, 12 * (Ne + 1)).reshape(Ne + 1, 4, 3)
        o["wen"] = r.take(np.float64, 12 * (Ne + 1)).reshape(Ne + 1, 4, 3)
        doms = []
        for _d in range(MN + 1):
            dd = {}
            dn = int(r.take(np.int32, 1)[0])
            dd["Ne"] = dn
            dd["sid"] = r.take(np.int32, dn + 1)
            for key, nc in (("U", 4), ("dU", 4), ("E", 3), ("dE", 3), ("H", 3), ("dH", 3)):
                dd[key] = r.take(np.complex128, (dn + 1) * 4 * nc).reshape(dn + 1, 4, nc)
            doms.append(dd)
        o["dom"] = doms
        objs.append(o)
    if r.p != len(raw):
        raise ValueError("trailing bytes in %s: %d" % (path, len(raw) - r.p))
    return objs


def parse_example1_log(path):
    """Numbers printed by example1.c: fields (type 0, 1), F/N per object and
    type, absorbed power per object and type."""
    import re
    txt = open(path).read().splitlines()
    num = r"([-+]?\d\.\d+e[-+]\d+)"
    out = {"field": {}, "force": {}, "absorb": {}}
    typ = None
    sect = None
    oid = None
    cur = {}
    for line in txt:
        m = re.match(r"type=(\d) setting", line)
        if m:
            typ = int(m.group(1))
            cur = {}
            if sect == "field" or sect is None:
                out["field"][typ] = cur
            continue
        if line.startswith("Total electromagnetic field"):
            sect = "field"
            continue
        if line.startswith("Radiation force"):
            sect = "force"
            continue
        if line.startswith("Absorbed energy"):
            sect = "absorb"
            continue
        m = re.match(r"(Ex|Ey|Ez|Hx|Hy|Hz) = " + r"\s*" + num + r"\s*" + num + " I", line)
        if m and sect == "field":
            out["field"][typ][m.group(1)] = complex(float(m.group(2)), float(m.group(3)))
            continue
        m = re.match(r"ob[jc]+e?c?t id = (\d+)", line.replace("objcet", "object"))
        if m:
            oid = int(m.group(1))
            continue
        m = re.match(r"F = \(\s*" + num + r",\s*" + num + r",\s*" + num + r"\)", line)
        if m:
            out["force"].setdefault(typ, {}).setdefault(oid, {})["F"] = np.array([float(m.group(i)) for i in (1, 2, 3)])
            continue
        m = re.match(r"N = \(\s*" + num + r",\s*" + num + r",\s*" + num + r"\)", line)
        if m:
            out["force"].setdefault(typ, {}).setdefault(oid, {})["N"] = np.array([float(m.group(i)) for i in (1, 2, 3)])
            continue
        m = re.match(r"P =\s*" + num, line)
        if m:
            out["absorb"].setdefault(typ, {})[oid] = float(m.group(1))
    return out

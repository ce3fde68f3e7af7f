"""Shared loaders for the reference-held solved-state fixtures: v1.dat (two
cones of verification1/cone_m1_p5.msh, 28 triangles + 553 quads each, a
plane wave; expected outputs in verification1/example1.log) and ex30.dat
(the glass hemisphere of analysis_sample3/hemisphere_m1.msh, 1008 quads, a
focused plane wave; expected outputs in analysis_sample3/example1_ex30.log)."""
import os

import numpy as np

import oracle as orc
from datfile import parse_example1_log, read_dat

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_cache = {}


def v1_raw():
    if "raw" not in _cache:
        _cache["raw"] = read_dat(os.path.join(GOLDEN, "v1.dat.gz"))
    return _cache["raw"]


def v1_log():
    return parse_example1_log(os.path.join(GOLDEN, "v1_example1.log"))


def ex30_raw():
    if "ex30" not in _cache:
        _cache["ex30"] = read_dat(os.path.join(GOLDEN, "ex30.dat.gz"))
    return _cache["ex30"]


def ex30_log():
    return parse_example1_log(os.path.join(GOLDEN, "ex30_example1.log"))


def ex30_oracle():
    return stored_oracle(ex30_raw())


def v1_oracle():
    return stored_oracle(v1_raw())


def stored_oracle(raw):
    """oracle objects loaded with the stored geometry, beams and solved state"""
    objs = []
    for o in raw:
        ob = orc.Obj.from_arrays(o["rn"], o["ed"], o["md"], o["sd"], o["gd"], o["n"])
        ob.setup(o["mfb"])
        ob.override(ren=o["ren"], wen=o["wen"], kn=o["kn"])
        for d in range(o["MN"] + 1):
            for key in ("U", "dU", "E", "dE", "H", "dH"):
                ob.state(d, key, o["dom"][d][key])
        objs.append(ob)
    return objs


def active_slots(ed, sid):
    """mask over padded (element, slot) pairs that carry a node"""
    tri = ed[np.abs(sid[1:]), 3] == 0
    act = np.ones((len(sid) - 1, 4), bool)
    act[tri, 3] = False
    return act.reshape(-1)


REF_BIN = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref")
DROPIN_HEADER = os.path.join(os.path.dirname(GOLDEN), "..", "include", "bem3_emf_b2.h")


def ref_example(name):
    """path of the reference's example program relinked against this library
    (make -C oracle ref, which records the hash of include/bem3_emf_b2.h it
    compiled against); skips the test when it was never built here (the
    reference tree is not mounted), and fails it when it was built against
    another version of the header"""
    import hashlib
    import pytest
    path = os.path.join(REF_BIN, name)
    if not os.path.exists(path):
        pytest.skip("reference %s not compiled against the drop-in library (make -C oracle ref)" % name)
    stamp = os.path.join(REF_BIN, "header.sha256")
    want = hashlib.sha256(open(DROPIN_HEADER, "rb").read()).hexdigest()
    got = open(stamp).read().split()[0] if os.path.exists(stamp) else None
    if got != want:
        pytest.fail("%s was compiled against another include/bem3_emf_b2.h (%s, now %s): make -C oracle ref"
                    % (name, got, want))
    return path

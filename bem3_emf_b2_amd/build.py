"""Build recipe of the HIP library (gfx950).  Used by __graft_entry__.build()
and by hand: ``python -m bem3_emf_b2_amd.build``."""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libbem3hip.so")
SOURCES = ["bem_asm.hip", "bem_la.hip", "bem_field.hip", "bem_bie.hip", "bem_api.hip", "bem_prof.hip", "bem_schur.hip", "bem_mo.hip", "bem_tune.hip"]
# host side of the drop-in API (include/bem3_emf_b2.h), plain C
HOST_SOURCES = ["host/bem_host.c"]
# command-line programs of the reference, built against the drop-in API
PROGRAMS = ["d3b2_create_matrix", "d3b2_bv_solver"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")
BIN = os.path.join(PKG, "bin")
CFLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wextra", "-Wno-unused-parameter", "-I", os.path.join(ROOT, "include")]

FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", os.path.join(ROOT, "include")]


def _compile(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=False):
    from concurrent.futures import ThreadPoolExecutor
    objs, cmds = [], []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        if not os.path.exists(src):
            continue
        obj = os.path.join(CSRC, s.replace(".hip", ".o"))
        cmds.append([HIPCC] + FLAGS + ["-c", src, "-o", obj])
        objs.append(obj)
    # the device sources compile independently (a few at a time: hipcc is memory hungry)
    with ThreadPoolExecutor(max_workers=min(len(cmds), max(1, min(8, os.cpu_count() or 1)))) as ex:
        for f in [ex.submit(_compile, c, verbose) for c in cmds]:
            f.result()
    for s in HOST_SOURCES:
        src = os.path.join(CSRC, s)
        obj = src[:-2] + ".o"
        cmd = [CC] + CFLAGS + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs + ["-lm"]
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    for o in objs:
        os.remove(o)
    os.makedirs(BIN, exist_ok=True)
    for p in PROGRAMS:
        out = os.path.join(BIN, p)
        cmd = [CC] + CFLAGS + [os.path.join(CSRC, "host", p + ".c"), "-o", out + ".tmp", "-L", PKG, "-lbem3hip",
                               "-Wl,-rpath,$ORIGIN/..", "-lm"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        os.replace(out + ".tmp", out)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)

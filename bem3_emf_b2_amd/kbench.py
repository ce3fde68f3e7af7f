"""Stand-alone runs of single device kernels of the library at a chosen
shape, for timing and hardware-counter collection without the rest of the
pipeline:

  python -m bem3_emf_b2_amd.kbench gemm --m 15360 --n 15360 --k 15360 --prec c128 --reps 3

runs C = A B (complex, row-major, device-resident operands of the given
shape) through the same dispatch the solver uses (the 128x128-tile kernel at
this size) and prints one JSON line with the average time per launch and the
achieved rate.  The 15360^3 product is the largest GEMM of the block
elimination of a 5120-triangle object (M = H1 + G1 Z)."""
import argparse
import ctypes as C
import json
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("kernel", choices=["gemm"])
    ap.add_argument("--m", type=int, default=15360)
    ap.add_argument("--n", type=int, default=15360)
    ap.add_argument("--k", type=int, default=15360)
    ap.add_argument("--prec", choices=["c128", "c64"], default="c128")
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()

    import bem3_emf_b2_amd as bem
    bem.init(0)
    L = bem.lib()
    L.bem_zgemm.restype = None
    L.bem_cgemm.restype = None
    L.bem_zgemm.argtypes = L.bem_cgemm.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p,
                                                   C.c_size_t, C.c_void_p, C.c_size_t, C.c_double, C.c_double,
                                                   C.c_void_p, C.c_size_t, C.c_void_p]
    # device buffers through the HIP runtime the library itself uses
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    esz = 16 if args.prec == "c128" else 8
    cdt = np.complex128 if args.prec == "c128" else np.complex64
    rng = np.random.default_rng(1)

    def dev(shape, fill=True):
        p = C.c_void_p()
        n = shape[0] * shape[1]
        assert hip.hipMalloc(C.byref(p), n * esz) == 0
        h = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(cdt) if fill else np.zeros(shape, cdt)
        assert hip.hipMemcpy(p, h.ctypes.data, n * esz, 1) == 0   # host to device
        return p, h

    A, hA = dev((args.m, args.k))
    B, hB = dev((args.k, args.n))
    Cm, _ = dev((args.m, args.n), fill=False)
    f = L.bem_zgemm if args.prec == "c128" else L.bem_cgemm
    st = L.bh_stream()
    kid = "zgemm" if args.prec == "c128" else "cgemm"
    call = lambda: f(args.m, args.n, args.k, 1.0, 0.0, A, args.k, B, args.n, 0.0, 0.0, Cm, args.n, st)
    call()
    bem.sync()
    bem.prof_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.reps):
        call()
    bem.sync()
    wall = (time.perf_counter() - t0) / args.reps
    ms, work, n = bem.prof_read(kid)
    bem.prof_enable(False)
    # spot check of the first row against numpy
    row = np.zeros(args.n, cdt)
    assert hip.hipMemcpy(row.ctypes.data, Cm, args.n * esz, 2) == 0   # device to host
    ref = hA[:1].astype(np.complex128) @ hB.astype(np.complex128)
    err = float(np.abs(row - ref[0]).max() / np.abs(ref).max())
    for p in (A, B, Cm):
        hip.hipFree(p)
    avg = ms / 1000.0 / max(n, 1)
    print(json.dumps({"kernel": "gemm", "prec": args.prec, "m": args.m, "n": args.n, "k": args.k, "launches": n,
                      "avg_launch_ms": 1000.0 * avg, "wall_ms": 1000.0 * wall,
                      "tflops": work / max(n, 1) / avg / 1e12 if n else None,
                      "bytes_operands": (args.m * args.k + args.k * args.n + args.m * args.n) * esz,
                      "row0_rel_err": err}), flush=True)


if __name__ == "__main__":
    main()

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
    import torch
    dt = torch.complex128 if args.prec == "c128" else torch.complex64
    g = torch.Generator(device="cuda").manual_seed(1)
    A = torch.randn(args.m, args.k, dtype=dt, device="cuda", generator=g)
    B = torch.randn(args.k, args.n, dtype=dt, device="cuda", generator=g)
    Cm = torch.zeros(args.m, args.n, dtype=dt, device="cuda")
    torch.cuda.synchronize()
    f = L.bem_zgemm if args.prec == "c128" else L.bem_cgemm
    st = L.bh_stream()
    kid = "zgemm" if args.prec == "c128" else "cgemm"
    f(args.m, args.n, args.k, 1.0, 0.0, A.data_ptr(), args.k, B.data_ptr(), args.n, 0.0, 0.0, Cm.data_ptr(), args.n, st)
    bem.sync()
    bem.prof_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.reps):
        f(args.m, args.n, args.k, 1.0, 0.0, A.data_ptr(), args.k, B.data_ptr(), args.n, 0.0, 0.0, Cm.data_ptr(),
          args.n, st)
    bem.sync()
    wall = (time.perf_counter() - t0) / args.reps
    ms, work, n = bem.prof_read(kid)
    bem.prof_enable(False)
    # spot check of one row against torch
    ref = (A[:1] @ B)
    err = float((Cm[:1] - ref).abs().max() / ref.abs().max())
    avg = ms / 1000.0 / max(n, 1)
    print(json.dumps({"kernel": "gemm", "prec": args.prec, "m": args.m, "n": args.n, "k": args.k, "launches": n,
                      "avg_launch_ms": 1000.0 * avg, "wall_ms": 1000.0 * wall,
                      "tflops": work / max(n, 1) / avg / 1e12 if n else None,
                      "bytes_operands": (args.m * args.k + args.k * args.n + args.m * args.n) * A.element_size(),
                      "row0_rel_err_vs_torch": err}), flush=True)


if __name__ == "__main__":
    main()

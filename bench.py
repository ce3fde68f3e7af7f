"""Benchmark of the bem3_emf_b2 hot path on one MI355X.

Workload (one "step" = one configuration solved end to end, FP64 complex):
a dielectric sphere of radius 0.5 um (n = 1.59) in water (n = 1.33), surface
meshed as a level-4 icosphere (5120 triangles, 15360 element nodes), one
x-polarised plane wave of vacuum wavelength 1.064 um along +z (E0 = 1):
  coefficient-matrix assembly of both domains (G, H, dG, dH, dF)
  + factorisation of the boundary-integral system
  + boundary solve (potentials, boundary E/H and their normal derivatives)
  + E/H at 1000 seeded points in a 4x4x4 um box (7-point rule)
  + radiation force and torque (4-point and 7-point rules).
The mesh and beam are uploaded once; every step recomputes everything above
from the resident geometry.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published peaks
PEAK_FP64_TFLOPS = 78.6     # FP64 vector and FP64 matrix (dense)
PEAK_HBM_GBS = 8000.0

METRIC = "assembly+solve+coupling s/config vs plain FP64 GPU at 1e-6 field error"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(level):
    from bem3_emf_b2_amd import meshgen as mg
    V, F = mg.icosphere(level, 0.5)
    arrays = mg.paired_arrays(V, F)
    beams = mg.xpol_plane_wave(lambda0=1.064, n0=1.33, E0=1.0)
    return arrays, beams


def points(n, seed=20240601):
    import numpy as np
    rng = np.random.default_rng(seed)
    return rng.uniform(-2.0, 2.0, size=(n, 3))


def gpu_step(bem, obj, pts, form="fp64"):
    """one configuration end to end; form "fp64" is the plain FP64 form,
    "mixed" the mixed-precision form (FP32 far-element quadrature and
    boundary field sums, complex-FP32 factorisation refined in FP64)"""
    obj.set_precision("mixed" if form == "mixed" else "fp64")
    obj.assemble(2)
    obj.factor(0)
    bem.mo_solve([obj])
    fld, _ = bem.mo_field([obj], pts, "EH", "t", 1)
    FN = [bem.mo_force([obj], 0, typ, (0.0, 0.0, 0.0))[1:] for typ in (0, 1)]
    bem.sync()
    return fld, FN


def kernel_table(bem, steps):
    kern = {}
    for name in bem.KERNELS:
        ms, work, n = bem.prof_read(name)
        kern[name] = (ms, work, n)
    tot = sum(v[0] for v in kern.values())
    for name, (ms, work, n) in sorted(kern.items(), key=lambda kv: -kv[1][0]):
        if n:
            log("  %-11s %9.1f ms total  %6d launches  %5.1f %% of tracked" % (name, ms, n, 100.0 * ms / max(tot, 1e-9)))
    return kern


def cpu_baseline(arrays, beams, npts, target_s):
    """The oracle (a C/OpenMP restatement of the reference's algorithms: the
    same quadrature rules, near-pair switch and dense 8NN system) timed on a
    bounded sample of the same workload and scaled to seconds per
    configuration:
      assembly: rows of the first T elements of both domains (all columns),
                scaled by Ne / T;
      dense solve: LU of an n_s x n_s complex system, scaled by flops to the
                reference's zgetrf + zgetri of the 8NN matrix (8 na^3) plus
                its two per-domain zgesv (2 x 8/3 NN^3);
      fields: 64 of the points, scaled to all of them;
      force: boundary fields at 32 element centroids, scaled to the
                Ne + 7 Ne boundary points of the two rules (the 4-point
                rule takes three of its points from the stored node values)."""
    import numpy as np

    import oracle as orc
    rn, ed, md, sd, gd = arrays
    o = orc.Obj.from_arrays(*arrays, [0j, 1.59])
    o.setup(beams)
    Ne = o.dNe[0]
    NN = 3 * Ne
    na = 8 * NN
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    # assembly sample: grow T until the sample takes about a third of the budget
    T = 8
    while True:
        t0 = time.perf_counter()
        for d in (0, 1):
            o.cmatrix_rows(d, 1, 1 + T)
        ta = time.perf_counter() - t0
        if ta > target_s / 3 or T >= Ne:
            break
        T = min(Ne, int(T * max(2.0, 0.9 * (target_s / 3) / max(ta, 1e-3))))
    t_asm = ta * Ne / T
    # dense LU sample
    ns = 1536
    rng = np.random.default_rng(7)
    A = rng.standard_normal((ns, ns)) + 1j * rng.standard_normal((ns, ns)) + ns * np.eye(ns)
    B = rng.standard_normal((ns, 1)) + 0j
    t0 = time.perf_counter()
    orc.zgesv(A, B)
    tl = time.perf_counter() - t0
    rate = (8.0 / 3.0) * ns ** 3 / tl
    t_lu = (8.0 * na ** 3 + 2 * (8.0 / 3.0) * NN ** 3) / rate
    t_lu_block = 40.0 * NN ** 3 / rate    # the same elimination as the GPU's block form
    # the solved state is not needed for timing the field sums: reuse the
    # oracle's own boundary data layout with zero values
    P = points(npts)[:64]
    t0 = time.perf_counter()
    orc.mo_field([o], P, "EH", "t", 1)
    t_fld = (time.perf_counter() - t0) * npts / len(P)
    t0 = time.perf_counter()
    for t in range(1, 33):
        orc.mo_EH_bd([o], 0, 0, t, 0.0, 0.0, 2, "t")
    t_bd = (time.perf_counter() - t0) / 32
    t_force = t_bd * 8 * Ne
    o.free()
    total = t_asm + t_lu + t_fld + t_force
    return {"value": total, "unit": "s/config", "cores": cores, "kind": "port",
            "value_block_form": t_asm + t_lu_block + t_fld + t_force,
            "sample": ("oracle on %d host threads: assembly rows of %d/%d elements x 2 domains (%.1f s, x%.0f), "
                       "LU n=%d (%.1f s, %.1f GFLOP/s, scaled to 8 na^3 + 2x8/3 NN^3 at na=%d), fields 64/%d points, "
                       "boundary fields 32 centroids (scaled to 8 Ne); extrapolated: assembly %.0f s, dense solve "
                       "%.0f s, fields %.1f s, force %.0f s"
                       % (cores, T, Ne, ta, Ne / T, ns, tl, rate / 1e9, na, npts, t_asm, t_lu, t_fld, t_force))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--level", type=int, default=4, help="icosphere level (4: 5120 triangles)")
    ap.add_argument("--points", type=int, default=1000)
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU-baseline sample budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mixed-steps", type=int, default=2,
                    help="timed steps of the mixed-precision form after the FP64 ones (0: skip)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a device copy of 1 GiB (scale of the HBM byte counters)")
    ap.add_argument("--scene16", action="store_true",
                    help="only time the 16-object uneven scene (both forms, one configuration each)")
    ap.add_argument("--scene16-forms", default="fp64,mixed",
                    help="forms of the 16-object scene to run (the comparison needs both)")
    ap.add_argument("--scene16-out", default=None, help="one form alone: also save its fields and forces (.npz)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        x = torch.ones(1 << 26, dtype=torch.complex128, device="cuda")   # 1 GiB
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 16, "copy_bytes_written": y.numel() * 16, "launches": 3}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import bem3_emf_b2_amd as bem
    if args.scene16:
        forms = args.scene16_forms.split(",")
        if set(forms) == {"fp64", "mixed"}:
            scene16_compare()      # each form in a fresh child process
        else:
            bem.init(0)
            scene16_main(bem, forms, args.scene16_out)
        return
    bem.init(0)
    arrays, beams = workload(args.level)
    rn, ed, md, sd, gd = arrays
    obj = bem.Object(rn, ed, md, sd, [beams[0][1], 1.59], beams[0][0])
    obj.set_beams(beams)
    pts = points(args.points)
    Ne, NN = obj.Ne, obj.NN
    log("workload: icosphere level %d, Ne=%d, NN=%d" % (args.level, Ne, NN))

    for w in range(args.warmup):
        t0 = time.perf_counter()
        gpu_step(bem, obj, pts)
        log("warmup %d: %.3f s" % (w, time.perf_counter() - t0))

    bem.prof_enable(True)
    try:
        import torch
        torch.cuda.synchronize()
    except Exception:
        torch = None
    bem.sync()
    t0 = time.perf_counter()
    for s in range(args.steps):
        ref = gpu_step(bem, obj, pts)
        log("step %d done (%.3f s)" % (s, time.perf_counter() - t0))
    bem.sync()
    if torch is not None:
        torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = 1000.0 * dt / args.steps

    # kernel table from the HIP-event registry
    kern = kernel_table(bem, args.steps)
    bem.prof_enable(False)
    # dominant kernel: the tracked kernel with the largest summed time
    dom = max(kern, key=lambda k: kern[k][0])
    roof = roofline(dom, kern[dom])

    # the mixed-precision form on the same workload, its error against the
    # FP64 form's fields and force, and its own dominant kernel
    mixed = None
    if args.mixed_steps > 0:
        gpu_step(bem, obj, pts, "mixed")
        bem.prof_enable(True)
        bem.sync()
        t0 = time.perf_counter()
        for s in range(args.mixed_steps):
            got = gpu_step(bem, obj, pts, "mixed")
            log("mixed step %d done (%.3f s)" % (s, time.perf_counter() - t0))
        bem.sync()
        dtm = (time.perf_counter() - t0) / args.mixed_steps
        kmix = kernel_table(bem, args.mixed_steps)
        bem.prof_enable(False)
        it, res = obj.solve_stats()
        fe = float(np.abs(got[0] - ref[0]).max() / np.abs(ref[0]).max())
        Fe = max(float(np.abs(got[1][t][0] - ref[1][t][0]).max() / np.abs(ref[1][t][0]).max()) for t in (0, 1))
        Ne_ = max(float(np.abs(got[1][t][1] - ref[1][t][1]).max() / max(np.abs(ref[1][t][1]).max(), np.abs(ref[1][t][0]).max()))
                  for t in (0, 1))
        dmx = max(kmix, key=lambda k: kmix[k][0])
        mixed = {"value": dtm, "unit": "s/config", "speedup_vs_fp64": (ms_step / 1000.0) / dtm,
                 "triangles_per_s": Ne / dtm,
                 "refinement_steps": it, "relative_residual": res,
                 "rel_err_vs_fp64": {"fields": fe, "force": Fe, "torque": Ne_},
                 "kernels_ms_per_step": {k: v[0] / args.mixed_steps for k, v in kmix.items() if v[2]},
                 "roofline": roofline(dmx, kmix[dmx])}
        log("mixed form: %.3f s/config (x%.2f), %d refinement steps, field error %.2e" % (dtm, mixed["speedup_vs_fp64"], it, fe))

    cpu = None
    if not args.no_cpu:
        os.environ.setdefault("OMP_NUM_THREADS", str(min(16, os.cpu_count() or 1)))
        t0 = time.perf_counter()
        cpu = cpu_baseline(arrays, beams, args.points, args.cpu_seconds)
        log("cpu baseline sample took %.1f s" % (time.perf_counter() - t0))

    out = {
        "metric": METRIC,
        "value": ms_step / 1000.0,
        "unit": "s/config",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {"workload": "sphere r=0.5um n=1.59 in water, icosphere L%d (%d triangles), x-pol plane wave "
                               "1.064um; plain FP64 form: assembly + factor + solve + boundary E/H + E/H at %d "
                               "points + force/torque" % (args.level, Ne, args.points),
                   "triangles": Ne, "element_nodes": NN, "points": args.points, "precision": "fp64"},
        "triangles_per_s": Ne / (ms_step / 1000.0),
        "kernels_ms_per_step": {k: v[0] / args.steps for k, v in kern.items() if v[2]},
        "roofline": roof,
        "cpu_baseline": cpu,
        "mixed_form": mixed,
    }
    print(json.dumps(out), flush=True)


def scene16_run(bem, objs, beams, pts, form):
    """one configuration of the 16-object scene in one form: assembly and
    factorisation of every object, then the coupling loop to a 1e-8 change
    (timed); fields at the given points and force/torque of every object
    afterwards (untimed, for the comparison of the forms)"""
    from bem3_emf_b2_amd import meshgen as mg
    G = []
    for o in objs:
        rn, ed, md, sd, gd = mg.paired_arrays(o["V"] + o["center"], o["F"])
        g = bem.Object(rn, ed, md, sd, [beams[0][1], o["n"]], beams[0][0])
        g.set_precision(form)
        g.set_beams(beams)
        G.append(g)
    bem.prof_enable(True)
    bem.sync()
    t0 = time.perf_counter()
    for i, g in enumerate(G):
        g.assemble(2)
        g.factor(0)
        bem.sync()
        log("%s: object %d (%d triangles) assembled and factorised at %.1f s" % (form, i, g.Ne, time.perf_counter() - t0))
    t1 = time.perf_counter()
    it, tr = bem.mo_solve(G, ccd=1e-8)
    bem.sync()
    t2 = time.perf_counter()
    kern = kernel_table(bem, 1)
    bem.prof_enable(False)
    log("%s: coupling loop %d iterations, %.1f s" % (form, it, t2 - t1))
    fld, oid = bem.mo_field(G, pts, "EH", "t", 1)
    FN = [bem.mo_force(G, m, 0, tuple(objs[m]["center"]))[1:] for m in range(len(G))]
    for g in G:
        g.close()
    return {"s": t2 - t0, "objects_s": t1 - t0, "coupling_s": t2 - t1, "iterations": it, "fld": fld, "oid": oid,
            "FN": FN, "kernels_ms": {k: v[0] for k, v in kern.items() if v[2]}}


def scene16_main(bem, forms, out_npz=None):
    """the batching configuration: 16 objects (seeded spheres, cubes and
    cones, 0.3-1.5 um, 500-5000 triangles, n 1.4-1.7) in a 10 um cube of
    water, x-polarised 1.064 um plane wave, coupling to a 1e-8 change; the
    plain FP64 form and the mixed-precision form, one configuration each"""
    import numpy as np

    from bem3_emf_b2_amd import meshgen as mg
    os.environ["BEM3_HIP_PROGRESS"] = "1"
    objs = mg.scene16()
    S = sum(len(o["F"]) for o in objs)
    beams = mg.xpol_plane_wave(lambda0=1.064, n0=1.33, E0=1.0)
    pts = np.random.default_rng(20240602).uniform(-5.0, 5.0, size=(200, 3))
    log("scene16: %d objects, %d triangles" % (len(objs), S))
    # one form (a child of scene16_compare, or a profiler run): its timing
    r = scene16_run(bem, objs, beams, pts, forms[0])
    if out_npz:
        np.savez(out_npz, fld=r["fld"], oid=r["oid"], F=np.array([f for f, n in r["FN"]]),
                 N=np.array([n for f, n in r["FN"]]))
    print(json.dumps({"form": forms[0], "s": r["s"], "objects_s": r["objects_s"], "coupling_s": r["coupling_s"],
                      "iterations": r["iterations"], "kernels_ms": r["kernels_ms"], "triangles": S,
                      "objects": len(objs)}), flush=True)


def scene16_compare():
    """both forms of the 16-object scene, each in a fresh child process (the
    device memory the first form freed does not weigh on the second), then
    the mixed form's errors against the FP64 form"""
    import subprocess
    import tempfile

    import numpy as np
    res = {}
    with tempfile.TemporaryDirectory() as tmp:
        for form in ("fp64", "mixed"):
            npz = os.path.join(tmp, form + ".npz")
            r = subprocess.run([sys.executable, "-u", os.path.abspath(__file__), "--scene16", "--scene16-forms", form,
                                "--scene16-out", npz], stdout=subprocess.PIPE, text=True)
            if r.returncode != 0:
                log("scene16: the %s form failed (exit %d)" % (form, r.returncode))
                sys.exit(r.returncode)
            line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
            d = json.loads(line)
            z = np.load(npz)
            d.update(fld=z["fld"], oid=z["oid"], FN=list(zip(z["F"], z["N"])))
            res[form] = d
    r64, rmx = res["fp64"], res["mixed"]
    S = r64["triangles"]
    keep = (r64["oid"] < 0) & (rmx["oid"] < 0)
    fe = float(np.abs(rmx["fld"][keep] - r64["fld"][keep]).max() / np.abs(r64["fld"][keep]).max())
    Fe = max(float(np.abs(a[0] - b[0]).max() / np.abs(b[0]).max()) for a, b in zip(rmx["FN"], r64["FN"]))
    Ne_ = max(float(np.abs(a[1] - b[1]).max() / max(np.abs(b[1]).max(), np.abs(b[0]).max()))
              for a, b in zip(rmx["FN"], r64["FN"]))
    out = {
        "metric": METRIC, "value": r64["s"], "unit": "s/config", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": 1000.0 * r64["s"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128", "data": "synthetic",
        "config": {"workload": "16-object uneven scene (seeded spheres, cubes, cones 0.3-1.5 um, 500-5000 triangles, "
                               "n 1.4-1.7) in a 10 um cube of water, x-pol plane wave 1.064 um, coupling to 1e-8; "
                               "plain FP64 form: assembly + factor of every object + coupling loop",
                   "objects": r64["objects"], "triangles": S, "precision": "fp64"},
        "fp64": {k: r64[k] for k in ("s", "objects_s", "coupling_s", "iterations", "kernels_ms")},
        "mixed_form": {"value": rmx["s"], "unit": "s/config", "speedup_vs_fp64": r64["s"] / rmx["s"],
                       **{k: rmx[k] for k in ("objects_s", "coupling_s", "iterations", "kernels_ms")},
                       "rel_err_vs_fp64": {"fields": fe, "force": Fe, "torque": Ne_, "field_points": int(keep.sum())}},
    }
    print(json.dumps(out), flush=True)


# FP64 operations per quadrature point of the far-pair kernels (counted on
# the gfx950 ISA of the point loop: v_fma/mul/add_f64 count 2/1/1, the
# sqrt/rcp/sin/cos sequences at their instruction counts)
FLOP_PER_POINT = {"asm_gl": 325.0, "field": 80.0, "scatter": 160.0}
PEAK_FP32_TFLOPS = 157.3    # FP32 vector and FP32 matrix (dense)
# k_gemm128<double> at 15360^3: FETCH_SIZE 323,492,022 KiB x 2.0 + WRITE_SIZE 3,830,400 KiB x 1.0
GEMM128_FP64_TRAFFIC = {"bytes": 323492022 * 1024 * 2.0 + 3830400 * 1024 * 1.0, "shape": "15360x15360x15360",
                        "algorithmic_bytes": 3 * 15360 ** 2 * 16}
# k_gemm128<float> at 15360^3: FETCH_SIZE 42,520,650 KiB x 2.0 + WRITE_SIZE 1,843,200 KiB x 1.0
GEMM128_FP32_TRAFFIC = {"bytes": 42520650 * 1024 * 2.0 + 1843200 * 1024 * 1.0, "shape": "15360x15360x15360",
                        "algorithmic_bytes": 3 * 15360 ** 2 * 8}


def roofline(dom, rec):
    """roofline entry of one tracked kernel from its (ms, work, launches)"""
    ms, work, n = rec
    avg_s = ms / 1000.0 / max(n, 1)
    if dom in ("zgemm", "cgemm"):
        peak = PEAK_FP64_TFLOPS if dom == "zgemm" else PEAK_FP32_TFLOPS
        ach = work / n / avg_s / 1e12
        # HBM bytes per launch from FETCH_SIZE / WRITE_SIZE counter runs of the
        # kernel alone at 15360^3 (the step's largest product), scaled by the
        # copy-kernel calibration (profiles/gemm128_fp64_15360_counters.txt,
        # profiles/gemm128_fp32_15360_counters.txt)
        t = GEMM128_FP64_TRAFFIC if dom == "zgemm" else GEMM128_FP32_TRAFFIC
        return {"kernel": "k_gemm128<%s>" % ("double" if dom == "zgemm" else "float"), "bound": "mfma",
                "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                "traffic": t["bytes"] if t else None, "traffic_shape": t["shape"] if t else None,
                "traffic_algorithmic_bytes": t["algorithmic_bytes"] if t else None,
                "avg_launch_ms": 1000.0 * avg_s, "work_per_launch": work / n,
                "work_unit": "real %s flops (8 m n k)" % ("FP64" if dom == "zgemm" else "FP32")}
    if dom in FLOP_PER_POINT:
        # far-pair quadrature: FP64 VALU; flops per (row node, element) pair
        # = 432 points x FLOP_PER_POINT (ISA count of the kernel's point loop)
        fpp = 432 * FLOP_PER_POINT[dom]
        ach = work / n * fpp / avg_s / 1e12
        return {"kernel": "k_" + dom, "bound": "valu-fp64", "achieved": ach, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": ach / PEAK_FP64_TFLOPS, "traffic": None,
                "avg_launch_ms": 1000.0 * avg_s, "work_per_launch": work / n, "work_unit": "pairs"}
    return {"kernel": dom, "bound": "mfma", "achieved": None, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": None, "traffic": None, "avg_launch_ms": 1000.0 * avg_s}

if __name__ == "__main__":
    main()

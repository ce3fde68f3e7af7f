"""Benchmark of the bem3_emf_b2 hot path on one MI355X.

Default workload (one "step" = one configuration solved end to end): the
16-object uneven scene -- seeded spheres, cubes and cones of size
0.3-1.5 um with 780-4800 triangles each (36648 in all), refractive index
1.4-1.7, in a 10 um cube of water (n = 1.33), one x-polarised 1.064 um plane
wave along +z.  A step is, for every object, coefficient-matrix assembly of
both domains and factorisation of its boundary-integral system, then the
mo_solve coupling loop to a 1e-8 change of the boundary values (with the
boundary E/H and their normal derivatives at the end).  The geometry and the
beam are uploaded once; every step recomputes everything above.

`value` is the mixed-precision form's seconds per configuration (far-element
rules and series coupling sums in FP64, FP32 boundary-field sums,
complex-FP32 block elimination on MFMA refined in FP64 to a 1e-10 relative
residual).  The same line carries the plain FP64 form on the same scene
(one configuration), the mixed form's relative errors against it (fields at
200 points outside the objects, force and torque of every object), the
dominant kernel's roofline, a CPU baseline of the reference algorithm on the
host cores (run while the FP64 form occupies the GPU), the single
level-4 sphere (5120 triangles) in both forms, four objects under 64
plane waves solved together as right-hand sides (force and torque per
beam, far-field intensity grid per beam), and the level-5 sphere (20480
triangles) in the low-memory mixed form.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published peaks (dense)
PEAK_FP64_TFLOPS = 78.6     # FP64 vector and FP64 matrix
PEAK_FP32_TFLOPS = 157.3    # FP32 matrix
PEAK_BF16_TFLOPS = 2500.0   # bf16 matrix (dense)
PEAK_HBM_GBS = 8000.0

METRIC = "assembly+solve+coupling s/config vs plain FP64 GPU at 1e-6 field error"
SCENE_SEED = 2024


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ 16-object scene
def scene_spec():
    from bem3_emf_b2_amd import meshgen as mg
    objs = mg.scene16(seed=SCENE_SEED)
    beams = mg.xpol_plane_wave(lambda0=1.064, n0=1.33, E0=1.0)
    return objs, beams


def scene_objects(bem, objs, beams, form):
    """device objects of the scene (geometry and beam uploaded once)"""
    from bem3_emf_b2_amd import meshgen as mg
    G = []
    for o in objs:
        rn, ed, md, sd, gd = mg.paired_arrays(o["V"] + o["center"], o["F"])
        g = bem.Object(rn, ed, md, sd, [beams[0][1], o["n"]], beams[0][0])
        g.set_precision(form)
        g.set_beams(beams)
        G.append(g)
    return G


def scene_step(bem, G, phases=None):
    """one configuration: assembly and factorisation of every object, the
    coupling loop to a 1e-8 change (boundary E/H included); phases (a list)
    receives the seconds of the two parts"""
    t0 = time.perf_counter()
    bem.mo_prepare(G, 2, "auto")
    bem.sync()
    t1 = time.perf_counter()
    it, _ = bem.mo_solve(G, ccd=1e-8)
    bem.sync()
    if phases is not None:
        phases.append((t1 - t0, time.perf_counter() - t1))
    return it


def scene_points():
    import numpy as np
    return np.random.default_rng(20240602).uniform(-5.0, 5.0, size=(200, 3))


def scene_outputs(bem, G, objs, pts):
    fld, oid = bem.mo_field(G, pts, "EH", "t", 1)
    FN = [bem.mo_force(G, m, 0, tuple(objs[m]["center"]))[1:] for m in range(len(G))]
    return fld, oid, FN


def rel_errors(fld, oid, FN, fld0, oid0, FN0):
    import numpy as np
    keep = (oid < 0) & (oid0 < 0)
    fe = float(np.abs(fld[keep] - fld0[keep]).max() / np.abs(fld0[keep]).max())
    Fe = max(float(np.abs(a[0] - b[0]).max() / np.abs(b[0]).max()) for a, b in zip(FN, FN0))
    Ne_ = max(float(np.abs(a[1] - b[1]).max() / max(np.abs(b[1]).max(), np.abs(b[0]).max())) for a, b in zip(FN, FN0))
    return {"fields": fe, "force": Fe, "torque": Ne_, "field_points": int(keep.sum())}


def kernel_table(bem):
    """name -> (ms, work, launches, algorithmic bytes) summed over launches"""
    kern = {}
    for name in bem.KERNELS:
        ms, work, by, n = bem.prof_read_bytes(name)
        kern[name] = (ms, work, n, by)
    tot = sum(v[0] for v in kern.values())
    for name, (ms, work, n, by) in sorted(kern.items(), key=lambda kv: -kv[1][0]):
        if n:
            log("  %-13s %9.1f ms total  %6d launches  %5.1f %% of tracked  %8.1f TFLOP/s  %7.0f GB/s"
                % (name, ms, n, 100.0 * ms / max(tot, 1e-9), work / max(ms, 1e-9) / 1e9, by / max(ms, 1e-9) / 1e6))
    return kern


# ------------------------------------------------------------------ CPU baseline
def cpu_scene_baseline(target_s=20.0):
    """The oracle (a C/OpenMP restatement of the reference's algorithms: the
    same quadrature rules, near-pair switch, dense 8NN system and coupling
    sums) timed on bounded samples of the scene, scaled per phase:
      assembly: rows of the first T elements of both domains of one object
        (all columns), scaled by the scene's element pairs (sum 6 Ne^2);
      dense solve: the LU of an n_s x n_s complex system, scaled by flops to
        every object's zgetrf + zgetri of its 8NN matrix (8 na^3, na = 24 Ne)
        and its two per-domain zgesv (2 x 8/3 NN^3);
      coupling: add_scattered (mEMP2_s at every slot, mo_solve.c:296-396) of
        one object pair, scaled by the scene's (slot, element) pairs per
        sweep, 4 (S^2 - sum Ne^2), and by the sweeps of the measured loop;
      the per-sweep solves X = invA B: one na^2 complex mat-vec per object
        and sweep at the LU's rate (a lower bound: the reference streams
        invA from disk)."""
    import numpy as np

    import oracle as orc
    from bem3_emf_b2_amd import meshgen as mg
    objs, beams = scene_spec()
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    Ne = np.array([len(o["F"]) for o in objs], dtype=np.float64)
    S = Ne.sum()
    # assembly sample on the smallest object
    small = int(np.argmin(Ne))
    O = []
    for i in sorted(set([small, int(np.argsort(Ne)[1])])):
        o = objs[i]
        ob = orc.Obj.from_arrays(*mg.paired_arrays(o["V"] + o["center"], o["F"]), [0j, o["n"]])
        ob.setup(beams)
        O.append(ob)
    oa = O[0]
    T = 8
    while True:
        t0 = time.perf_counter()
        for d in (0, 1):
            oa.cmatrix_rows(d, 1, 1 + T)
        ta = time.perf_counter() - t0
        if ta > target_s / 4 or T >= oa.dNe[0]:
            break
        T = min(oa.dNe[0], int(T * max(2.0, 0.9 * (target_s / 4) / max(ta, 1e-3))))
    per_pair = ta / (2.0 * 3 * T * oa.dNe[0])          # (row node, element) pairs of 2 domains
    t_asm = per_pair * float((6.0 * Ne ** 2).sum())
    # dense LU sample
    ns = 1536
    rng = np.random.default_rng(7)
    A = rng.standard_normal((ns, ns)) + 1j * rng.standard_normal((ns, ns)) + ns * np.eye(ns)
    B = rng.standard_normal((ns, 1)) + 0j
    t0 = time.perf_counter()
    orc.zgesv(A, B)
    tl = time.perf_counter() - t0
    rate = (8.0 / 3.0) * ns ** 3 / tl
    na = 24.0 * Ne
    NN = 3.0 * Ne
    t_lu = float((8.0 * na ** 3 + 2 * (8.0 / 3.0) * NN ** 3).sum()) / rate
    t_lu_block = float((43.0 * NN ** 3).sum()) / rate
    # the same factorisation by numpy's LAPACK (an optimised BLAS: the closest
    # stand-in here for the reference's MKL zgetrf/zgetri)
    nl = 3072
    A = rng.standard_normal((nl, nl)) + 1j * rng.standard_normal((nl, nl)) + nl * np.eye(nl)
    B = rng.standard_normal((nl, 1)) + 0j
    t0 = time.perf_counter()
    np.linalg.solve(A, B)
    tll = time.perf_counter() - t0
    del A
    rate_lapack = (8.0 / 3.0) * nl ** 3 / tll
    # coupling sample: the field of O[1] at the slots of O[0]
    t0 = time.perf_counter()
    orc.add_scattered(O[0], O[1])
    tc = time.perf_counter() - t0
    pairs_sample = 4.0 * O[0].dNe[0] * O[1].dNe[0]
    per_cpair = tc / pairs_sample
    sweeps = 13     # the measured loop: a first solve, 11 sweeps of the coupling sums and solves, E/H
    cpairs = 4.0 * (S ** 2 - float((Ne ** 2).sum()))
    t_coup = per_cpair * cpairs * 11
    t_solve = float((8.0 * na ** 2).sum()) * sweeps / rate
    for ob in O:
        ob.free()
    total = t_asm + t_lu + t_coup + t_solve
    t_lu_l = float((8.0 * na ** 3 + 2 * (8.0 / 3.0) * NN ** 3).sum()) / rate_lapack
    t_solve_l = float((8.0 * na ** 2).sum()) * sweeps / rate_lapack
    return {"value": total, "unit": "s/config", "cores": cores, "nproc": os.cpu_count(), "kind": "port",
            "value_block_form": t_asm + t_lu_block + t_coup + float((8.0 * (8 * NN) ** 2).sum()) * sweeps / rate,
            "value_lapack_lu": t_asm + t_lu_l + t_coup + t_solve_l,
            "lapack_lu_sample": "numpy.linalg.solve n=%d complex128: %.2f s, %.1f GFLOP/s (the dense solve and the "
                                "per-sweep mat-vecs of value at this rate: %.0f s + %.0f s)"
                                % (nl, tll, rate_lapack / 1e9, t_lu_l, t_solve_l),
            "sample": ("oracle on %d host threads (nproc %s): assembly rows of %d/%d elements x 2 domains of one "
                       "object (%.1f s; %.3g s per pair, scaled to the scene's sum 6 Ne^2 = %.3g pairs), LU n=%d "
                       "(%.1f s, %.1f GFLOP/s, scaled to 8 na^3 + 2x8/3 NN^3 per object), add_scattered of one "
                       "object pair (%.0f pairs, %.1f s, scaled to 4 (S^2 - sum Ne^2) = %.3g pairs x 11 sweeps), "
                       "per-sweep invA mat-vecs at the LU rate; extrapolated: assembly %.0f s, dense solve %.0f s, "
                       "coupling %.0f s, solves %.0f s"
                       % (cores, os.cpu_count(), T, oa.dNe[0], ta, per_pair, float((6 * Ne ** 2).sum()), ns, tl,
                          rate / 1e9, pairs_sample, tc, cpairs, t_asm, t_lu, t_coup, t_solve))}


def start_cpu_baseline(target_s):
    """the CPU baseline in a child process (no GPU use): it runs on the host
    cores while the FP64 form runs on the GPU"""
    code = ("import json, sys; sys.path.insert(0, %r); import bench; "
            "print(json.dumps(bench.cpu_scene_baseline(%r)))" % (ROOT, target_s))
    # one host core stays with the process driving the GPU
    env = dict(os.environ)
    n = int(env.get("OMP_NUM_THREADS") or 0) or os.cpu_count() or 2
    env["OMP_NUM_THREADS"] = str(max(1, n - 1))
    return subprocess.Popen([sys.executable, "-c", code], stdout=subprocess.PIPE, text=True, env=env)


def finish_cpu_baseline(p):
    out, _ = p.communicate(timeout=600)
    if p.returncode != 0:
        log("cpu baseline failed (exit %d)" % p.returncode)
        return None
    return json.loads([l for l in out.splitlines() if l.startswith("{")][-1])


# ------------------------------------------------------------------ single L4 sphere
def l4_extra(bem, npts=1000):
    """the level-4 sphere (5120 triangles) end to end in both forms: assembly,
    factorisation, solve, boundary E/H, E/H at 1000 points, force/torque"""
    import numpy as np

    from bem3_emf_b2_amd import meshgen as mg
    V, F = mg.icosphere(4, 0.5)
    rn, ed, md, sd, gd = mg.paired_arrays(V, F)
    beams = mg.xpol_plane_wave(lambda0=1.064, n0=1.33, E0=1.0)
    pts = np.random.default_rng(20240601).uniform(-2.0, 2.0, size=(npts, 3))
    obj = bem.Object(rn, ed, md, sd, [beams[0][1], 1.59], beams[0][0])
    obj.set_beams(beams)
    out = {}
    res = {}
    for form in ("fp64", "mixed"):
        obj.set_precision(form)

        def step():
            obj.assemble(2)
            obj.factor(0)
            bem.mo_solve([obj])
            fld, _ = bem.mo_field([obj], pts, "EH", "t", 1)
            FN = [bem.mo_force([obj], 0, typ, (0.0, 0.0, 0.0))[1:] for typ in (0, 1)]
            bem.sync()
            return fld, FN
        step()
        t0 = time.perf_counter()
        res[form] = step()
        out[form] = time.perf_counter() - t0
        log("L4 sphere, %s form: %.3f s/config" % (form, out[form]))
    obj.close()
    (f64, FN64), (fmx, FNmx) = res["fp64"], res["mixed"]
    return {"workload": "sphere r=0.5um n=1.59 in water, icosphere L4 (5120 triangles), x-pol plane wave 1.064um: "
                        "assembly + factor + solve + boundary E/H + E/H at %d points + force/torque" % npts,
            "fp64_s": out["fp64"], "mixed_s": out["mixed"], "speedup": out["fp64"] / out["mixed"],
            "rel_err_vs_fp64": {"fields": float(np.abs(fmx - f64).max() / np.abs(f64).max()),
                                "force": max(float(np.abs(a[0] - b[0]).max() / np.abs(b[0]).max())
                                             for a, b in zip(FNmx, FN64))}}


def l5_extra(bem):
    """config 3: the level-5 sphere (20480 triangles, NN 61440) in the
    low-memory mixed form, one configuration end to end: assembly,
    factorisation, solve with boundary E/H, E/H at 1000 points (7-point
    rule) and force/torque (rule 0)"""
    import numpy as np

    from bem3_emf_b2_amd import meshgen as mg
    V, F = mg.icosphere(5, 0.5)
    rn, ed, md, sd, gd = mg.paired_arrays(V, F)
    beams = mg.xpol_plane_wave(lambda0=1.064, n0=1.33, E0=1.0)
    pts = np.random.default_rng(20240601).uniform(-2.0, 2.0, size=(1000, 3))
    g = bem.Object(rn, ed, md, sd, [beams[0][1], 1.59], beams[0][0])
    g.set_precision("mixed_lm")
    g.set_beams(beams)
    bem.sync()
    t = [time.perf_counter()]
    g.assemble(2)
    bem.sync()
    t.append(time.perf_counter())
    g.factor(0)
    bem.sync()
    t.append(time.perf_counter())
    bem.mo_solve([g])
    bem.sync()
    t.append(time.perf_counter())
    bem.mo_field([g], pts, "EH", "t", 1)
    bem.mo_force([g], 0, 0)
    bem.sync()
    t.append(time.perf_counter())
    steps, res = g.solve_stats()
    g.close()
    ph = [t[i + 1] - t[i] for i in range(4)]
    log("L5 sphere, low-memory mixed form: %.1f s (assembly %.1f, factorisation %.1f, solve + boundary E/H %.1f, "
        "fields + force %.1f; %d refinement steps, residual %.1e)" % (t[-1] - t[0], *ph, steps, res))
    nn2 = 61440.0 ** 2
    return {"workload": "sphere r=0.5um n=1.59 in water, icosphere L5 (20480 triangles, NN 61440), x-pol plane wave "
                        "1.064um, low-memory mixed form: assembly + factor + solve + boundary E/H + E/H at 1000 points + "
                        "force/torque",
            "s_total": t[-1] - t[0],
            "phases_s": {"assembly": ph[0], "factorisation": ph[1], "solve_boundary_EH": ph[2], "fields_force": ph[3]},
            "refinement_steps": steps, "fp64_relative_residual": res,
            "device_bytes_matrices": {"fp32_storage_and_factors": 6 * nn2 * 8, "fp64_row_blocks": 16 * 2 ** 30,
                                      "fp64_matrices_not_stored": 12 * nn2 * 16}}


def beams_extra(bem, nb=64):
    """four objects under nb plane waves as right-hand sides (seeded random
    directions and polarisation angles), end to end: assembly and
    factorisation of every object, the coupling loop of all beams together
    to a 1e-8 change with the boundary E/H, force and torque of every object
    under every beam (both rules), and the far-field intensity grid of
    180 x 360 points at 2000 wavelengths for every beam"""
    from bem3_emf_b2_amd import meshgen as mg
    beams = mg.random_plane_waves(nb, 64)
    objs = mg.scene_beams()
    G = []
    for o in objs:
        rn, ed, md, sd, gd = mg.paired_arrays(o["V"] + o["center"], o["F"])
        g = bem.Object(rn, ed, md, sd, [1.33, o["n"]], 1.064)
        g.set_precision("mixed")
        g.set_beams(beams[:1])
        G.append(g)
    rc = [o["center"] for o in objs]
    pts = bem.far_field_grid(1.064, 2000.0, 180, 360)
    bem.sync()
    t = [time.perf_counter()]
    bem.mo_prepare(G, 2, "auto")
    bem.sync()
    t.append(time.perf_counter())
    it = bem.mo_solve_beams(G, beams, ccd=1e-8)
    bem.sync()
    t.append(time.perf_counter())
    for typ in (0, 1):
        bem.mo_force_beams(G, typ, rc)
    bem.sync()
    t.append(time.perf_counter())
    bem.mo_far_field_beams(G, pts, 1, "I")
    bem.sync()
    t.append(time.perf_counter())
    for g in G:
        g.close()
    ph = [t[i + 1] - t[i] for i in range(4)]
    log("%d beams: %.2f s (objects %.2f, coupled solve %.2f in %d iterations, force %.2f, far field %.2f)"
        % (nb, t[-1] - t[0], *ph[:2], it, *ph[2:]))
    return {"workload": "4 objects (two 2028-triangle spheres r=1um, 3072-triangle cube 1.5um, 1520-triangle cone; "
                        "n=1.59 in water) under %d seeded random plane waves 1.064 um as right-hand sides: assembly + "
                        "factor + coupled solve of all beams to a 1e-8 change + boundary E/H + force/torque of every "
                        "object and beam (both rules) + 180x360 far-field intensity grid at 2000 wavelengths per beam"
                        % nb,
            "s_total": t[-1] - t[0], "s_per_beam": (t[-1] - t[0]) / nb, "iterations": it,
            "phases_s": {"objects": ph[0], "coupled_solve": ph[1], "force_torque": ph[2], "far_field": ph[3]}}


# ------------------------------------------------------------------ roofline
# k_cgemm_bf6 over the 16-object scene's configuration (every launch of every
# shape the mixed form's factorisations and solves take; one stream): HBM
# bytes per launch on average from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE,
# each scaled by the reading of a 1 GiB device copy
# (profiles/cgemm_bf6_scene16_counters.txt)
SCENE_GEMM_BF6_TRAFFIC = {"bytes": (173814.5 * 2.0 + 37164.5 * 1.0) * 1024,
                          "shape": "16-object scene with the shipped tuning table, mean over 7740 launches"}
# kernels whose registry work is real flops: (name, bound, peak TFLOP/s)
FLOP_KERNELS = {"zgemm": ("k_gemm128<double>", PEAK_FP64_TFLOPS), "cgemm": ("k_cgemm_bf6", PEAK_BF16_TFLOPS),
                "zgemm_small": ("k_gemm64<double>", PEAK_FP64_TFLOPS), "cgemm_small": ("k_gemm64<float>", PEAK_FP32_TFLOPS),
                "zgemv": ("k_gemm_skinny<double>", PEAK_FP64_TFLOPS), "cgemv": ("k_gemm_skinny<float>", PEAK_FP32_TFLOPS),
                "trsm": ("triangular solves (FP64)", PEAK_FP64_TFLOPS), "trsm_c": ("triangular solves (FP32)", PEAK_FP32_TFLOPS),
                "lu_panel": ("k_panel_coop<double>", PEAK_FP64_TFLOPS), "lu_panel_c": ("k_panel_slot", PEAK_FP32_TFLOPS),
                "mp_moments": ("k_mp_moments (FP64 VALU)", PEAK_FP64_TFLOPS)}
# FP64 operations per quadrature point of the far-pair kernels (counted on
# the gfx950 ISA of the point loop: v_fma/mul/add_f64 count 2/1/1, the
# sqrt/rcp/sin/cos sequences at their instruction counts)
FLOP_PER_POINT = {"asm_gl": 325.0, "field": 80.0, "scatter_far": 160.0}


def roofline(dom, rec):
    """roofline entry of one tracked kernel from its (ms, work, launches)"""
    ms, work, n, by = rec
    avg_s = ms / 1000.0 / max(n, 1)
    if dom == "cgemm":
        # complex FP32 product in the 3M form, each real product as six bf16
        # part products: the MFMA pipes execute 36 m n k flops per 8 m n k of
        # the complex product
        name, peak = FLOP_KERNELS[dom]
        alg = work / n / avg_s / 1e12
        ach = 4.5 * alg
        t = SCENE_GEMM_BF6_TRAFFIC
        return {"kernel": name, "bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                "traffic": t["bytes"] if t else None, "traffic_shape": t["shape"] if t else None,
                "algorithmic_bytes_per_launch": by / n if by else None,
                "avg_launch_ms": 1000.0 * avg_s, "work_per_launch": 4.5 * work / n,
                "work_unit": "bf16 MFMA flops (36 m n k: 3 real products of the 3M form x 6 bf16 part products)",
                "complex_fp32_tflops_8mnk": alg, "vs_fp32_matrix_peak": alg / PEAK_FP32_TFLOPS}
    if dom in FLOP_KERNELS:
        name, peak = FLOP_KERNELS[dom]
        ach = work / n / avg_s / 1e12
        return {"kernel": name, "bound": "mfma" if "VALU" not in name else "valu-fp64", "achieved": ach, "peak": peak,
                "unit": "TFLOP/s", "frac": ach / peak, "traffic": None,
                "algorithmic_bytes_per_launch": by / n if by else None,
                "avg_launch_ms": 1000.0 * avg_s, "work_per_launch": work / n, "work_unit": "real flops"}
    if dom in FLOP_PER_POINT:
        fpp = 432 * FLOP_PER_POINT[dom]
        ach = work / n * fpp / avg_s / 1e12
        return {"kernel": "k_" + dom, "bound": "valu-fp64", "achieved": ach, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": ach / PEAK_FP64_TFLOPS, "traffic": None,
                "avg_launch_ms": 1000.0 * avg_s, "work_per_launch": work / n, "work_unit": "pairs x 432 points"}
    return {"kernel": dom, "bound": None, "achieved": None, "peak": None, "unit": None, "frac": None, "traffic": None,
            "avg_launch_ms": 1000.0 * avg_s}


# ------------------------------------------------------------------ main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU-baseline sample budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparator (profiling runs)")
    ap.add_argument("--no-l4", action="store_true", help="skip the single L4 sphere")
    ap.add_argument("--no-beams", action="store_true", help="skip the 64-beam workload")
    ap.add_argument("--no-l5", action="store_true", help="skip the level-5 sphere")
    ap.add_argument("--lanes", type=int, default=0, help="device streams of the per-object work (0: default)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a device copy of 1 GiB (scale of the HBM byte counters)")
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
        print("bench.py measures one GPU: every GPU box here has one, and this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import bem3_emf_b2_amd as bem
    bem.init(0)
    objs, beams = scene_spec()
    S = sum(len(o["F"]) for o in objs)
    pts = scene_points()
    log("scene: %d objects, %d triangles" % (len(objs), S))
    G = scene_objects(bem, objs, beams, "mixed")
    for w in range(args.warmup):
        # the first warm-up runs on one stream: the work areas of the
        # kernel-timing configuration below exist before it
        bem.set_lanes(1 if w == 0 else args.lanes)
        t0 = time.perf_counter()
        scene_step(bem, G)
        log("warmup %d: %.3f s" % (w, time.perf_counter() - t0))
    bem.set_lanes(args.lanes)
    bem.sync()
    t0 = time.perf_counter()
    it = 0
    phases = []
    for s in range(args.steps):
        it = scene_step(bem, G, phases)
        log("step %d done (%.3f s; objects %.3f s, coupling loop %.3f s)" % (s, time.perf_counter() - t0, *phases[-1]))
    bem.sync()
    dt = time.perf_counter() - t0
    s_step = dt / args.steps
    series, direct, lmax = bem.plan_stats()
    split = bem.plan_split()
    # kernel timings: one more configuration with the per-object work on one
    # stream, so that the events around each launch bracket that kernel alone
    # (with several streams they would also count the other streams' kernels)
    bem.set_lanes(1)
    bem.prof_enable(True)
    os.environ["BEM3_HIP_G1_AHEAD"] = "0"   # the G1 factorisations on the same stream too
    t1 = time.perf_counter()
    scene_step(bem, G)
    s_serial = time.perf_counter() - t1
    del os.environ["BEM3_HIP_G1_AHEAD"]
    kern = kernel_table(bem)
    bem.prof_enable(False)
    bem.set_lanes(args.lanes)
    log("kernel-timing configuration (one stream): %.3f s" % s_serial)
    dom = max(kern, key=lambda k: kern[k][0])
    roof = roofline(dom, kern[dom])
    t1 = time.perf_counter()
    fld, oid, FN = scene_outputs(bem, G, objs, pts)
    bem.sync()
    s_outputs = time.perf_counter() - t1
    log("outputs (fields at %d points, force and torque of every object): %.3f s" % (len(pts), s_outputs))
    for g in G:
        g.close()

    # the plain FP64 form, one configuration, while the CPU baseline runs on the host cores
    cpu_proc = None if args.no_cpu else start_cpu_baseline(args.cpu_seconds)
    fp64 = None
    if not args.no_fp64:
        G64 = scene_objects(bem, objs, beams, "fp64")
        bem.sync()
        t1 = time.perf_counter()
        it64 = scene_step(bem, G64)
        t64 = time.perf_counter() - t1
        log("fp64 form: %.1f s/config (%d iterations)" % (t64, it64))
        t1 = time.perf_counter()
        f64, oid64, FN64 = scene_outputs(bem, G64, objs, pts)
        bem.sync()
        t64_out = time.perf_counter() - t1
        for g in G64:
            g.close()
        fp64 = {"value": t64, "unit": "s/config", "iterations": it64, "outputs_s": t64_out,
                "rel_err_of_mixed": rel_errors(fld, oid, FN, f64, oid64, FN64)}
    cpu = finish_cpu_baseline(cpu_proc) if cpu_proc else None
    l4 = None if args.no_l4 else l4_extra(bem)
    beams64 = None if args.no_beams else beams_extra(bem)
    l5 = None if args.no_l5 else l5_extra(bem)

    out = {
        "metric": METRIC,
        "value": s_step,
        "unit": "s/config",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * s_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+c128 (mixed)",
        "data": "synthetic",
        "config": {"workload": "16-object uneven scene (seeded spheres, cubes, cones 0.3-1.5 um, 780-4800 triangles, "
                               "n 1.4-1.7) in a 10 um cube of water, x-pol plane wave 1.064 um; assembly + factor of "
                               "every object + mo_solve coupling to a 1e-8 change; mixed-precision form",
                   "objects": len(objs), "triangles": S, "coupling_iterations": it,
                   "coupling_pairs": {"series": series, "element_sums": direct, "split": split,
                                      "largest_order": lmax}},
        "triangles_per_s": S / s_step,
        "phases_s": {"objects": sum(p[0] for p in phases) / len(phases),
                     "coupling_loop": sum(p[1] for p in phases) / len(phases)},
        "outputs_s": {"mixed": s_outputs, "fp64": fp64["outputs_s"] if fp64 else None,
                      "what": "E, H at 200 points outside the objects and force + torque of every object (rule 0), "
                              "after the configuration; not part of value"},
        "fp64_form": fp64,
        "speedup_vs_fp64": (fp64["value"] / s_step) if fp64 else None,
        "kernels_ms_per_step": {k: v[0] for k, v in kern.items() if v[2]},
        "kernel_timing_step": {"lanes": 1, "s": s_serial,
                               "note": "kernels_ms_per_step and roofline come from one extra configuration with the "
                                       "per-object work on one stream (events bracket one kernel each); value uses "
                                       "the default streams"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (cpu["value"] / s_step) if cpu else None,
        "l4_sphere": l4,
        "beams64": beams64,
        "l5_sphere": l5,
        "tuning": {"source": os.path.relpath(bem.tune_source(), ROOT) if os.path.isabs(bem.tune_source())
                   else bem.tune_source(), "entries": bem.tune_get()},
    }
    print(json.dumps(out), flush=True)
    bem.finalize()


if __name__ == "__main__":
    main()

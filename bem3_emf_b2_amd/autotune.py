"""Autotuner of the library's tuning table (include/bem3_hip.h, BH_TUNE_*).

It measures on the device, in this order:

1. the product kernels: at a grid of shapes (tile counts 9..256, k from 64
   to 2048) the 64x64-tile kernel against the 128x128-tile kernel in FP64,
   and in complex FP32 the FP32-MFMA against the bf16-split 128x128 kernel,
   then the 64x64 kernel against the faster of the two; the thresholds of
   the dispatch rule (the 128x128 kernel from `gemm_tiles_*` tiles with
   k >= 32, or from `gemm_longk_tiles` with k >= `gemm_longk_k`) are those
   that minimise the summed time over the grid;
2. the schedule of the multi-object configuration on the 16-object scene
   (mixed form, one x-polarised plane wave): assembly as one batch or per
   object, and the number of device streams, by the time of whole
   configurations (bh_mo_prepare + bh_mo_solve to a 1e-8 change).

The table is written as "name value" lines with the measurements as
comments; bh_init reads tune_gfx950.txt beside the library.

    python -m bem3_emf_b2_amd.autotune [--out PATH] [--quick]

--quick takes a few shapes and a 3-object scene (a test of the tuner itself).
"""
import argparse
import itertools
import os
import sys
import time

import bem3_emf_b2_amd as bem
from bem3_emf_b2_amd import meshgen as mg

DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune_gfx950.txt")
SCENE_SEED = 2024


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def tiles(m, n):
    return ((m + 127) // 128) * ((n + 127) // 128)


def rule_big(T, k, t1, t2, kl):
    return (T >= t1 and k >= 32) or (T >= t2 and k >= kl)


def fit_thresholds(rows, defaults):
    """rows: {prec: [(T, k, t_small, t_big)]}; defaults: {prec: t1} and the
    shared (t2, kl).  The per-precision tile threshold t1 and the shared
    long-k rule (t2 tiles, k >= kl) that minimise the summed time of the
    kernels the rule picks over both precisions' grids (ties keep the
    defaults).  Returns ({prec: t1}, (t2, kl), cost, cost of the defaults)."""
    Ts = sorted({r[0] for rs in rows.values() for r in rs})
    cand_t = sorted(set(Ts + [Ts[-1] + 1] + list(defaults[0].values()) + [defaults[1][0]]))
    cand_k = sorted({r[1] for rs in rows.values() for r in rs} | {defaults[1][1]})

    def cost(prec, t1, t2, kl):
        return sum(tb if rule_big(T, k, t1, t2, kl) else ts for T, k, ts, tb in rows[prec])

    def total(t1s, t2, kl):
        return sum(cost(p, t1s[p], t2, kl) for p in rows)

    c0 = total(defaults[0], *defaults[1])
    best = (c0, dict(defaults[0]), tuple(defaults[1]))
    for t2, kl in itertools.product(cand_t, cand_k):
        # the long-k rule only lowers a threshold: t1 >= t2
        t1s = {p: min((t for t in cand_t if t >= t2), key=lambda t: cost(p, t, t2, kl)) for p in rows}
        c = total(t1s, t2, kl)
        if c < best[0] * (1 - 1e-3):
            best = (c, t1s, (t2, kl))
    return best[1], best[2], best[0], c0


def tune_gemm(quick, notes):
    sizes = [384, 768, 1280] if quick else [384, 512, 640, 768, 1024, 1280, 1536, 1792, 2048]
    ks = [64, 1024] if quick else [64, 128, 256, 512, 1024, 2048]
    reps = 2 if quick else 5
    out = {}
    # complex FP32: the bf16-split against the FP32-MFMA 128x128 kernel where the 128x128 tiles run
    t_mfma = t_bf6 = 0.0
    for m in sizes[-3:]:
        for k in ks:
            t_mfma += bem.gemm_time("fp32", "128x128", m, m, k, reps)
            t_bf6 += bem.gemm_time("fp32", "bf16split", m, m, k, reps)
    out["gemm_bf6"] = 1 if t_bf6 < t_mfma else 0
    notes.append("complex-FP32 128x128 tiles over %d shapes: FP32-MFMA %.2f ms, bf16-split %.2f ms -> gemm_bf6 %d"
                 % (3 * len(ks), t_mfma, t_bf6, out["gemm_bf6"]))
    big32 = "bf16split" if out["gemm_bf6"] else "128x128"
    bem.tune_set("gemm_bf6", out["gemm_bf6"])
    keys = {"fp64": "gemm_tiles_fp64", "fp32": "gemm_tiles_fp32"}
    rows = {}
    for prec, big in (("fp64", "128x128"), ("fp32", big32)):
        rows[prec] = []
        for m in sizes:
            for k in ks:
                ts = bem.gemm_time(prec, "64x64", m, m, k, reps)
                tb = bem.gemm_time(prec, big, m, m, k, reps)
                rows[prec].append((tiles(m, m), k, ts, tb))
                log("autotune: %s %5d x %5d x %5d  64x64 %.4f ms  %s %.4f ms" % (prec, m, m, k, ts, big, tb))
    dflt = ({p: bem.tune_get(keys[p]) for p in keys}, (bem.tune_get("gemm_longk_tiles"), bem.tune_get("gemm_longk_k")))
    t1s, (t2, kl), c, c0 = fit_thresholds(rows, dflt)
    for p in keys:
        out[keys[p]] = t1s[p]
    out["gemm_longk_tiles"], out["gemm_longk_k"] = t2, kl
    notes.append("products over %d shapes per precision: 128x128 tiles from %d (FP64) / %d (FP32) tiles, with k >= %d "
                 "from %d: %.2f ms (the defaults %.2f ms)" % (len(rows["fp64"]), t1s["fp64"], t1s["fp32"], kl, t2, c, c0))
    for k_, v in out.items():
        bem.tune_set(k_, v)
    return out


def scene_device_objects(spec, beams):
    G = []
    for o in spec:
        rn, ed, md, sd, gd = mg.paired_arrays(o["V"] + o["center"], o["F"])
        g = bem.Object(rn, ed, md, sd, [beams[0][1], o["n"]], beams[0][0])
        g.set_precision("mixed")
        g.set_beams(beams)
        G.append(g)
    return G


def tune_schedule(quick, notes):
    spec = mg.scene16(seed=SCENE_SEED, n_obj=3 if quick else 16)
    beams = mg.xpol_plane_wave(lambda0=1.064, n0=1.33, E0=1.0)
    G = scene_device_objects(spec, beams)
    try:
        def config():
            t0 = time.perf_counter()
            bem.mo_prepare(G, 2, "auto")
            bem.sync()
            bem.mo_solve(G, ccd=1e-8)
            bem.sync()
            return time.perf_counter() - t0

        bem.set_lanes(0)
        config()   # warm-up: work areas, code objects
        lanes = [1, 2] if quick else [2, 4, 6, 8]
        best = None
        for batch in (1, 0):
            for nl in lanes:
                bem.tune_set("asm_batch", batch)
                bem.tune_set("lanes", nl)
                t = min(config() for _ in range(1 if quick else 2))
                log("autotune: scene asm_batch %d lanes %d: %.3f s/config" % (batch, nl, t))
                notes.append("scene (%d objects) asm_batch %d lanes %d: %.3f s/config" % (len(G), batch, nl, t))
                # the batched assembly (4 launches and one synchronisation for all
                # objects) unless the per-object form is faster by more than the
                # run-to-run spread of a configuration (2 %)
                if best is None or t < best[0] * (0.98 if (batch == 0 and best[1] == 1) else 1.0):
                    best = (t, batch, nl)
        bem.tune_set("asm_batch", best[1])
        bem.tune_set("lanes", best[2])
        return {"asm_batch": best[1], "lanes": best[2]}
    finally:
        for g in G:
            g.close()


def write_table(path, vals, notes):
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        f.write("# tuning table of libbem3hip.so (gfx950), written by bem3_emf_b2_amd/autotune.py\n")
        f.write("# measured %s\n" % time.strftime("%Y-%m-%d %H:%M:%S"))
        for n in notes:
            f.write("#   %s\n" % n)
        for k in bem.TUNE_KEYS:
            f.write("%s %d\n" % (k, vals.get(k, bem.tune_get(k))))
    os.replace(tmp, path)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--no-scene", action="store_true", help="products only")
    a = ap.parse_args(argv)
    bem.init(0)
    notes = []
    vals = tune_gemm(a.quick, notes)
    if not a.no_scene:
        vals.update(tune_schedule(a.quick, notes))
    write_table(a.out, vals, notes)
    log("autotune: wrote %s" % a.out)
    for n in notes:
        log("  " + n)
    return vals


if __name__ == "__main__":
    main()
    bem.finalize()

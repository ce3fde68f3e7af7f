"""Batched assembly and the tuning table on the device.

bh_mo_prepare assembles every object and domain as one batch (one launch per
assembly pass over all of them); each job's arithmetic is that of the object
assembled alone, so the matrices must agree bit for bit with per-object
assembly, for objects of different element counts, wavenumbers and element
types (the reference's quadrangle cone of verification1 among triangle
meshes; create_cmatrix_domain, solve_bieq2.c:98-184).  The schedule entries of
the tuning table (streams, batched or per-object assembly) leave the coupled
solution unchanged; the product-kernel thresholds change which kernel takes a
product, not its result beyond rounding."""
import os

import numpy as np
import pytest

from bem3_emf_b2_amd import meshgen as mg

pytestmark = pytest.mark.gpu

BEAM = np.array([[1.064, 1.33, 0.665, 1.0, 0.0, 0.3, 0.1, 0.0, 0.0, 0.0, 0.4, 0.2]])


def _tri_scene():
    return [
        (mg.icosphere(1, 0.3, center=(-0.9, 0.0, 0.0)), 1.45),
        (mg.box(0.5, 2, center=(0.15, 0.65, 0.05)), 1.60),
        (mg.cone(0.3, 0.6, 8, 2, center=(0.3, -0.65, 0.2)), 1.55 + 0.01j),
        (mg.icosphere(2, 0.25, center=(0.95, 0.1, -0.3)), 1.70),
    ]


def _objects(precs):
    """the verification1 quadrangle cone, then the triangle scene; precs[i]
    is object i's form"""
    import bem3_emf_b2_amd as bem
    from fixtures import v1_raw
    o = v1_raw()[0]
    G = [bem.Object(o["rn"], o["ed"], o["md"], o["sd"], list(o["n"]), float(o["lambda0"]))]
    for (V, F), n in _tri_scene():
        rn, ed, md, sd, gd = mg.paired_arrays(V, F)
        G.append(bem.Object(rn, ed, md, sd, [BEAM[0][1], n], BEAM[0][0]))
    for g, p in zip(G, precs):
        g.set_precision(p)
        g.set_beams(BEAM)
    return G


@pytest.fixture
def tune_restore():
    import bem3_emf_b2_amd as bem
    bem.init(0)
    saved = bem.tune_get()
    yield
    for k, v in saved.items():
        bem.tune_set(k, v)
    bem.set_lanes(0)


@pytest.mark.parametrize("precs", [["fp64"] * 5, ["mixed"] * 5, ["fp64", "mixed", "fp64", "mixed", "mixed"]])
def test_batched_assembly_equals_per_object(precs, tune_restore):
    import bem3_emf_b2_amd as bem
    print("batched assembly, forms", precs, flush=True)
    bem.tune_set("asm_batch", 1)
    A = _objects(precs)
    bem.mo_prepare(A, 2, "auto")
    B = _objects(precs)
    for g in B:
        g.assemble(2)
    try:
        for i, (a, b) in enumerate(zip(A, B)):
            for did in (0, 1):
                assert a.gh_count(did) == b.gh_count(did), (i, did)
                for key in ("G", "H", "dG", "dH", "dF"):
                    x, y = a.matrix(did, key), b.matrix(did, key)
                    assert np.array_equal(x, y), (i, did, key, np.abs(x - y).max())
    finally:
        for g in A + B:
            g.close()


def test_schedule_entries_leave_the_solution_unchanged(tune_restore):
    """one stream with per-object assembly against four streams with the
    batched assembly: the coupled mixed-form state of every object (the
    coupling loop to a 1e-8 change) agrees to the last bit"""
    import bem3_emf_b2_amd as bem
    states = []
    for lanes, batch in ((1, 0), (4, 1)):
        bem.tune_set("lanes", lanes)
        bem.tune_set("asm_batch", batch)
        bem.set_lanes(0)
        G = _objects(["mixed"] * 5)[1:]
        bem.mo_prepare(G, 2, "auto")
        it, _ = bem.mo_solve(G, ccd=1e-8)
        states.append((it, [g.state(0, "U") for g in G], [g.state(1, "dU") for g in G]))
        for g in G:
            g.close()
    (i0, u0, d0), (i1, u1, d1) = states
    assert i0 == i1
    for a, b in zip(u0 + d0, u1 + d1):
        assert np.array_equal(a, b), np.abs(a - b).max()


def test_product_thresholds(tune_restore):
    """the tile thresholds send one product to either kernel; both agree with
    numpy (FP64 to 1e-13, complex FP32 to its rounding), and so do the
    bf16-split and FP32-MFMA 128x128 kernels"""
    import bem3_emf_b2_amd as bem
    rng = np.random.default_rng(5)
    m, n, k = 700, 650, 300
    A = rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))
    Bm = rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))
    ref = A @ Bm
    for tiles in (1, 1 << 20):
        bem.tune_set("gemm_tiles_fp64", tiles)
        bem.tune_set("gemm_longk_tiles", tiles)
        got = bem.zgemm(A, Bm)
        assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-13, tiles
    ref32 = A.astype(np.complex64).astype(np.complex128) @ Bm.astype(np.complex64).astype(np.complex128)
    for tiles, bf6 in ((1, 1), (1, 0), (1 << 20, 1)):
        bem.tune_set("gemm_tiles_fp32", tiles)
        bem.tune_set("gemm_longk_tiles", tiles)
        bem.tune_set("gemm_bf6", bf6)
        got = bem.cgemm(A, Bm)
        assert np.abs(got - ref32).max() / np.abs(ref32).max() < 2e-6, (tiles, bf6)


def test_gemm_time_probe(tune_restore):
    import bem3_emf_b2_amd as bem
    for prec, kernels in (("fp64", ("64x64", "128x128")), ("fp32", ("64x64", "128x128", "bf16split"))):
        for kern in kernels:
            t = bem.gemm_time(prec, kern, 512, 384, 256, reps=2)
            assert 0.0 < t < 1000.0, (prec, kern, t)
    with pytest.raises(bem.BemError):
        bem.gemm_time("fp64", "bf16split", 256, 256, 256)


def test_autotuner_writes_a_loadable_table(tmp_path, tune_restore):
    """the autotuner's quick mode (a few product shapes, a 3-object scene)
    writes a table that bh_tune_load accepts and whose entries are in force"""
    import bem3_emf_b2_amd as bem
    from bem3_emf_b2_amd import autotune
    out = str(tmp_path / "tune.txt")
    vals = autotune.main(["--quick", "--out", out])
    assert os.path.exists(out)
    bem.tune_load(out)
    for k, v in vals.items():
        assert bem.tune_get(k) == v, k

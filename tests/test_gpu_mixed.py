"""Mixed-precision form on the device: the complex-FP32 GEMM / LU kernels
against numpy, and the FP32 block elimination refined in FP64 (factor mode
3) against the CPU oracle and against the device's FP64 block form.

Tolerances: FP32 kernels to a few units of FP32 round-off times the problem
size; refined solutions to 1e-8 / 1e-7 relative (refinement stops at a
relative residual of 1e-10, times the system's condition); with the
reduced far-pair rules and the FP32 boundary sums as well, fields, force and
torque to the 1e-6 relative tolerance the mixed form is specified for."""
import numpy as np
import pytest

import oracle as orc
from bem3_emf_b2_amd import meshgen as mg

pytestmark = pytest.mark.gpu

PTS = np.array([[0.0, 0.0, 0.0], [0.3, 0.1, 0.2], [1.0, 0.5, 0.3], [0.0, 0.0, 2.0], [-1.5, 0.7, -0.4]])


def _rand(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


# (3001, 2999, 77) and (4096, 4096, 64): 128x128-tile kernel; n <= 8 with
# k >= 256: the streaming kernel; 128x128 tiles (the bf16-split kernel) from
# 160 tiles, or 64 tiles with k >= 1024 (partial tiles and a k that is no
# multiple of the 16-value step included); the rest: 64x64 tiles
@pytest.mark.parametrize("m,n,k", [(64, 64, 16), (100, 70, 37), (3001, 2999, 77), (4096, 4096, 64), (3001, 1, 500),
                                   (1000, 6, 2048), (17, 8, 999), (257, 130, 64), (2049, 2050, 1031), (130, 9000, 2000),
                                   (1537, 1100, 33)])
def test_cgemm_matches_numpy(m, n, k):
    import bem3_emf_b2_amd as bem
    rng = np.random.default_rng(m * 1000 + n + k + 1)
    A, B, C0 = _rand(rng, m, k), _rand(rng, k, n), _rand(rng, m, n)
    al, be = -1.0 + 0.5j, 0.3 - 1.0j
    got = bem.cgemm(A, B, C0, alpha=al, beta=be)
    want = al * (A.astype(np.complex128) @ B.astype(np.complex128)) + be * C0.astype(np.complex128)
    # FP32 products and sums: error ~ sqrt(k) * eps32 relative to |A||B|
    scale = np.abs(A).max() * np.abs(B).max() * k
    assert np.abs(got - want).max() < 4e-7 * scale


@pytest.mark.parametrize("n,nrhs", [(1, 1), (63, 2), (300, 6), (1000, 3), (3100, 2)])
def test_cgesv_backward_stable(n, nrhs):
    import bem3_emf_b2_amd as bem
    rng = np.random.default_rng(n + 11 * nrhs)
    A = _rand(rng, n, n) + np.complex64(n * 0.1) * np.eye(n, dtype=np.complex64)
    B = _rand(rng, n, nrhs)
    X = bem.cgesv(A, B).astype(np.complex128)
    A64, B64 = A.astype(np.complex128), B.astype(np.complex128)
    res = np.linalg.norm(A64 @ X - B64) / (np.linalg.norm(A64) * np.linalg.norm(X))
    assert res < 2e-6
    want = np.linalg.solve(A64, B64)
    assert np.abs(X - want).max() < 1e-4 * np.abs(want).max() * max(1.0, np.log10(np.linalg.cond(A64)))


def _objects(spheres, beams):
    import bem3_emf_b2_amd as bem
    O, G = [], []
    for lvl, rad, n, c in spheres:
        V, F = mg.icosphere(lvl, rad, center=c)
        arrays = mg.paired_arrays(V, F)
        o = orc.Obj.from_arrays(*arrays, [0j, n])
        o.setup(beams)
        O.append(o)
        rn, ed, md, sd, gd = arrays
        g = bem.Object(rn, ed, md, sd, [beams[0][1], n], beams[0][0])
        g.assemble(2)
        g.set_beams(beams)
        G.append(g)
    return O, G


def test_mixed_single_object_vs_oracle():
    import bem3_emf_b2_amd as bem
    beams = mg.xpol_plane_wave()
    O, G = _objects([(1, 0.5, 1.59, (0, 0, 0))], beams)
    S = orc.System(O[0])
    orc.mo_solve(O, [S])
    G[0].factor("mixed")
    bem.mo_solve(G)
    it, res = G[0].solve_stats()
    assert 1 <= it <= 30 and res <= 1e-10, (it, res)
    for d in (0, 1):
        for key in ("U", "dU", "E", "dE", "H", "dH"):
            err = _rel(G[0].state(d, key), O[0].state(d, key))
            assert err < 1e-8, (d, key, err)
    for typ in (0, 1, 2, 3):
        want, _ = orc.mo_field(O, PTS, "EH", "t", typ)
        got, _ = bem.mo_field(G, PTS, "EH", "t", typ)
        assert _rel(got, want) < 1e-8, typ
    for typ in (0, 1):
        _, F0, N0 = orc.mo_force(O, 0, typ, (0.1, -0.2, 0.05))
        _, F1, N1 = bem.mo_force(G, 0, typ, (0.1, -0.2, 0.05))
        assert np.abs(F1 - F0).max() < 1e-8 * np.abs(F0).max()
        assert np.abs(N1 - N0).max() < 1e-8 * max(np.abs(N0).max(), np.abs(F0).max())


def test_mixed_matches_fp64_block_form_level3():
    """1280 triangles (NN = 3840): the refined FP32 solve against the FP64
    block form on the same device matrices (the FP64 form is checked
    against the oracle at level 1)."""
    import bem3_emf_b2_amd as bem
    beams = mg.xpol_plane_wave()
    V, F = mg.icosphere(3, 0.5)
    rn, ed, md, sd, gd = mg.paired_arrays(V, F)
    g = bem.Object(rn, ed, md, sd, [1.33, 1.59], 1.064)
    g.assemble(2)
    g.set_beams(beams)
    g.factor("block")
    bem.mo_solve([g])
    ref = {(d, k): g.state(d, k) for d in (0, 1) for k in ("U", "dU", "E", "dE", "H", "dH")}
    rng = np.random.default_rng(5)
    pts = rng.uniform(-2, 2, (64, 3))
    f_ref, _ = bem.mo_field([g], pts, "EH", "t", 1)
    _, F_ref, N_ref = bem.mo_force([g], 0, 1)
    g.factor("mixed")
    bem.mo_solve([g])
    it, res = g.solve_stats()
    assert res <= 1e-10
    # the refinement's 1e-10 residual times the condition of G, ~1e3 here
    for (d, k), v in ref.items():
        assert _rel(g.state(d, k), v) < 1e-7, (d, k)
    f, _ = bem.mo_field([g], pts, "EH", "t", 1)
    assert _rel(f, f_ref) < 1e-7
    _, F1, N1 = bem.mo_force([g], 0, 1)
    assert np.abs(F1 - F_ref).max() < 1e-7 * np.abs(F_ref).max()
    g.close()


def test_reduced_far_rule_matrices():
    """The mixed form's assembly (far pairs with the reduced 5x5 / 6x6 rules,
    FP64 arithmetic): every coefficient within 1e-8 of the FP64 assembly
    (relative to the matrix's largest entry; the rules' own error bound is
    5e-9 per pair), the same near-pair (GHN) set."""
    import bem3_emf_b2_amd as bem
    V, F = mg.icosphere(2, 0.5)
    rn, ed, md, sd, gd = mg.paired_arrays(V, F)
    g = bem.Object(rn, ed, md, sd, [1.33, 1.45 + 0.02j], 1.064)
    g.assemble(2)
    ref = {(d, w): g.matrix(d, w) for d in (0, 1) for w in ("G", "H", "dG", "dH", "dF")}
    nref = [g.gh_count(d) for d in (0, 1)]
    g.set_precision("mixed")
    g.assemble(2)
    assert [g.gh_count(d) for d in (0, 1)] == nref
    for (d, w), v in ref.items():
        got = g.matrix(d, w)
        assert _rel(got, v) < 1e-8, (d, w, _rel(got, v))
    g.close()


def test_mixed_force_on_low_contrast_cube():
    """A weakly scattering cube (n = 1.413 in water, 6 % contrast): its force
    comes from the scattered part of the field, so noise on the total field
    is magnified by the inverse contrast.  The mixed form (FP64 assembly
    with reduced far rules, FP32 factors refined in FP64, FP32 boundary
    sums) stays within 1e-6 of the FP64 form for force and torque."""
    import bem3_emf_b2_amd as bem
    beams = mg.xpol_plane_wave()
    V, F = mg.box(1.02, 8)
    rn, ed, md, sd, gd = mg.paired_arrays(V, F)
    res = {}
    for prec in ("fp64", "mixed"):
        g = bem.Object(rn, ed, md, sd, [1.33, 1.413], 1.064)
        g.set_precision(prec)
        g.set_beams(beams)
        g.assemble(2)
        g.factor(0)
        bem.mo_solve([g])
        res[prec] = [bem.mo_force([g], 0, t)[1:] for t in (0, 1)]
        g.close()
    for t in (0, 1):
        (F0, N0), (F1, N1) = res["fp64"][t], res["mixed"][t]
        assert np.abs(F1 - F0).max() < 1e-6 * np.abs(F0).max(), (t, np.abs(F1 - F0).max() / np.abs(F0).max())
        assert np.abs(N1 - N0).max() < 1e-6 * max(np.abs(N0).max(), np.abs(F0).max()), t


def test_mixed_form_end_to_end_vs_oracle():
    """The whole mixed-precision form (reduced far-pair rules, FP32 boundary
    field sums, FP32 block elimination refined in FP64) against the FP64
    oracle: fields, force and torque within the 1e-6 relative tolerance
    the form is specified for."""
    import bem3_emf_b2_amd as bem
    beams = mg.xpol_plane_wave()
    O, G = _objects([(2, 0.5, 1.59, (0, 0, 0))], beams)
    S = orc.System(O[0])
    orc.mo_solve(O, [S])
    g = G[0]
    g.set_precision("mixed")
    g.assemble(2)
    g.factor(0)
    bem.mo_solve(G)
    it, res = g.solve_stats()
    assert res <= 1e-10 and it >= 1
    rng = np.random.default_rng(20240601)
    pts = np.r_[PTS, rng.uniform(-2.0, 2.0, (40, 3))]
    for typ in (0, 1):
        want, _ = orc.mo_field(O, pts, "EH", "t", typ)
        got, _ = bem.mo_field(G, pts, "EH", "t", typ)
        assert _rel(got, want) < 1e-6, (typ, _rel(got, want))
    for typ in (0, 1):
        _, F0, N0 = orc.mo_force(O, 0, typ)
        _, F1, N1 = bem.mo_force(G, 0, typ)
        assert np.abs(F1 - F0).max() < 1e-6 * np.abs(F0).max(), typ
        assert np.abs(N1 - N0).max() < 1e-6 * max(np.abs(N0).max(), np.abs(F0).max()), typ


def test_mixed_two_object_coupling_loop():
    import bem3_emf_b2_amd as bem
    beams = np.array([[1.064, 1.33, 0.665, 1.0, 0.0, 0.3, 0.1, 0.0, 0.0, 0.0, 0.4, 0.2]])
    spec = [(1, 0.35, 1.59, (0.0, 0.0, 0.45)), (1, 0.3, 1.45 + 0.01j, (0.1, 0.0, -0.45))]
    O, G = _objects(spec, beams)
    S = [orc.System(o) for o in O]
    it_o, tr_o = orc.mo_solve(O, S, ccd=1e-8)
    for g in G:
        g.factor("mixed")
    it_g, tr_g = bem.mo_solve(G, ccd=1e-8)
    assert it_g == it_o
    assert np.abs(tr_g[2:it_o + 1] - tr_o[2:it_o + 1]).max() / tr_o[2] < 1e-6
    for m in range(2):
        for key in ("U", "E", "H"):
            assert _rel(G[m].state(0, key), O[m].state(0, key)) < 1e-8
    want, _ = orc.mo_field(O, PTS, "EH", "t", 1)
    got, _ = bem.mo_field(G, PTS, "EH", "t", 1)
    assert _rel(got, want) < 1e-8

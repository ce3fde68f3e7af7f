"""End-to-end device path vs the CPU oracle: factor/solve (dense 8NN form and
the one-medium block elimination), boundary E/H, fields, force, absorbed
power and the multi-object coupling loop.  FP64 throughout; tolerances are
stated per quantity (solver round-off amplified by the system's condition)."""
import numpy as np
import pytest

import oracle as orc
from bem3_emf_b2_amd import meshgen as mg

pytestmark = pytest.mark.gpu

PTS = np.array([[0.0, 0.0, 0.0], [0.3, 0.1, 0.2], [1.0, 0.5, 0.3], [0.0, 0.0, 2.0], [-1.5, 0.7, -0.4]])


def _objects(spheres, beams):
    """spheres: list of (level, radius, n, center)"""
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


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def single():
    beams = mg.xpol_plane_wave()
    O, G = _objects([(1, 0.5, 1.59, (0, 0, 0))], beams)
    S = orc.System(O[0])
    orc.mo_solve(O, [S])
    return O, G


@pytest.mark.parametrize("mode", [1, 2])
def test_single_object_boundary_values(single, mode):
    import bem3_emf_b2_amd as bem
    O, G = single
    G[0].factor(mode)
    bem.mo_solve(G)
    tol = 1e-10 if mode == 1 else 1e-9
    for d in (0, 1):
        for key in ("U", "dU", "E", "dE", "H", "dH"):
            err = _rel(G[0].state(d, key), O[0].state(d, key))
            assert err < tol, (mode, d, key, err)


def test_single_solve_uses_gmres_on_schur_complement(single):
    """one object, one solve: S is not formed, GMRES converges in a few
    dozen steps (cond(S) is about 2); the result equals the dense form's"""
    import bem3_emf_b2_amd as bem
    O, G = single
    G[0].factor(1)
    bem.mo_solve(G)
    ref = {k: G[0].state(0, k) for k in ("U", "dU")}
    G[0].factor(2)
    bem.mo_solve(G)
    _, _, gm = G[0].solve_stats(gmres=True)
    assert 3 <= gm <= 60, gm
    for k, v in ref.items():
        assert _rel(G[0].state(0, k), v) < 1e-11, k


@pytest.mark.parametrize("typ", [0, 1, 2, 3])
def test_single_object_fields(single, typ):
    import bem3_emf_b2_amd as bem
    O, G = single
    G[0].factor(2)
    bem.mo_solve(G)
    for mode in ("s", "t", "i"):
        want, oid_w = orc.mo_field(O, PTS, "EH", mode, typ)
        got, oid_g = bem.mo_field(G, PTS, "EH", mode, typ)
        assert (oid_w == oid_g).all()
        assert _rel(got, want) < 1e-9, (typ, mode)
    want, _ = orc.mo_field(O, PTS, "mEMP", "t", typ)
    got, _ = bem.mo_field(G, PTS, "mEMP", "t", typ)
    assert _rel(got, want) < 1e-9


@pytest.mark.parametrize("typ", [0, 1, 2, 3])
def test_gradient_field_form(single, typ):
    """E, H from the potentials' gradients (EH_mEMP_s): near points, points
    inside, and far-field points at 2000 wavelengths as the reference's
    far-field example uses them."""
    import bem3_emf_b2_amd as bem
    O, G = single
    G[0].factor(2)
    bem.mo_solve(G)
    th = np.linspace(0.1, 3.0, 7)
    far = 2000 * 1.064 * np.c_[np.sin(th) * np.cos(2 * th), np.sin(th) * np.sin(2 * th), np.cos(th)]
    pts = np.r_[PTS, far]
    for mode in ("s", "t", "i"):
        want, oid_w = orc.mo_field(O, pts, "EH_mEMP", mode, typ)
        got, oid_g = bem.mo_field(G, pts, "EH_mEMP", mode, typ)
        assert (oid_w == oid_g).all()
        for p in range(len(pts)):
            scale = max(np.abs(want[p]).max(), 1e-300)
            assert np.abs(got[p] - want[p]).max() < 1e-9 * scale, (typ, mode, p)


@pytest.mark.parametrize("typ", [0, 1])
def test_single_object_force_absorb(single, typ):
    import bem3_emf_b2_amd as bem
    O, G = single
    G[0].factor(2)
    bem.mo_solve(G)
    r0, F0, N0 = orc.mo_force(O, 0, typ, (0.1, -0.2, 0.05))
    r1, F1, N1 = bem.mo_force(G, 0, typ, (0.1, -0.2, 0.05))
    assert r0 == r1 == 0
    assert np.abs(F1 - F0).max() < 1e-9 * np.abs(F0).max()
    assert np.abs(N1 - N0).max() < 1e-9 * max(np.abs(N0).max(), np.abs(F0).max())
    a0, P0 = orc.mo_absorb(O, 0, typ)
    a1, P1 = bem.mo_absorb(G, 0, typ)
    assert a0 == a1
    assert abs(P1 - P0) < 1e-9 * abs(P0)


def test_two_object_coupling_loop():
    import bem3_emf_b2_amd as bem
    beams = np.array([[1.064, 1.33, 0.665, 1.0, 0.0, 0.3, 0.1, 0.0, 0.0, 0.0, 0.4, 0.2]])
    spec = [(1, 0.35, 1.59, (0.0, 0.0, 0.45)), (1, 0.3, 1.45 + 0.01j, (0.1, 0.0, -0.45))]
    O, G = _objects(spec, beams)
    S = [orc.System(o) for o in O]
    it_o, tr_o = orc.mo_solve(O, S, ccd=1e-8)
    for g in G:
        g.factor(2)
    it_g, tr_g = bem.mo_solve(G, ccd=1e-8)
    assert it_g == it_o
    assert all(g.solve_stats(gmres=True)[2] == 0 for g in G)   # the loop formed S
    assert np.abs(tr_g[2:it_o + 1] - tr_o[2:it_o + 1]).max() / tr_o[2] < 1e-6
    for m in range(2):
        for key in ("U", "E", "H"):
            assert _rel(G[m].state(0, key), O[m].state(0, key)) < 1e-8
    want, _ = orc.mo_field(O, PTS, "EH", "t", 1)
    got, _ = bem.mo_field(G, PTS, "EH", "t", 1)
    assert _rel(got, want) < 1e-8
    for m in range(2):
        _, F0, N0 = orc.mo_force(O, m, 0)
        _, F1, N1 = bem.mo_force(G, m, 0)
        assert np.abs(F1 - F0).max() < 1e-8 * np.abs(F0).max()

"""Many plane waves as separate right-hand sides (bh_mo_solve_beams) against
one coupling loop per beam (bh_mo_solve with that beam alone), both in the
mixed-precision form: three objects, two of them close enough for element
sums between them, four beams of seeded random directions and
polarisations.  Both loops are run to round-off (a change below 1e-24 or
40 iterations), so the boundary values U, dU, E, H of every domain, object
and beam agree to 1e-8, and so do the per-beam force and torque of every
object (bh_mo_force_beams against bh_mo_force of the one-beam run, both
rules) to 1e-6: the batched force forms the field integrals in FP64 with
the reduced far rules (bounded at 3e-8), the one-beam force of the mixed
form in FP32 for far pairs.  Against the oracle's mo_force_FN
(mo_force.c:10-240) with the same boundary values loaded, the batched
force agrees to 1e-7 for two of the beams.  The far field of every beam by the objects' series
(bh_mo_far_field_beams, 7-point rule) against the element sums of
EH_mEMP_s (bh_mo_field after bh_object_select_beam) at points 2000
wavelengths and 12 um away: 1e-8 relative."""
import numpy as np
import pytest

import oracle as orc
from bem3_emf_b2_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def _beams(nb, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nb):
        th, ph, psi = np.arccos(1 - 2 * rng.uniform()), 2 * np.pi * rng.uniform(), np.pi * rng.uniform()
        out.append([1.064, 1.33, 0.665, np.cos(psi), 0.0, np.sin(psi), 0.0, 0.0, 0.0, 0.0, th, ph])
    return np.array(out)


def _scene():
    return [
        (mg.icosphere(2, 0.35, center=(-1.4, 0.0, 0.0)), 1.5),
        (mg.box(0.5, 3, center=(1.2, 0.6, 0.2)), 1.6),
        (mg.icosphere(1, 0.2, center=(1.2, 0.6, 0.75)), 1.45),   # close to the cube: element sums
    ]


def _objects(beam):
    import bem3_emf_b2_amd as bem
    G = []
    for (V, F), n in _scene():
        rn, ed, md, sd, gd = mg.paired_arrays(V, F)
        g = bem.Object(rn, ed, md, sd, [1.33, n], 1.064)
        g.set_precision("mixed")
        g.set_beams(beam)
        g.assemble(2)
        g.factor(0)
        G.append(g)
    return G


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _far_points():
    rng = np.random.default_rng(12)
    u = rng.normal(size=(150, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    u[:100] *= 2000.0 * 1.064
    u[100:] *= 12.0
    return u


def test_many_beams_match_one_loop_per_beam():
    import bem3_emf_b2_amd as bem
    beams = _beams(4, 8)
    G = _objects(beams[:1])
    it = bem.mo_solve_beams(G, beams, ccd=1e-24, iter_max=40)
    series, direct, _ = bem.plan_stats()
    assert series >= 2 and direct >= 2
    got = [[{(d, k): g.beam_state(b, d, k) for d in (0, 1) for k in ("U", "dU", "E", "H")} for g in G]
           for b in range(len(beams))]
    rc = np.array([(-1.4, 0.0, 0.0), (1.2, 0.6, 0.2), (1.2, 0.6, 0.75)])   # the objects' centres
    FN = {t: bem.mo_force_beams(G, t, rc) for t in (0, 1)}
    # the oracle's force from the same boundary values, beams 0 and 2
    worst_o = 0.0
    for b in (0, 2):
        O = []
        for ((V, F), n), g in zip(_scene(), G):
            ob = orc.Obj.from_arrays(*mg.paired_arrays(V, F), [0j, n])
            ob.setup(beams[b:b + 1])
            for k in ("E", "dE", "H", "dH"):
                ob.state(0, k, g.beam_state(b, 0, k))
            O.append(ob)
        for m in range(len(O)):
            for t in (0, 1):
                _, F0, N0 = orc.mo_force(O, m, t, tuple(rc[m]))
                s = np.abs(F0).max()
                e = max(np.abs(FN[t][0][b, m] - F0).max() / s,
                        np.abs(FN[t][1][b, m] - N0).max() / max(np.abs(N0).max(), s))
                worst_o = max(worst_o, e)
                assert e < 1e-7, (b, m, t, e)
        for ob in O:
            ob.free()
    print("batched force vs the oracle's: %.2e" % worst_o, flush=True)
    pts = _far_points()
    far = bem.mo_far_field_beams(G, pts, 1, "EH")
    far_i = bem.mo_far_field_beams(G, pts, 1, "I")
    assert np.allclose(far_i[..., 0], (np.abs(far[..., :3]) ** 2).sum(-1), rtol=1e-13, atol=0)
    worst_far = 0.0
    for b in range(len(beams)):
        bem.select_beam(G, b)
        ref, oid = bem.mo_field(G, pts, "EH_mEMP", "s", 1)
        assert (oid < 0).all()
        e = max(_rel(far[b, :100], ref[:100]), _rel(far[b, 100:], ref[100:]))   # each radius on its own scale
        worst_far = max(worst_far, e)
        assert e < 1e-8, (b, e)
    print("far field by series vs element sums: %.2e" % worst_far, flush=True)
    for g in G:
        g.close()
    worst = worst_f = 0.0
    for b in range(len(beams)):
        G1 = _objects(beams[b:b + 1])
        bem.mo_solve(G1, ccd=1e-24, iter_max=40)
        for m, g in enumerate(G1):
            for d in (0, 1):
                for k in ("U", "dU", "E", "H"):
                    e = _rel(got[b][m][(d, k)], g.state(d, k))
                    worst = max(worst, e)
                    assert e < 1e-8, (b, m, d, k, e)
            for t in (0, 1):
                _, F0, N0 = bem.mo_force(G1, m, t, tuple(rc[m]))
                F1, N1 = FN[t][0][b, m], FN[t][1][b, m]
                assert FN[t][2][b, m] == 0
                s = np.abs(F0).max()
                e = max(np.abs(F1 - F0).max() / s, np.abs(N1 - N0).max() / max(np.abs(N0).max(), s))
                worst_f = max(worst_f, e)
                assert e < 1e-6, (b, m, t, e)
        for g in G1:
            g.close()
    print("many beams vs one loop per beam: states %.2e, force/torque %.2e (%d iterations)" % (worst, worst_f, it),
          flush=True)

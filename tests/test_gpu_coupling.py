"""Coupling sums of the mixed-precision form by per-object multipole series
(bem_mo.hip) against the element sums of the same form (the reduced far
rules of mEMP2_s, mo_solve.c:296-396): five objects, four of them far
enough apart for the series, one pair so close that it is split (the series
on the 64-slot tiles whose slots keep their distance, element sums on the
others).  The series reproduces the discrete sums to its truncation (about
1e-10 relative) plus the solid-angle terms it takes as zero (quadrature
level); the element sums of the mixed form add their pairs in FP32 (about
1e-8 relative on the scattered field, measured 1.4e-8 on the solved state),
so the solved state, fields outside the objects and forces agree to 1e-7
and the iteration counts are equal."""
import numpy as np
import pytest

from bem3_emf_b2_amd import meshgen as mg

pytestmark = pytest.mark.gpu

BEAM = np.array([[1.064, 1.33, 0.665, 1.0, 0.0, 0.3, 0.1, 0.0, 0.0, 0.0, 0.4, 0.2]])


def _scene():
    return [
        (mg.icosphere(2, 0.4, center=(-1.6, 0.0, 0.0)), 1.45),
        (mg.box(0.6, 3, center=(0.2, 1.5, 0.1)), 1.60),
        (mg.cone(0.35, 0.7, 12, 3, center=(0.4, -1.5, 0.3)), 1.55),
        (mg.icosphere(2, 0.3, center=(1.7, 0.2, -0.4)), 1.70),
        # close to the last sphere: the pair is split (series on the tiles of
        # slots that keep their distance, element sums on the others)
        (mg.icosphere(1, 0.2, center=(1.7, 0.2, 0.25)), 1.50),
    ]


def _solve(mode):
    import bem3_emf_b2_amd as bem
    bem.set_coupling(mode)
    G = []
    for (V, F), n in _scene():
        rn, ed, md, sd, gd = mg.paired_arrays(V, F)
        g = bem.Object(rn, ed, md, sd, [BEAM[0][1], n], BEAM[0][0])
        g.set_precision("mixed")
        g.assemble(2)
        g.set_beams(BEAM)
        g.factor(0)
        G.append(g)
    it, tr = bem.mo_solve(G, ccd=1e-8)
    stats = bem.plan_stats() + (bem.plan_split(),)
    U = [g.state(0, "U") for g in G]
    rng = np.random.default_rng(3)
    pts = rng.uniform(-2.5, 2.5, (64, 3))
    fld, oid = bem.mo_field(G, pts, "EH", "t", 1)
    FN = [bem.mo_force(G, m, 0)[1:] for m in range(len(G))]
    for g in G:
        g.close()
    bem.set_coupling("auto")
    return dict(it=it, tr=tr, stats=stats, U=U, fld=fld[oid < 0], FN=FN)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_series_coupling_matches_element_sums():
    ref = _solve("direct")
    got = _solve("auto")
    series, direct, lmax, split = got["stats"]
    print("pairs by series %d, by element sums %d, split %d, largest order %d" % (series, direct, split, lmax),
          flush=True)
    # the close pair is split: the series takes the far tiles of its slots,
    # the element sums the near ones
    assert series >= 10 and split >= 1 and direct + split >= 2 and lmax > 0
    assert ref["stats"][0] == 0 and ref["stats"][3] == 0
    assert got["it"] == ref["it"]
    for a, b in zip(got["U"], ref["U"]):
        assert _rel(a, b) < 1e-7, _rel(a, b)
    assert _rel(got["fld"], ref["fld"]) < 1e-7, _rel(got["fld"], ref["fld"])
    for (F1, N1), (F0, N0) in zip(got["FN"], ref["FN"]):
        s = np.abs(F0).max()
        assert np.abs(F1 - F0).max() < 1e-7 * s
        assert np.abs(N1 - N0).max() < 1e-7 * max(np.abs(N0).max(), s)

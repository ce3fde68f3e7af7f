"""The coupling loop's solves of all objects in lock step (the mixed form's
batched refinement: every triangular solve of every object in one launch
per factor and triangle) against the same loop with each object solved on
its own (BEM3_HIP_BATCH_SOLVE=0), and a second loop on the same objects
against the first: each loop starts from the incident field
(mo_solve_bv, mo_solve.c:10-88, after mo_initialize), so it repeats the
first one's iterations instead of starting at its converged state."""
import os

import numpy as np
import pytest

from bem3_emf_b2_amd import meshgen as mg

pytestmark = pytest.mark.gpu

BEAM = np.array([[1.064, 1.33, 0.665, 1.0, 0.0, 0.3, 0.1, 0.0, 0.0, 0.0, 0.4, 0.2]])


def _scene():
    return [
        (mg.icosphere(2, 0.3, center=(-0.9, 0.0, 0.0)), 1.45),
        (mg.box(0.5, 2, center=(0.15, 0.65, 0.05)), 1.60),
        (mg.cone(0.3, 0.6, 12, 3, center=(0.3, -0.65, 0.2)), 1.55 + 0.01j),
        (mg.icosphere(2, 0.25, center=(0.95, 0.1, -0.3)), 1.70),
        (mg.icosphere(1, 0.2, center=(0.0, 0.0, 1.2)), 1.50),
    ]


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _objects():
    import bem3_emf_b2_amd as bem
    G = []
    for (V, F), n in _scene():
        rn, ed, md, sd, gd = mg.paired_arrays(V, F)
        g = bem.Object(rn, ed, md, sd, [BEAM[0][1], n], BEAM[0][0])
        g.set_precision("mixed")
        g.assemble(2)
        g.set_beams(BEAM)
        g.factor(0)
        G.append(g)
    return G


def _run(bem, G, batch, pts):
    old = os.environ.get("BEM3_HIP_BATCH_SOLVE")
    os.environ["BEM3_HIP_BATCH_SOLVE"] = "1" if batch else "0"
    try:
        it, tr = bem.mo_solve(G, ccd=1e-10)
    finally:
        if old is None:
            del os.environ["BEM3_HIP_BATCH_SOLVE"]
        else:
            os.environ["BEM3_HIP_BATCH_SOLVE"] = old
    fld, oid = bem.mo_field(G, pts, "EH", "t", 1)
    stats = [g.solve_stats() for g in G]
    return it, tr, fld, oid, stats


def test_batched_solves_match_per_object_solves():
    import bem3_emf_b2_amd as bem
    G = _objects()
    pts = np.random.default_rng(5).uniform(-1.8, 1.8, (64, 3))
    it1, tr1, f1, oid1, st1 = _run(bem, G, True, pts)
    # every object's refinement reached its 1e-10 target (the last sweep's
    # warm start may already be there: zero steps)
    for steps, res in st1:
        assert res <= 1e-10 and steps >= 0
    it0, tr0, f0, oid0, _ = _run(bem, G, False, pts)
    keep = (oid1 < 0) & (oid0 < 0)
    assert keep.sum() > 10
    print("iterations batched %d, per object %d; fields %.2e" % (it1, it0, _rel(f1[keep], f0[keep])), flush=True)
    assert abs(it1 - it0) <= 1
    assert _rel(f1[keep], f0[keep]) < 1e-8
    # the change measures of the two loops follow each other down to the
    # refinement's accuracy
    n = min(it0, it1)
    assert np.allclose(tr1[2:n], tr0[2:n], rtol=1e-4, atol=1e-9)
    # a third loop on the same objects repeats the first (it starts from the
    # incident field, not from the converged state of the loop before it)
    it2, tr2, f2, _, _ = _run(bem, G, True, pts)
    assert it2 == it1
    assert tr2[2] > 1e-6                     # the first sweep changes the incident data again
    assert np.allclose(tr2[2:it2], tr1[2:it1], rtol=1e-4, atol=1e-9)
    assert _rel(f2[keep], f1[keep]) < 1e-9
    for g in G:
        g.close()


def test_graph_of_the_sweep_sums_matches_launches():
    """The coupling sums of every sweep after the first launched as one
    captured graph (the default) against the same kernels launched one by
    one (BEM3_HIP_GRAPH=0): the same kernels in the same order, so the
    change measures and the fields agree to the last bit."""
    import bem3_emf_b2_amd as bem
    G = _objects()
    pts = np.random.default_rng(6).uniform(-1.8, 1.8, (48, 3))
    res = {}
    for graph in (True, False):
        old = os.environ.get("BEM3_HIP_GRAPH")
        os.environ["BEM3_HIP_GRAPH"] = "1" if graph else "0"
        try:
            it, tr = bem.mo_solve(G, ccd=1e-10)
        finally:
            if old is None:
                del os.environ["BEM3_HIP_GRAPH"]
            else:
                os.environ["BEM3_HIP_GRAPH"] = old
        fld, oid = bem.mo_field(G, pts, "EH", "t", 1)
        res[graph] = (it, tr, fld, [g.state(0, "U") for g in G])
    (i1, t1, f1, u1), (i0, t0, f0, u0) = res[True], res[False]
    assert i1 == i0 and i1 > 3
    assert np.array_equal(t1, t0)
    assert np.array_equal(f1, f0)
    for a, b in zip(u1, u0):
        assert np.array_equal(a, b)
    for g in G:
        g.close()

"""The benchmark configurations at their full sizes on the device.

* Level-2 sphere (320 triangles): the plain FP64 form end to end against
  the oracle (the reference's dense 8NN system): E/H at 1000 seeded points
  in a 4 um box with the 7-point rule, force and torque with both rules
  (the example1.c protocol), 1e-9.
* Level-4 sphere (5120 triangles, 15360 element nodes): the device's
  coefficient matrices at 48 sampled row nodes of each domain against the
  oracle's rows (create_cmatrix_domain, solve_bieq2.c:98-184, through
  dcoef_bd_node_t2, d3b1_elem.c:482-539), 1e-12; the plain FP64 form end to
  end, and the mixed-precision form against it at 1e-6 (E/H at 1000 points
  with the 7-point rule, force and torque with both rules).
* The 16-object uneven scene (36648 triangles): both forms end to end, the
  mixed form against the FP64 form at 1e-6 for the fields at 200 points
  outside the objects and the force and torque of every object; and the
  coupling sums of the converged state at sampled node slots against the
  oracle's mEMP2_s (mo_solve.c:296-396) from the same boundary values:
  1e-10 for the FP64 form's element sums, 1e-7 for the mixed form's series
  and FP32 sums (their stated accuracy)."""
import numpy as np
import pytest

import oracle as orc
from bem3_emf_b2_amd import meshgen as mg
from conftest import progress

pytestmark = pytest.mark.gpu

BEAM = mg.xpol_plane_wave(lambda0=1.064, n0=1.33, E0=1.0)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


# ------------------------------------------------------------------ level-2 sphere
def test_l2_fp64_form_vs_oracle():
    import bem3_emf_b2_amd as bem
    V, F = mg.icosphere(2, 0.5)
    arrays = mg.paired_arrays(V, F)
    o = orc.Obj.from_arrays(*arrays, [0j, 1.59])
    o.setup(BEAM)
    S = orc.System(o)
    orc.mo_solve([o], [S])
    rn, ed, md, sd, gd = arrays
    g = bem.Object(rn, ed, md, sd, [1.33, 1.59], 1.064)
    g.set_beams(BEAM)
    g.assemble(2)
    g.factor(0)
    bem.mo_solve([g])
    pts = np.random.default_rng(20240601).uniform(-2.0, 2.0, size=(1000, 3))
    f_ref, oid_ref = orc.mo_field([o], pts, "EH", "t", 1)
    f_got, oid_got = bem.mo_field([g], pts, "EH", "t", 1)
    assert (oid_ref == oid_got).all()
    assert _rel(f_got, f_ref) < 1e-9, _rel(f_got, f_ref)
    for typ in (0, 1):
        _, F0, N0 = orc.mo_force([o], 0, typ)
        _, F1, N1 = bem.mo_force([g], 0, typ)
        assert _rel(F1, F0) < 1e-9, (typ, _rel(F1, F0))
        assert np.abs(N1 - N0).max() < 1e-9 * max(np.abs(N0).max(), np.abs(F0).max()), typ
    S.free()
    o.free()
    g.close()


# ------------------------------------------------------------------ level-4 sphere
@pytest.fixture(scope="module")
def l4():
    V, F = mg.icosphere(4, 0.5)
    arrays = mg.paired_arrays(V, F)
    return arrays


def test_l4_sampled_rows_match_oracle(l4):
    import bem3_emf_b2_amd as bem
    rn, ed, md, sd, gd = l4
    g = bem.Object(rn, ed, md, sd, [1.33, 1.59], 1.064)
    g.assemble(2)
    o = orc.Obj.from_arrays(*l4, [0j, 1.59])
    o.setup(BEAM)
    rng = np.random.default_rng(44)
    for did in (0, 1):
        ne = o.dNe[did]
        elems = np.sort(rng.choice(np.arange(1, ne + 1), 16, replace=False))
        rows = np.array([3 * (t - 1) + i for t in elems for i in range(3)])
        cols = np.array([4 * s + i for s in range(ne) for i in range(3)])   # padded -> compact columns
        ref = {k: [] for k in ("G", "H", "dG", "dH", "dF")}
        for t in elems:
            r = o.cmatrix_rows(did, int(t), int(t) + 1)
            for k in ref:
                ref[k].append(r[k][:3])
        for key in ("G", "H"):
            want = np.concatenate(ref[key])[:, cols]
            got = g.matrix_rows(did, key, rows)
            assert _rel(got, want) < 1e-12, (did, key, _rel(got, want))
        for key in ("dG", "dH"):
            want = np.concatenate(ref[key])[:, :, cols]
            got = g.matrix_rows(did, key, rows)
            assert _rel(got, want) < 1e-12, (did, key, _rel(got, want))
        want = np.concatenate(ref["dF"])
        got = g.matrix_rows(did, "dF", rows)
        assert np.abs(got - want).max() < 1e-12 * np.abs(want).max(), did
    o.free()
    g.close()


def test_l4_mixed_form_vs_fp64(l4):
    import bem3_emf_b2_amd as bem
    rn, ed, md, sd, gd = l4
    g = bem.Object(rn, ed, md, sd, [1.33, 1.59], 1.064)
    g.set_beams(BEAM)
    pts = np.random.default_rng(20240601).uniform(-2.0, 2.0, size=(1000, 3))
    res = {}
    for form in ("fp64", "mixed"):
        g.set_precision(form)
        g.assemble(2)
        g.factor(0)
        bem.mo_solve([g])
        fld, oid = bem.mo_field([g], pts, "EH", "t", 1)
        FN = [bem.mo_force([g], 0, typ)[1:] for typ in (0, 1)]
        res[form] = (fld, oid, FN)
    (f64, oid, FN64), (fmx, _, FNmx) = res["fp64"], res["mixed"]
    e_f = _rel(fmx, f64)
    e_F = max(_rel(a[0], b[0]) for a, b in zip(FNmx, FN64))
    e_N = max(np.abs(a[1] - b[1]).max() / max(np.abs(b[1]).max(), np.abs(b[0]).max()) for a, b in zip(FNmx, FN64))
    print("L4 mixed vs fp64: fields %.2e force %.2e torque %.2e" % (e_f, e_F, e_N), flush=True)
    assert e_f < 1e-6 and e_F < 1e-6 and e_N < 1e-6
    g.close()


# ------------------------------------------------------------------ 16-object scene
DSTS = (0, 5, 7, 12)   # destination objects of the coupling-sum check


def _scene_run(form):
    """one configuration of the scene in one form; the objects are closed
    before returning (both forms at once would not fit in HBM)"""
    import bem3_emf_b2_amd as bem
    objs = mg.scene16()
    G = []
    for o in objs:
        rn, ed, md, sd, gd = mg.paired_arrays(o["V"] + o["center"], o["F"])
        g = bem.Object(rn, ed, md, sd, [BEAM[0][1], o["n"]], BEAM[0][0])
        g.set_precision(form)
        g.set_beams(BEAM)
        G.append(g)
    progress("scene %s: objects" % form)
    bem.mo_prepare(G, 2, "auto")
    it, tr = bem.mo_solve(G, ccd=1e-8)
    pts = np.random.default_rng(20240602).uniform(-5.0, 5.0, size=(200, 3))
    fld, oid = bem.mo_field(G, pts, "EH", "t", 1)
    FN = []
    for m in range(len(G)):
        FN.append(bem.mo_force(G, m, 0, tuple(objs[m]["center"]))[1:])
        progress("scene %s: force of object %d" % (form, m))
    states = [(g.state(0, "U"), g.state(0, "dU")) for g in G]
    sums = {d: bem.mo_coupling_sums(G, d) for d in DSTS}
    for g in G:
        g.close()
    return dict(objs=objs, it=it, fld=fld, oid=oid, FN=FN, states=states, sums=sums)


@pytest.fixture(scope="module")
def scene():
    return _scene_run("fp64"), _scene_run("mixed")


def test_scene_mixed_form_vs_fp64(scene):
    r64, rmx = scene
    assert abs(rmx["it"] - r64["it"]) <= 1
    keep = (r64["oid"] < 0) & (rmx["oid"] < 0)
    assert keep.sum() > 150
    e_f = _rel(rmx["fld"][keep], r64["fld"][keep])
    worst_F = worst_N = 0.0
    for (F1, N1), (F0, N0) in zip(rmx["FN"], r64["FN"]):
        worst_F = max(worst_F, _rel(F1, F0))
        worst_N = max(worst_N, np.abs(N1 - N0).max() / max(np.abs(N0).max(), np.abs(F0).max()))
    print("scene mixed vs fp64: fields %.2e force %.2e torque %.2e" % (e_f, worst_F, worst_N), flush=True)
    assert e_f < 1e-6 and worst_F < 1e-6 and worst_N < 1e-6


@pytest.mark.parametrize("form,tol", [("fp64", 1e-10), ("mixed", 1e-7)])
def test_scene_coupling_sums_match_oracle(scene, form, tol):
    """the loop's coupling sums at the converged boundary values, at 8 node
    slots of each of 4 destination objects, against the oracle's mEMP2_s
    summed over the other 15 objects (the same boundary values loaded)"""
    r = scene[0] if form == "fp64" else scene[1]
    objs = r["objs"]
    O = []
    for o, (U, dU) in zip(objs, r["states"]):
        ob = orc.Obj.from_arrays(*mg.paired_arrays(o["V"] + o["center"], o["F"]), [0j, o["n"]])
        ob.setup(BEAM)
        ob.state(0, "U", U)
        ob.state(0, "dU", dU)
        O.append(ob)
    rng = np.random.default_rng(5)
    worst = 0.0
    for dst in DSTS:
        progress("coupling sums of object %d vs oracle" % dst)
        U, dU = r["sums"][dst]
        geo = O[dst].geometry()
        open_faces = np.nonzero(geo["md"][1:] == 0)[0] + 1   # elements whose main side is the open domain
        for t in rng.choice(open_faces, 8, replace=False):
            j = int(rng.integers(0, 4))
            rt = geo["ren"][t, j]
            v = geo["wen"][t, j] / np.linalg.norm(geo["wen"][t, j])
            u0, du0 = np.zeros(4, complex), np.zeros(4, complex)
            for src in range(len(O)):
                if src == dst:
                    continue
                did, u, du = O[src].mEMP2_s(rt, v)
                assert did == 0
                u0 += u
                du0 += du
            worst = max(worst, np.abs(U[t, j] - u0).max() / np.abs(u0).max(),
                        np.abs(dU[t, j] - du0).max() / np.abs(du0).max())
    for ob in O:
        ob.free()
    print("%s coupling sums vs oracle: %.2e" % (form, worst), flush=True)
    assert worst < tol, worst

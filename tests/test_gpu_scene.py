"""Uneven multi-object scene on the device (the 16-object configuration's
shape, scaled down so that the CPU oracle solves it in seconds): a sphere,
a cube, a cone and a second sphere of different sizes and refractive
indices, coupled through mo_solve to a 1e-8 change.  The FP64 block form
against the oracle (iteration count, fields outside the objects, force and
torque on every object, 1e-8), and the mixed-precision form against the
oracle at its 1e-6 field tolerance (force and torque 1e-5)."""
import numpy as np
import pytest

import oracle as orc
from bem3_emf_b2_amd import meshgen as mg

pytestmark = pytest.mark.gpu

BEAM = np.array([[1.064, 1.33, 0.665, 1.0, 0.0, 0.3, 0.1, 0.0, 0.0, 0.0, 0.4, 0.2]])


def _scene():
    return [
        (mg.icosphere(1, 0.3, center=(-0.9, 0.0, 0.0)), 1.45),
        (mg.box(0.5, 1, center=(0.15, 0.65, 0.05)), 1.60),
        (mg.cone(0.3, 0.6, 8, 2, center=(0.3, -0.65, 0.2)), 1.55 + 0.01j),
        (mg.icosphere(1, 0.25, center=(0.95, 0.1, -0.3)), 1.70),
    ]


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def oracle_scene():
    O, S = [], []
    for (V, F), n in _scene():
        o = orc.Obj.from_arrays(*mg.paired_arrays(V, F), [0j, n])
        o.setup(BEAM)
        O.append(o)
        S.append(orc.System(o))
    it, tr = orc.mo_solve(O, S, ccd=1e-8)
    rng = np.random.default_rng(11)
    pts = rng.uniform(-1.6, 1.6, (48, 3))
    fld, oid = orc.mo_field(O, pts, "EH", "t", 1)
    keep = oid < 0    # points outside every object
    FN = [orc.mo_force(O, m, t)[1:] for m in range(len(O)) for t in (0, 1)]
    return dict(it=it, tr=tr, pts=pts[keep], fld=fld[keep], FN=FN, n=len(O))


def _device_scene(prec):
    import bem3_emf_b2_amd as bem
    G = []
    for (V, F), n in _scene():
        rn, ed, md, sd, gd = mg.paired_arrays(V, F)
        g = bem.Object(rn, ed, md, sd, [BEAM[0][1], n], BEAM[0][0])
        g.set_precision(prec)
        g.assemble(2)
        g.set_beams(BEAM)
        g.factor(0)
        G.append(g)
    return G


# mixed form: 1e-6, the tolerance the form is specified for, on the fields
# and on force and torque (net values of cancelling stress integrals)
@pytest.mark.parametrize("prec,tol,ftol", [("fp64", 1e-8, 1e-8), ("mixed", 1e-6, 1e-6)])
def test_uneven_scene_coupling(oracle_scene, prec, tol, ftol):
    import bem3_emf_b2_amd as bem
    ref = oracle_scene
    G = _device_scene(prec)
    it, tr = bem.mo_solve(G, ccd=1e-8)
    if prec == "fp64":
        assert it == ref["it"]
        assert np.abs(tr[2:it + 1] - ref["tr"][2:it + 1]).max() / ref["tr"][2] < 1e-6
    else:
        assert abs(it - ref["it"]) <= 1
    fld, oid = bem.mo_field(G, ref["pts"], "EH", "t", 1)
    assert (oid < 0).all()
    assert _rel(fld, ref["fld"]) < tol, _rel(fld, ref["fld"])
    worst = 0.0
    k = 0
    for m in range(ref["n"]):
        for t in (0, 1):
            F0, N0 = ref["FN"][k]
            _, F1, N1 = bem.mo_force(G, m, t)
            scale = np.abs(F0).max()
            worst = max(worst, np.abs(F1 - F0).max() / scale, np.abs(N1 - N0).max() / max(np.abs(N0).max(), scale))
            print("%s object %d rule %d: force %.2e torque %.2e" % (prec, m, t, np.abs(F1 - F0).max() / scale,
                  np.abs(N1 - N0).max() / max(np.abs(N0).max(), scale)), flush=True)
            assert np.abs(F1 - F0).max() < ftol * scale, (m, t)
            assert np.abs(N1 - N0).max() < ftol * max(np.abs(N0).max(), scale), (m, t)
            k += 1
    for g in G:
        g.close()

"""Config 3: the level-5 sphere (20480 triangles, NN = 61440, n = 1.59 in
water, x-polarised 1.064 um plane wave) on one device in the low-memory
mixed form: G, H assembled in complex FP32 and factorised in place (the
FP64 matrices of both domains would take 242 GB, their tangential
derivatives another 483 GB), the FP64 right-hand side, refinement residual
and boundary E/H rows from blocks of rows assembled again.

Checks: the refinement reaches a relative FP64 residual of 1e-10; the
solved boundary values satisfy the oracle's FP64 rows of both domains'
boundary integral equations (H U = G dU and H E = G dE at the rows of 8
sampled elements per domain, create_cmatrix_domain, solve_bieq2.c:98-184)
to 1e-6 of their terms (the device's mixed form takes reduced far rules,
bounded at 3e-8 per pair); the fields at 16 points outside the sphere match
the oracle's mo_EH_t on the same boundary values to 1e-9."""
import time

import numpy as np
import pytest

import oracle as orc
from bem3_emf_b2_amd import meshgen as mg
from conftest import progress

pytestmark = pytest.mark.gpu

BEAM = mg.xpol_plane_wave(lambda0=1.064, n0=1.33, E0=1.0)


def test_l5_low_memory_form():
    import bem3_emf_b2_amd as bem
    V, F = mg.icosphere(5, 0.5)
    arrays = mg.paired_arrays(V, F)
    rn, ed, md, sd, gd = arrays
    g = bem.Object(rn, ed, md, sd, [1.33, 1.59], 1.064)
    assert g.NN == 61440
    g.set_precision("mixed_lm")
    g.set_beams(BEAM)
    t = [time.time()]
    g.assemble(2)
    t.append(time.time())
    progress("L5: assembled in %.1f s" % (t[-1] - t[-2]))
    g.factor(0)
    t.append(time.time())
    progress("L5: factorised in %.1f s" % (t[-1] - t[-2]))
    bem.mo_solve([g])
    t.append(time.time())
    steps, res = g.solve_stats()
    progress("L5: solved with boundary E/H in %.1f s (%d refinement steps, residual %.1e)" % (t[-1] - t[-2], steps, res))
    assert res <= 1e-10
    st = {(d, k): g.state(d, k) for d in (0, 1) for k in ("U", "dU", "E", "dE", "H", "dH")}
    pts = np.random.default_rng(5).normal(size=(16, 3))
    pts = 1.5 * pts / np.linalg.norm(pts, axis=1)[:, None]
    fld, oid = bem.mo_field([g], pts, "EH", "t", 1)
    assert (oid < 0).all()
    g.close()
    o = orc.Obj.from_arrays(*arrays, [0j, 1.59])
    o.setup(BEAM)
    rng = np.random.default_rng(55)
    worst = 0.0
    for d in (0, 1):
        U, dU = st[(d, "U")][1:].reshape(-1, 4), st[(d, "dU")][1:].reshape(-1, 4)
        E, dE = st[(d, "E")][1:].reshape(-1, 3), st[(d, "dE")][1:].reshape(-1, 3)
        for te in rng.choice(np.arange(1, o.dNe[d] + 1), 8, replace=False):
            r = o.cmatrix_rows(d, int(te), int(te) + 1)
            for A, dA in ((U, dU), (E, dE)):
                hu, gdu = r["H"] @ A, r["G"] @ dA
                e = np.abs(hu - gdu).max() / max(np.abs(hu).max(), np.abs(gdu).max())
                worst = max(worst, e)
        progress("L5: oracle rows of domain %d checked" % d)
    print("L5 boundary values vs the oracle's FP64 rows: %.2e" % worst, flush=True)
    assert worst < 1e-6
    for d in (0, 1):
        for k in ("U", "dU", "E", "dE", "H", "dH"):
            o.state(d, k, st[(d, k)])
    ref, oid_ref = orc.mo_field([o], pts, "EH", "t", 1)
    o.free()
    e = np.abs(fld - ref).max() / np.abs(ref).max()
    print("L5 fields vs the oracle's on the same boundary values: %.2e; phases %s s" %
          (e, ", ".join("%.1f" % (t[i + 1] - t[i]) for i in range(3))), flush=True)
    assert e < 1e-9

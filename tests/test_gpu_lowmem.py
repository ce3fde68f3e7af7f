"""The low-memory mixed form (BH_PREC_MIXED_LM: G, H stored only in complex
FP32 and factorised in place, every FP64 product -- right-hand side,
refinement residual, boundary E/H rows, the dE/dH refinement -- from rows
assembled again a block at a time) against the mixed form that keeps the
FP64 matrices.  Both assemble the same FP64 entries (the low-memory form
rounds each block into its FP32 storage), so they differ only by the
rounding of the in-place factorisation, which the FP64 refinement removes:
boundary values of both domains, fields and force agree to 1e-9.  Blocks of
100 rows (the last one partial) exercise the block bookkeeping; a second
object with one block per domain checks the default block size."""
import numpy as np
import pytest

from bem3_emf_b2_amd import meshgen as mg

pytestmark = pytest.mark.gpu

BEAM = np.array([[1.064, 1.33, 0.665, 1.0, 0.0, 0.3, 0.1, 0.0, 0.0, 0.0, 0.4, 0.2]])


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _run(prec, arrays, rows=0):
    import bem3_emf_b2_amd as bem
    rn, ed, md, sd, gd = arrays
    g = bem.Object(rn, ed, md, sd, [1.33, 1.59], 1.064)
    g.set_precision(prec)
    if rows:
        g.set_block_rows(rows)
    g.set_beams(BEAM)
    g.assemble(2)
    g.factor(0)
    bem.mo_solve([g])
    st = {(d, k): g.state(d, k) for d in (0, 1) for k in ("U", "dU", "E", "dE", "H", "dH")}
    pts = np.random.default_rng(3).uniform(-1.5, 1.5, (64, 3))
    fld, _ = bem.mo_field([g], pts, "EH", "t", 1)
    FN = [bem.mo_force([g], 0, t)[1:] for t in (0, 1)]
    ngh = (g.gh_count(0), g.gh_count(1))
    g.close()
    return st, fld, FN, ngh


@pytest.mark.parametrize("rows", [100, 0])
def test_low_memory_form_matches_mixed(rows):
    V, F = mg.icosphere(2, 0.5)
    arrays = mg.paired_arrays(V, F)
    ref = _run("mixed", arrays)
    got = _run("mixed_lm", arrays, rows)
    assert got[3] == ref[3]   # the same near pairs take the 64-point rule
    worst = 0.0
    for key, want in ref[0].items():
        e = _rel(got[0][key], want)
        worst = max(worst, e)
        assert e < 1e-9, (key, e)
    assert _rel(got[1], ref[1]) < 1e-9
    for (F1, N1), (F0, N0) in zip(got[2], ref[2]):
        s = np.abs(F0).max()
        assert np.abs(F1 - F0).max() < 1e-9 * s
        assert np.abs(N1 - N0).max() < 1e-9 * max(np.abs(N0).max(), s)
    print("low-memory form (%s rows per block) vs mixed: boundary values %.2e, fields %.2e"
          % (rows or "default", worst, _rel(got[1], ref[1])), flush=True)

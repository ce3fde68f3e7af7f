"""Coefficient-matrix assembly on the GPU vs the CPU oracle (restating
create_cmatrix_domain, solve_bieq2.c:98-184, and dcoef_bd_node_t2,
d3b1_elem.c:482-539).  Tolerance: 1e-12 of the largest entry of each
matrix, FP64 throughout."""
import numpy as np
import pytest

import oracle as orc
from bem3_emf_b2_amd import meshgen as mg

pytestmark = pytest.mark.gpu

CASES = {
    "sphere_L1_dielectric": dict(level=1, radius=0.5, n=1.59),
    "sphere_L2_dielectric": dict(level=2, radius=0.5, n=1.59),
    "sphere_L1_metal": dict(level=1, radius=0.3, n=0.06 + 4.0j),
}


def _pair(case):
    import bem3_emf_b2_amd as bem
    V, F = mg.icosphere(case["level"], case["radius"])
    arrays = mg.paired_arrays(V, F)
    beams = mg.xpol_plane_wave()
    o = orc.Obj.from_arrays(*arrays, [0j, case["n"]])
    o.setup(beams)
    rn, ed, md, sd, gd = arrays
    g = bem.Object(rn, ed, md, sd, [1.33, case["n"]], 1.064)
    return o, g


def _compact(ne):
    """padded (element, 4 slots) -> compact (element, 3 slots) index map"""
    return np.array([4 * t + i for t in range(ne) for i in range(3)])


@pytest.mark.parametrize("name", list(CASES))
def test_assembly_matches_oracle(name):
    o, g = _pair(CASES[name])
    g.assemble(2)
    for did in (0, 1):
        ref = o.cmatrix(did, 2, want_d=True)
        idx = _compact(o.dNe[did])
        assert g.gh_count(did) == ref["ngh"]
        for key in ("G", "H"):
            want = ref[key][np.ix_(idx, idx)]
            got = g.matrix(did, key)
            err = np.abs(got - want).max() / np.abs(want).max()
            assert err < 1e-12, (did, key, err)
        for key in ("dG", "dH"):
            want = ref[key][idx][:, :, idx]
            got = g.matrix(did, key)
            err = np.abs(got - want).max() / np.abs(want).max()
            assert err < 1e-12, (did, key, err)
        dF = g.matrix(did, "dF")
        assert np.abs(dF - ref["dF"][idx]).max() < 1e-12 * np.abs(ref["dF"]).max()


def test_quad_mesh_assembly_matches_oracle():
    """The reference's verification1 cone (28 triangles and 553 bi-linear
    quadrangles, metal n = 0.06+4i): G, H, dG, dH, dF of both domains with
    the quadrangles' far pairs (12x12 rule), near pairs (64x64 rule) and
    self element (polar rule, principal-value derivatives of
    bil_dcoef.c:369-740, H and F terms of warped quadrangles)."""
    import bem3_emf_b2_amd as bem
    from fixtures import active_slots, v1_oracle, v1_raw
    o = v1_raw()[0]
    ob = v1_oracle()[0]
    g = bem.Object(o["rn"], o["ed"], o["md"], o["sd"], list(o["n"]), float(o["lambda0"]))
    g.assemble(2)
    for did in (0, 1):
        ref = ob.cmatrix(did, 2, want_d=True)
        idx = active_slots(o["ed"], o["dom"][did]["sid"])
        assert g.gh_count(did) == ref["ngh"], did
        for key in ("G", "H"):
            want = ref[key][np.ix_(idx, idx)]
            got = g.matrix(did, key)
            err = np.abs(got - want).max() / np.abs(want).max()
            assert err < 1e-12, (did, key, err)
        for key in ("dG", "dH"):
            want = ref[key][idx][:, :, idx]
            got = g.matrix(did, key)
            err = np.abs(got - want).max() / np.abs(want).max()
            assert err < 1e-12, (did, key, err)
        # the dF/dt row sums of the rows at the cone's sharp edges cancel
        # near-pair terms far larger than the sums themselves, so their
        # summation order shows at 1e-11 of max |dF|
        dF = g.matrix(did, "dF")
        assert np.abs(dF - ref["dF"][idx]).max() < 1e-10 * np.abs(ref["dF"]).max()
    g.close()

"""Oracle vs the reference's combined-object fixture (verification1/
combined_object): one object made of two separate cones
along z (z in [-0.325, -0.075] and [0.075, 0.325]), each its own medium
(MN = 2, both n = 0.06+4i), meshed together in 2cone.msh (56 triangles, 1106
quads); the solved state in vc.dat and the fields, forces, torques and
absorbed powers example1.c printed from it in example1.log.  This pins the
oracle's field, force and power code on an object with two interior
domains."""
import os

import numpy as np
import pytest

import oracle as orc
from datfile import parse_example1_log, read_dat
from fixtures import GOLDEN, stored_oracle

REF_PT = [0.3, 0.0, 1.1]  # example1.c:14-16


@pytest.fixture(scope="module")
def vc():
    return stored_oracle(read_dat(os.path.join(GOLDEN, "vc.dat.gz")))


@pytest.fixture(scope="module")
def vc_log():
    return parse_example1_log(os.path.join(GOLDEN, "vc_example1.log"))


def test_fixture_layout(vc):
    assert vc[0].MN == 2 and list(vc[0].dNe) == [1162, 581, 581]


@pytest.mark.parametrize("typ", [0, 1])
def test_total_field_matches_example1(vc, vc_log, typ):
    ref = vc_log["field"][typ]
    out, oid = orc.mo_field(vc, [REF_PT], "EH", "t", typ)
    want = np.array([ref[k] for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")])
    assert oid[0] == -1
    assert np.abs(out[0] - want).max() < 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("typ", [0, 1])
def test_force_torque_match_example1(vc, vc_log, typ):
    ref = vc_log["force"][typ][0]
    ret, F, N = orc.mo_force(vc, 0, typ)
    assert ret == 0
    assert np.abs(F - ref["F"]).max() < 1e-12 * np.abs(ref["F"]).max()
    assert np.abs(N - ref["N"]).max() < 1e-12 * np.abs(ref["N"]).max()


@pytest.mark.parametrize("typ", [0, 1])
def test_absorbed_power_matches_example1(vc, vc_log, typ):
    ref = vc_log["absorb"][typ][0]
    ret, P = orc.mo_absorb(vc, 0, typ)
    assert ret == 0
    assert abs(P - ref) < 1e-12 * ref


def test_points_in_each_cone_get_their_medium(vc):
    """domain_id (setup2.c:871-978) over the stored geometry: medium 1 is the
    lower cone, medium 2 the upper one, the gap between them is outside"""
    for pt, want in (([0, 0, -0.2], 1), ([0, 0, 0.2], 2), ([0, 0, 0.0], 0), ([0.02, 0.01, 0.28], 2),
                     ([0.02, 0.01, -0.28], 1), (REF_PT, 0)):
        assert vc[0].domain_id(pt) == want, pt

"""The reference's own golden fixture on the device: verification1/v1.dat
(two metal cones n = 0.06+4i, 28 triangles + 553 bi-linear quadrangles
each, the solved boundary state written by the reference) replayed through
the GPU field, boundary-field, force and absorbed-power kernels, against
the numbers the reference printed in verification1/example1.log (15
significant digits).  Also the reference's example1.c, compiled unchanged
against the drop-in library, run on the same .dat file.

Tolerance 1e-10 relative: the state is read from the file; what differs
from the reference is summation order, and the element constants and node
points that the library recomputes from the stored node coordinates."""
import gzip
import os
import shutil
import subprocess

import numpy as np
import pytest

from datfile import parse_example1_log
from fixtures import GOLDEN, v1_log, v1_raw, ref_example

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_PT = [0.3, 0.0, 1.1]  # example1.c:14-16
TOL = 1e-10


@pytest.fixture(scope="module")
def v1_device():
    import bem3_emf_b2_amd as bem
    G = []
    for o in v1_raw():
        g = bem.Object(o["rn"], o["ed"], o["md"], o["sd"], list(o["n"]), float(o["lambda0"]))
        g.set_beams(o["beams"])
        for d in range(o["MN"] + 1):
            for key in ("U", "dU", "E", "dE", "H", "dH"):
                g.state(d, key, o["dom"][d][key])
        G.append(g)
    yield G
    for g in G:
        g.close()


def test_device_geometry_matches_stored_node_data(v1_device):
    """node points and normals of every slot recomputed on the device from
    the stored node coordinates, triangles and quadrangles alike"""
    for g, o in zip(v1_device, v1_raw()):
        geo = g.geometry()
        for key in ("ren", "wen"):
            got, want = geo[key][1:], o[key][1:]
            assert np.abs(got - want).max() < 1e-14 * np.abs(want).max(), key


@pytest.mark.parametrize("typ", [0, 1])
def test_total_field_matches_example1_log(v1_device, typ):
    import bem3_emf_b2_amd as bem
    ref = v1_log()["field"][typ]
    out, oid = bem.mo_field(v1_device, [REF_PT], "EH", "t", typ)
    want = np.array([ref[k] for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")])
    assert oid[0] == -1
    assert np.abs(out[0] - want).max() < TOL * np.abs(want).max()


@pytest.mark.parametrize("typ", [0, 1])
@pytest.mark.parametrize("oid", [0, 1])
def test_force_torque_match_example1_log(v1_device, typ, oid):
    import bem3_emf_b2_amd as bem
    ref = v1_log()["force"][typ][oid]
    ret, F, N = bem.mo_force(v1_device, oid, typ)
    assert ret == 0
    assert np.abs(F - ref["F"]).max() < TOL * np.abs(ref["F"]).max()
    assert np.abs(N - ref["N"]).max() < TOL * np.abs(ref["N"]).max()


@pytest.mark.parametrize("typ", [0, 1])
@pytest.mark.parametrize("oid", [0, 1])
def test_absorbed_power_matches_example1_log(v1_device, typ, oid):
    import bem3_emf_b2_amd as bem
    ref = v1_log()["absorb"][typ][oid]
    ret, P = bem.mo_absorb(v1_device, oid, typ)
    assert ret == 0
    assert abs(P - ref) < TOL * ref


def test_reference_example1_on_v1_dat(tmp_path):
    """example1.c unchanged, relinked against the drop-in library, on the
    reference's v1.dat: every number it prints against the reference's own
    printout of the same program on the same file"""
    dat = tmp_path / "v1.dat"
    with gzip.open(os.path.join(GOLDEN, "v1.dat.gz"), "rb") as src, open(dat, "wb") as dst:
        shutil.copyfileobj(src, dst)
    print("run: example1 v1.dat", flush=True)
    r = subprocess.run([ref_example("example1"), str(dat)], cwd=str(tmp_path), capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    (tmp_path / "ex1.log").write_text(r.stdout)
    got = parse_example1_log(str(tmp_path / "ex1.log"))
    want = v1_log()
    for typ in (0, 1):
        w = np.array([want["field"][typ][k] for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")])
        g = np.array([got["field"][typ][k] for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")])
        assert np.abs(g - w).max() < TOL * np.abs(w).max(), typ
        for m in (0, 1):
            for q in ("F", "N"):
                a, b = got["force"][typ][m][q], want["force"][typ][m][q]
                assert np.abs(a - b).max() < TOL * np.abs(b).max(), (typ, m, q)
            assert abs(got["absorb"][typ][m] - want["absorb"][typ][m]) < TOL * want["absorb"][typ][m], (typ, m)


@pytest.mark.parametrize("typ", [0, 1, 2])
def test_EH_mEMP_on_quad_mesh_matches_oracle(v1_device, typ):
    """E, H from the gradients of the potentials (EH_mEMP_s,
    d3b1_field.c:68-170) over quadrangle sources (bil_dcoef_4p/9p/GL with
    the per-direction GHN switch) against the oracle on the stored state"""
    import bem3_emf_b2_amd as bem
    import oracle as orc
    from fixtures import v1_oracle
    pts = np.array([REF_PT, [0.0, 0.5, 0.1], [0.2, -0.3, 0.6], [-0.4, 0.1, -0.9], [0.1, 0.2, 3.0], [0.0, 0.0, 20.0]])
    got, gid = bem.mo_field(v1_device, pts, "EH_mEMP", "t", typ)
    want, wid = orc.mo_field(v1_oracle(), pts, "EH_mEMP", "t", typ)
    assert (gid == wid).all()
    for p in range(len(pts)):
        assert np.abs(got[p] - want[p]).max() < 1e-10 * np.abs(want[p]).max(), (p, typ)

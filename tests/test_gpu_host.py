"""The drop-in C host layer end to end on the GPU: the command-line programs
d3b2_create_matrix and d3b2_bv_solver (built from the reference's main
programs' call sequence over include/bem3_emf_b2.h) run on generated input
files, and the .dat file they write is checked against the CPU oracle run on
the same mesh, media and beam.  Where the reference's own example1.c has been
compiled against the drop-in library (oracle/_ref/example1), its printed
fields, force and absorbed power are checked too."""
import os
import subprocess

import numpy as np
import pytest

import oracle as orc
from bem3_emf_b2_amd import meshgen as mg
from datfile import parse_example1_log, read_dat

from fixtures import ref_example

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "bem3_emf_b2_amd", "bin")

BEAM = np.array([[1.064, 1.33, 0.665, 1.0, 0.0, 0.2, 0.1, 0.0, 0.0, 0.0, 0.3, 0.1]])
N_OBJ = 1.59
SHIFTS = [(0.0, 0.0, 0.45), (0.05, 0.0, -0.45)]


def _run(cmd, cwd):
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (cmd, r.stdout[-3000:], r.stderr[-3000:])
    return r.stdout


@pytest.fixture(scope="module")
def solved(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("host"))
    V, F = mg.icosphere(1, 0.35)
    mg.write_msh(os.path.join(d, "sphere.msh"), V, F)
    mg.write_medium(os.path.join(d, "medium.txt"), [N_OBJ])
    mg.write_ipw(os.path.join(d, "ipw.txt"), BEAM)
    _run([os.path.join(BIN, "d3b2_create_matrix"), "medium.txt", "sphere.msh", "sphere.obj"], d)
    with open(os.path.join(d, "objects.txt"), "w") as fp:
        fp.write("# number_of_objects reset_flag\n%d 0\n# file shift_x shift_y shift_z\n" % len(SHIFTS))
        for s in SHIFTS:
            fp.write("sphere.obj %.17g %.17g %.17g\n" % s)
    log = _run([os.path.join(BIN, "d3b2_bv_solver"), "objects.txt", "out.dat"], d)
    # the oracle on the same input files
    O = []
    for s in SHIFTS:
        o = orc.Obj.from_msh(os.path.join(d, "sphere.msh"), [N_OBJ])
        o.setup(BEAM, tv=np.array(s))
        O.append(o)
    S = [orc.System(o) for o in O]
    orc.mo_solve(O, S)
    return d, O, log


def test_obj_and_dat_files(solved):
    d, O, log = solved
    assert os.path.exists(os.path.join(d, "sphere.obj"))
    assert os.path.exists(os.path.join(d, "out.particles"))
    objs = read_dat(os.path.join(d, "out.dat"))
    assert len(objs) == 2
    for m, o in enumerate(objs):
        assert o["mod_fn"] == "sphere.obj"
        np.testing.assert_allclose(o["rs"], SHIFTS[m])
        np.testing.assert_allclose(o["beams"][0], BEAM[0])
        assert abs(o["lambda0"] - 1.064) < 1e-15 and abs(o["n0"] - 1.33) < 1e-15
        for did in (0, 1):
            for key in ("U", "dU", "E", "dE", "H", "dH"):
                want = O[m].state(did, key)
                got = o["dom"][did][key]
                err = np.abs(got - want).max() / np.abs(want).max()
                # coupling converged to 1e-6 change: both sides stop at the
                # same iteration, the remaining difference is solver round-off
                assert err < 1e-8, (m, did, key, err)


def test_solver_log_reports_convergence(solved):
    d, O, log = solved
    assert "iterative operation start" in log
    assert "count =" in log
    assert "finished." in log


def test_reference_example1_relinked(solved):
    """example1.c of the reference, compiled unchanged against the drop-in
    header and library, reads the .dat and prints fields, force, torque and
    absorbed power; the oracle computes the same from its own solution."""
    d, O, log = solved
    out = _run([ref_example("example1"), "out.dat"], d)
    with open(os.path.join(d, "ex1.log"), "w") as fp:
        fp.write(out)
    got = parse_example1_log(os.path.join(d, "ex1.log"))
    r = np.array([0.3, 0.0, 1.1])
    for typ in (0, 1):
        EH, _ = orc.mo_field(O, r[None, :], "EH", "t", typ)
        for i, key in enumerate(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")):
            assert abs(got["field"][typ][key] - EH[0, i]) <= 1e-8 * np.abs(EH[0]).max() + 1e-13, (typ, key)
        for m in range(2):
            ret, F, N = orc.mo_force(O, m, typ, (0.0, 0.0, 0.0))
            np.testing.assert_allclose(got["force"][typ][m]["F"], F, rtol=1e-7, atol=1e-7 * np.abs(F).max())
            np.testing.assert_allclose(got["force"][typ][m]["N"], N, rtol=1e-7, atol=1e-7 * np.abs(N).max() + 1e-15)
            ret, P = orc.mo_absorb(O, m, typ)
            # lossless spheres: P is the discretisation error of the flux
            # balance on this coarse mesh, the same number on both sides
            assert abs(got["absorb"][typ][m] - P) <= 1e-7 * abs(P) + 1e-14, (typ, m)




def test_reference_example4_far_field_relinked(solved):
    """example4.c (far-field intensity from E, H of the potentials' gradients,
    called from an OpenMP loop) relinked against the drop-in library: the
    printed intensities on the x=0 plane at 2000 wavelengths against the
    oracle's mo_EH_mEMP_s at the same points."""
    d, O, log = solved
    sn, mf, ty = 12, 2000.0, 1
    _run([ref_example("example4"), "out.dat", str(sn), str(mf), str(ty)], d)
    rows = [l.split() for l in open(os.path.join(d, "fsIe_yz.txt")) if not l.startswith("#")]
    assert len(rows) == sn
    ra = mf * 1.064
    dth = 2 * np.pi / sn
    th = 0.5 * dth + dth * np.arange(sn)
    pts = np.c_[np.zeros(sn), -ra * np.sin(th), ra * np.cos(th)]
    EH, _ = orc.mo_field(O, pts, "EH_mEMP", "s", ty)
    ie = (np.abs(EH[:, :3]) ** 2).sum(axis=1)
    got = np.array([float(r[1]) for r in rows])
    assert np.abs(got - ie).max() < 1e-8 * ie.max()

"""The reference's verification2 case run from its input files on the GPU,
against the outputs the reference itself printed for it.

Inputs (verification2/): the sphere mesh sphere_m0.msh (600 triangles),
medium_data.txt (n = 0.06+4i, a metal), ipw.txt (a plane wave along +z) and
object_settings.txt (three copies at x = 0, +0.5, -0.5).  d3b2_create_matrix
and d3b2_bv_solver of this library run on them as the reference's programs
were run (verification2/d3b2_create_matrix.log, d3b2_bv_solver.log); then the
reference's example1.c, compiled unchanged against this library, reads the
.dat file written here.  Checked: the iteration count and the convergence
trace of verification2/d3b2_bv_solver.log (6 printed digits), and every
number of verification2/example1.log (total field at (0.3, 0, 1.1) with both
rules, force, torque and absorbed power of each sphere, 15 printed digits).

Tolerances: both runs iterate the same fixed-point map from the same start
and stop at the same count, so what differs is the propagated round-off of
the factorisations: 1e-9 of the largest component of each printed vector
(torques: of 0.5 |F|, the centre sphere's torque being zero by symmetry),
1e-9 relative for the absorbed powers.  cc values are held to their printed
precision (relative 2e-5, 6 digits on both sides) plus 1e-10 absolute: cc is
a mean relative change of the state, so state differences at 1e-11 show in
the last iterations' values."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from datfile import parse_example1_log
from fixtures import GOLDEN, ref_example

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "bem3_emf_b2_amd", "bin")
OBJ = "sphere_m0_md.obj"
IPW = "## plane wave\n# number of waves\n1\n#lambda0 ni power real(e0x) imag(e0x) real(e0y) imag(e0y) fx fy fz theta phi\n" \
      "1.064 1.33 1.0e-3 1.0 0.0 0.0 0.0 0.0 0.0 0.0 0.0 0.0\n"
MEDIUM = "## medium data file for bem3_emf_b2\n# number of medium\n1\n# real(refractive_index) imag(refractive_index)\n" \
         "0.06 4.0\n"
SETTINGS = "# number_of_objects reset_flag_of_incident_field(0:no reset, 1:reset)\n3 0\n" \
           "# object_data_file_name shitf_x shift_y shit_z\n" \
           "%s +0.00 +0.00 +0.00\n%s +0.50 +0.00 +0.00\n%s -0.50 +0.00 +0.00\n" % (OBJ, OBJ, OBJ)


def _run(cmd, cwd):
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, (cmd, r.stdout[-3000:], r.stderr[-3000:])
    return r.stdout


def _trace(text):
    return {int(c): float(v) for c, v in re.findall(r"count =\s*(\d+), cc = ([0-9.eE+-]+)", text)}


@pytest.fixture(scope="module")
def v2_run(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("v2"))
    shutil.copy(os.path.join(GOLDEN, "sphere_m0.msh"), d)
    for name, text in (("ipw.txt", IPW), ("medium_data.txt", MEDIUM), ("object_settings.txt", SETTINGS)):
        with open(os.path.join(d, name), "w") as fp:
            fp.write(text)
    _run([os.path.join(BIN, "d3b2_create_matrix"), "medium_data.txt", "sphere_m0.msh", OBJ], d)
    log = _run([os.path.join(BIN, "d3b2_bv_solver"), "object_settings.txt", "v2.dat"], d)
    return d, log


def test_convergence_trace_matches_reference_log(v2_run):
    _, log = v2_run
    with open(os.path.join(GOLDEN, "v2_bv_solver.log")) as fp:
        want = _trace(fp.read())
    got = _trace(log)
    assert sorted(got) == sorted(want), (sorted(got), sorted(want))
    for c, w in want.items():
        assert abs(got[c] - w) <= 2e-5 * w + 1e-10, (c, got[c], w)


def test_reference_example1_matches_reference_log(v2_run):
    d, _ = v2_run
    out = _run([ref_example("example1"), "v2.dat"], d)
    with open(os.path.join(d, "ex1.log"), "w") as fp:
        fp.write(out)
    got = parse_example1_log(os.path.join(d, "ex1.log"))
    want = parse_example1_log(os.path.join(GOLDEN, "v2_example1.log"))
    keys = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
    for typ in (0, 1):
        w = np.array([want["field"][typ][k] for k in keys])
        g = np.array([got["field"][typ][k] for k in keys])
        print("field type %d: %.2e" % (typ, np.abs(g - w).max() / np.abs(w).max()), flush=True)
        assert np.abs(g - w).max() < 1e-9 * np.abs(w).max(), typ
        for m in range(3):
            # the centre sphere's torque vanishes by symmetry (printed values
            # are round-off): torques are scaled by |F| x 0.5 (the spacing)
            fs = np.abs(want["force"][typ][m]["F"]).max()
            for q, sc in (("F", fs), ("N", 0.5 * fs)):
                a, b = got["force"][typ][m][q], want["force"][typ][m][q]
                scale = max(np.abs(b).max(), sc)
                print("type %d object %d %s: %.2e" % (typ, m, q, np.abs(a - b).max() / scale), flush=True)
                assert np.abs(a - b).max() < 1e-9 * scale, (typ, m, q)
            pa, pb = got["absorb"][typ][m], want["absorb"][typ][m]
            assert abs(pa - pb) < 1e-9 * pb, (typ, m, pa, pb)

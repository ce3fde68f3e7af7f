"""The reference's verification1 case run from its input files on the GPU,
against the outputs the reference itself wrote for it.

Inputs (verification1/): the cone mesh cone_m1_p5.msh (28 triangles and 553
bi-linear quadrangles), medium_data.txt (n = 0.06+4i, a metal), ipw.txt (a
plane wave at theta = 0.3) and object_settings_v.txt (two copies shifted by
+-0.2 in z).  d3b2_create_matrix and d3b2_bv_solver of this library run on
them as the reference's programs were run (verification1/
d3b2_create_matrix.log, d3b2_bv_solver.log).  Checked: the iteration count
and the convergence trace cc of d3b2_bv_solver.log (6 printed digits), and
every solved boundary value (U, dU, E, dE, H, dH of both domains and both
objects) against the state the reference stored in v1.dat.

Tolerances: both runs iterate the same fixed-point map from the same start
and stop at the same count, so their states differ by the propagated
round-off of the two factorisations (MKL's and this library's): 1e-10 of
each quantity's largest entry (measured: 5e-15 for U, up to 3e-12 for the
normal derivatives of H).  The printed cc values are compared to their
printed precision (relative 2e-5, 6 digits on both sides) plus 1e-10
absolute: cc is a mean relative change of the state, so state differences at
1e-11 show in the last iterations' values."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from datfile import read_dat
from fixtures import GOLDEN, v1_raw

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "bem3_emf_b2_amd", "bin")
OBJ = "cone_m1_p5_md_x00.obj"
# verification1/ipw.txt: lambda0 ni power re(e0x) im(e0x) re(e0y) im(e0y) fx fy fz theta phi
IPW = "## plane wave\n# number of waves\n1\n#lambda0 ni power real(e0x) imag(e0x) real(e0y) imag(e0y) fx fy fz theta phi\n" \
      "1.064 1.33 1.0e-3 1.0 0.0 0.0 0.0 0.0 0.0 0.0 0.3 0.0\n"
MEDIUM = "## medium data file for bem3_emf_b2\n# number of medium\n1\n# real(refractive_index) imag(refractive_index)\n" \
         "0.06 4.0\n"
SETTINGS = "# number_of_objects reset_flag_of_incident_field(0:no reset, 1:reset)\n2 1\n" \
           "# object_data_file_name shitf_x shift_y shit_z\n" \
           "%s +0.00 +0.00 +0.20\n%s +0.00 +0.00 -0.20\n" % (OBJ, OBJ)


def _run(cmd, cwd):
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, (cmd, r.stdout[-3000:], r.stderr[-3000:])
    return r.stdout


def _trace(text):
    return {int(c): float(v) for c, v in re.findall(r"count =\s*(\d+), cc = ([0-9.eE+-]+)", text)}


@pytest.fixture(scope="module")
def v1_run(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("v1"))
    shutil.copy(os.path.join(GOLDEN, "cone_m1_p5.msh"), d)
    for name, text in (("ipw.txt", IPW), ("medium_data.txt", MEDIUM), ("object_settings_v.txt", SETTINGS)):
        with open(os.path.join(d, name), "w") as fp:
            fp.write(text)
    _run([os.path.join(BIN, "d3b2_create_matrix"), "medium_data.txt", "cone_m1_p5.msh", OBJ], d)
    log = _run([os.path.join(BIN, "d3b2_bv_solver"), "object_settings_v.txt", "v1.dat"], d)
    return read_dat(os.path.join(d, "v1.dat")), log


def test_convergence_trace_matches_reference_log(v1_run):
    _, log = v1_run
    with open(os.path.join(GOLDEN, "v1_bv_solver.log")) as fp:
        want = _trace(fp.read())
    got = _trace(log)
    assert sorted(got) == sorted(want), (sorted(got), sorted(want))
    for c, w in want.items():
        assert abs(got[c] - w) <= 2e-5 * w + 1e-10, (c, got[c], w)


def test_solved_state_matches_reference_v1_dat(v1_run):
    objs, _ = v1_run
    ref = v1_raw()
    assert len(objs) == 2
    for m in range(2):
        assert objs[m]["mod_fn"] == OBJ
        np.testing.assert_allclose(objs[m]["rs"], ref[m]["rs"], rtol=0, atol=1e-15)
        for did in (0, 1):
            for key in ("U", "dU", "E", "dE", "H", "dH"):
                got, want = objs[m]["dom"][did][key], ref[m]["dom"][did][key]
                err = np.abs(got - want).max() / np.abs(want).max()
                print("object %d domain %d %-2s %.2e" % (m, did, key, err), flush=True)
                assert err < 1e-10, (m, did, key, err)

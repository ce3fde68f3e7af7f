"""The reference's combined-object case (verification1/combined_object) run
from its input files on the GPU, against the outputs the reference itself
wrote for it.

Inputs: 2cone.msh (one object made of two separate cones, media 1 and 2, 56
triangles and 1106 bi-linear quads), medium_data.txt (two media, both
n = 0.06+4i), ipw.txt (a plane wave at theta = 0.3) and object_settings.txt
(the single object, unshifted).  With two interior media the library takes
the dense 8NN form (na = 36736).  d3b2_create_matrix and d3b2_bv_solver run
as the reference's programs were run (combined_object/d3b2_create_matrix.log,
d3b2_bv_solver.log); the reference's example1.c, relinked against this
library, then reads the .dat written here.  Checked: every solved boundary
value of the three domains against vc.dat, and every number of
combined_object/example1.log.

Tolerances: one object, so no iteration: the states differ by the round-off
of the two factorisations (MKL's LU of the same dense matrix and this
library's), 1e-10 of each quantity's largest entry; fields, forces, torques
and powers 1e-9 (of the largest component of each printed vector; torques of
|F| x 0.3, the cones' size, as their x and z components are round-off)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from datfile import parse_example1_log, read_dat
from fixtures import GOLDEN, ref_example

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "bem3_emf_b2_amd", "bin")
OBJ = "2cone_md.obj"
IPW = "## plane wave\n# number of waves\n1\n#lambda0 ni power real(e0x) imag(e0x) real(e0y) imag(e0y) fx fy fz theta phi\n" \
      "1.064 1.33 1.0e-3 1.0 0.0 0.0 0.0 0.0 0.0 0.0 0.3 0.0\n"
MEDIUM = "## medium data file for bem3_emf_b2\n# number of medium\n2\n# real(refractive_index) imag(refractive_index)  \n" \
         "0.06 4.0\n0.06 4.0\n"
SETTINGS = "# number_of_objects reset_flag_of_incident_field(0:no reset, 1:reset)\n1 0\n" \
           "# object_data_file_name shitf_x shift_y shit_z\n%s 0.0 0.0 0.0\n" % OBJ


def _run(cmd, cwd):
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, (cmd, r.stdout[-3000:], r.stderr[-3000:])
    return r.stdout


@pytest.fixture(scope="module")
def vc_run(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("vc"))
    shutil.copy(os.path.join(GOLDEN, "2cone.msh"), d)
    for name, text in (("ipw.txt", IPW), ("medium_data.txt", MEDIUM), ("object_settings.txt", SETTINGS)):
        with open(os.path.join(d, name), "w") as fp:
            fp.write(text)
    _run([os.path.join(BIN, "d3b2_create_matrix"), "medium_data.txt", "2cone.msh", OBJ], d)
    log = _run([os.path.join(BIN, "d3b2_bv_solver"), "object_settings.txt", "vc.dat"], d)
    return d, log


def test_solver_log_matches_reference(vc_run):
    _, log = vc_run
    assert "single object" in log and "iterative operation" not in log


def test_solved_state_matches_reference_vc_dat(vc_run):
    d, _ = vc_run
    got, want = read_dat(os.path.join(d, "vc.dat")), read_dat(os.path.join(GOLDEN, "vc.dat.gz"))
    assert len(got) == 1 and got[0]["MN"] == 2
    for did in (0, 1, 2):
        for key in ("U", "dU", "E", "dE", "H", "dH"):
            g, w = got[0]["dom"][did][key], want[0]["dom"][did][key]
            err = np.abs(g - w).max() / np.abs(w).max()
            print("domain %d %-2s %.2e" % (did, key, err), flush=True)
            assert err < 1e-10, (did, key, err)


def test_reference_example1_matches_reference_log(vc_run):
    d, _ = vc_run
    out = _run([ref_example("example1"), "vc.dat"], d)
    with open(os.path.join(d, "ex1.log"), "w") as fp:
        fp.write(out)
    got = parse_example1_log(os.path.join(d, "ex1.log"))
    want = parse_example1_log(os.path.join(GOLDEN, "vc_example1.log"))
    keys = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
    for typ in (0, 1):
        w = np.array([want["field"][typ][k] for k in keys])
        g = np.array([got["field"][typ][k] for k in keys])
        print("field type %d: %.2e" % (typ, np.abs(g - w).max() / np.abs(w).max()), flush=True)
        assert np.abs(g - w).max() < 1e-9 * np.abs(w).max(), typ
        fs = np.abs(want["force"][typ][0]["F"]).max()
        for q, sc in (("F", fs), ("N", 0.3 * fs)):
            a, b = got["force"][typ][0][q], want["force"][typ][0][q]
            scale = max(np.abs(b).max(), sc)
            print("type %d %s: %.2e" % (typ, q, np.abs(a - b).max() / scale), flush=True)
            assert np.abs(a - b).max() < 1e-9 * scale, (typ, q)
        pa, pb = got["absorb"][typ][0], want["absorb"][typ][0]
        assert abs(pa - pb) < 1e-9 * pb, (typ, pa, pb)

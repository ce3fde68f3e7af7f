"""The beams of multi_fbeam other than the plane wave on the device (focused
plane wave, focused with spiral phase, Bessel, Bessel with spiral phase,
radially / azimuthally polarised; fpw.c, lgb.c, bsb.c, bslgb.c, rab.c):

* incident E, H at points and the incident boundary data (E and dE/dn at
  every open-domain node slot) against the oracle, every type and a mixed
  beam set, 1e-12 of the largest value;
* the reference's analysis_sample3 fixture (ex30.dat: a glass hemisphere of
  1008 quadrangles solved under a focused plane wave) replayed on the
  device: total field, force and torque against example1_ex30.log;
* the same case run from its input files (hemisphere_m1.msh,
  medium_data1.txt n = 1.515, fpw.txt, object_settings0.txt) through
  d3b2_create_matrix and d3b2_bv_solver: the solved state against ex30.dat
  (1e-10) and the relinked example1 against example1_ex30.log;
* the two-object case of the same sample (the hemisphere and a metal cone
  turned by pi about x, coupled) from its input files: the convergence
  trace of d3b2_bv_solver.log and the relinked example1 against
  example1_ex3.log.

The absorbed powers of example1_ex30.log are round-off of a lossless object
(the reference prints its "loss of significance" warning with them), so
they are only checked to be small against the incident power."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as orc
from bem3_emf_b2_amd import meshgen as mg
from datfile import parse_example1_log, read_dat
from fixtures import GOLDEN, ex30_log, ex30_raw, ref_example
from test_oracle_beams import BEAMS, LAM

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "bem3_emf_b2_amd", "bin")
REF_PT = [0.3, 0.0, 1.1]
SETS = {k: [v] for k, v in BEAMS.items()}
SETS["mixed"] = [[0, LAM, 1.33, 0.0, 1e-3, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.3, 0.2, 0, 0], BEAMS["fpw"], BEAMS["bsb"],
                 BEAMS["rab"]]


def _pair(beams):
    import bem3_emf_b2_amd as bem
    V, F = mg.icosphere(1, 0.4)
    arrays = mg.paired_arrays(V, F)
    b = np.array(beams, dtype=np.float64)
    o = orc.Obj.from_arrays(*arrays, [0j, 1.5])
    o.setup(b)
    rn, ed, md, sd, gd = arrays
    g = bem.Object(rn, ed, md, sd, [1.33, 1.5], LAM)
    g.set_beams(b)
    return o, g


@pytest.mark.parametrize("name", list(SETS))
def test_incident_fields_match_oracle(name):
    import bem3_emf_b2_amd as bem
    o, g = _pair(SETS[name])
    for key in ("Ui0", "dUi0"):
        want, got = o.incident(key), g.incident(key)
        assert np.abs(got - want).max() < 1e-12 * np.abs(want).max(), (name, key)
    pts = np.random.default_rng(9).uniform(-1.5, 1.5, (40, 3))
    got, gid = bem.mo_field([g], pts, "EH", "i", 0)
    want, wid = orc.mo_field([o], pts, "EH", "i", 0)
    assert (gid == wid).all()
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max(), name
    g.close()
    o.free()


@pytest.fixture(scope="module")
def ex30_device():
    import bem3_emf_b2_amd as bem
    G = []
    for o in ex30_raw():
        g = bem.Object(o["rn"], o["ed"], o["md"], o["sd"], list(o["n"]), float(o["lambda0"]))
        g.set_beams(o["mfb"])
        for d in range(o["MN"] + 1):
            for key in ("U", "dU", "E", "dE", "H", "dH"):
                g.state(d, key, o["dom"][d][key])
        G.append(g)
    yield G
    for g in G:
        g.close()


@pytest.mark.parametrize("typ", [0, 1])
def test_ex30_total_field(ex30_device, typ):
    import bem3_emf_b2_amd as bem
    ref = ex30_log()["field"][typ]
    out, oid = bem.mo_field(ex30_device, [REF_PT], "EH", "t", typ)
    want = np.array([ref[k] for k in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")])
    assert oid[0] == -1
    assert np.abs(out[0] - want).max() < 1e-10 * np.abs(want).max()


@pytest.mark.parametrize("typ", [0, 1])
def test_ex30_force_torque(ex30_device, typ):
    import bem3_emf_b2_amd as bem
    ref = ex30_log()["force"][typ][0]
    ret, F, N = bem.mo_force(ex30_device, 0, typ)
    assert ret == 0
    assert np.abs(F - ref["F"]).max() < 1e-10 * np.abs(ref["F"]).max()
    assert np.abs(N - ref["N"]).max() < 1e-10 * np.abs(ref["F"]).max()


FPW = "## focused plane wave\n# number of beams\n1\n#lambda0 ni NA power real(e0x) imag(e0x) real(e0y) imag(e0y) fx fy fz theta phi sn\n" \
      "0.6328 1.0 0.35 1.0 1.0 0.0 0.0 0.0 +0.0 0.0 -0.01 +1.1 0.0 64\n"
MEDIUM1 = "## medium data file for bem3_emf_b2\n# number of medium\n1\n# real(refractive_index) imag(refractive_index)\n" \
          "1.515 0.0\n"
SETTINGS0 = "# number_of_objects reset_flag_of_incident_field(0:no reset, 1:reset)\n1 0\n" \
            "# object_data_file_name shitf_x shift_y shit_z\nhemisphere_m1_md1.obj +0.0 +0.0 -0.0\n"


def _run(cmd, cwd):
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, (cmd, r.stdout[-3000:], r.stderr[-3000:])
    return r.stdout


@pytest.fixture(scope="module")
def ex30_run(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("ex30"))
    shutil.copy(os.path.join(GOLDEN, "hemisphere_m1.msh"), d)
    for name, text in (("fpw.txt", FPW), ("medium_data1.txt", MEDIUM1), ("object_settings0.txt", SETTINGS0)):
        with open(os.path.join(d, name), "w") as fp:
            fp.write(text)
    _run([os.path.join(BIN, "d3b2_create_matrix"), "medium_data1.txt", "hemisphere_m1.msh", "hemisphere_m1_md1.obj"], d)
    log = _run([os.path.join(BIN, "d3b2_bv_solver"), "object_settings0.txt", "ex30.dat"], d)
    return d, log


def test_ex30_from_input_files_state(ex30_run):
    d, log = ex30_run
    assert "-- focused plane wave --" in log
    got, ref = read_dat(os.path.join(d, "ex30.dat")), ex30_raw()
    assert len(got) == 1
    np.testing.assert_array_equal(got[0]["mfb"], ref[0]["mfb"])
    for did in (0, 1):
        for key in ("U", "dU", "E", "dE", "H", "dH"):
            a, b = got[0]["dom"][did][key], ref[0]["dom"][did][key]
            err = np.abs(a - b).max() / np.abs(b).max()
            print("domain %d %-2s %.2e" % (did, key, err), flush=True)
            assert err < 1e-10, (did, key, err)


def test_ex30_from_input_files_example1(ex30_run):
    d, _ = ex30_run
    out = _run([ref_example("example1"), "ex30.dat"], d)
    with open(os.path.join(d, "ex1.log"), "w") as fp:
        fp.write(out)
    got = parse_example1_log(os.path.join(d, "ex1.log"))
    want = ex30_log()
    keys = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
    for typ in (0, 1):
        w = np.array([want["field"][typ][k] for k in keys])
        g = np.array([got["field"][typ][k] for k in keys])
        assert np.abs(g - w).max() < 1e-9 * np.abs(w).max(), typ
        for q in ("F", "N"):
            a, b = got["force"][typ][0][q], want["force"][typ][0][q]
            assert np.abs(a - b).max() < 1e-9 * np.abs(want["force"][typ][0]["F"]).max(), (typ, q)
        # round-off of a lossless object against an incident power of 1
        assert abs(got["absorb"][typ][0]) < 1e-3


# analysis_sample3's two-object case: the hemisphere and a metal cone
# (cone_m2.msh, n = 0.06+4i) turned by pi about x at create_matrix time and
# shifted by 0.13 in z, under the same focused plane wave
MEDIUM2 = "## medium data file for bem3_emf_b2\n# number of medium\n1\n# real(refractive_index) imag(refractive_index)\n" \
          "0.06 4.0\n"
SETTINGS = "# number_of_objects reset_flag_of_incident_field(0:no reset, 1:reset)\n2 0\n" \
           "# object_data_file_name shitf_x shift_y shit_z\nhemisphere_m1_md1.obj +0.0 +0.0 -0.0\n" \
           "cone_m2_md2_xpi.obj +0.0 +0.0 +0.130\n"


@pytest.fixture(scope="module")
def ex3_run(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("ex3"))
    for m in ("hemisphere_m1.msh", "cone_m2.msh"):
        shutil.copy(os.path.join(GOLDEN, m), d)
    for name, text in (("fpw.txt", FPW), ("medium_data1.txt", MEDIUM1), ("medium_data2.txt", MEDIUM2),
                       ("object_settings.txt", SETTINGS)):
        with open(os.path.join(d, name), "w") as fp:
            fp.write(text)
    _run([os.path.join(BIN, "d3b2_create_matrix"), "medium_data1.txt", "hemisphere_m1.msh", "hemisphere_m1_md1.obj"], d)
    _run([os.path.join(BIN, "d3b2_create_matrix"), "medium_data2.txt", "cone_m2.msh", "cone_m2_md2_xpi.obj",
          "1", "0", "0", "3.14159265", "0", "0", "0"], d)
    log = _run([os.path.join(BIN, "d3b2_bv_solver"), "object_settings.txt", "ex3.dat"], d)
    return d, log


def test_ex3_convergence_trace(ex3_run):
    import re
    _, log = ex3_run
    tr = lambda t: {int(c): float(v) for c, v in re.findall(r"count =\s*(\d+), cc = ([0-9.eE+-]+)", t)}
    with open(os.path.join(GOLDEN, "ex3_bv_solver.log")) as fp:
        want = tr(fp.read())
    got = tr(log)
    assert sorted(got) == sorted(want), (sorted(got), sorted(want))
    for c, w in want.items():
        print("count %2d: %.6g vs %.6g" % (c, got[c], w), flush=True)
        assert abs(got[c] - w) <= 2e-5 * w + 1e-10, (c, got[c], w)


def test_ex3_example1(ex3_run):
    d, _ = ex3_run
    out = _run([ref_example("example1"), "ex3.dat"], d)
    with open(os.path.join(d, "ex1.log"), "w") as fp:
        fp.write(out)
    got = parse_example1_log(os.path.join(d, "ex1.log"))
    want = parse_example1_log(os.path.join(GOLDEN, "ex3_example1.log"))
    keys = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
    for typ in (0, 1):
        w = np.array([want["field"][typ][k] for k in keys])
        g = np.array([got["field"][typ][k] for k in keys])
        print("field type %d: %.2e" % (typ, np.abs(g - w).max() / np.abs(w).max()), flush=True)
        assert np.abs(g - w).max() < 1e-8 * np.abs(w).max(), typ
        for m in (0, 1):
            fs = np.abs(want["force"][typ][m]["F"]).max()
            for q in ("F", "N"):
                a, b = got["force"][typ][m][q], want["force"][typ][m][q]
                print("type %d object %d %s: %.2e" % (typ, m, q, np.abs(a - b).max() / fs), flush=True)
                assert np.abs(a - b).max() < 1e-8 * fs, (typ, m, q)

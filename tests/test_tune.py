"""The tuning table's host side (no GPU needed: nothing reaches the device):
names and keys agree between the header, the library and the Python front
end, files load with their entries in force, bad files are refused, and the
committed table of the package loads."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lib():
    import bem3_emf_b2_amd as bem
    if not os.path.exists(bem.LIB_PATH):
        pytest.skip("library not built")
    return bem


@pytest.fixture
def bem_defaults(tmp_path):
    bem = _lib()
    yield bem
    empty = tmp_path / "empty.txt"
    empty.write_text("")
    bem.tune_load(str(empty))   # every entry back at its default


def test_keys_agree_with_header_and_library(bem_defaults):
    bem = bem_defaults
    hdr = open(os.path.join(ROOT, "include", "bem3_hip.h")).read()
    keys = dict(re.findall(r"#define BH_TUNE_([A-Z0-9_]+)\s+(\d+)", hdr))
    assert int(keys.pop("COUNT")) == len(bem.TUNE_KEYS)
    assert sorted(int(v) for v in keys.values()) == sorted(bem.TUNE_KEYS.values())
    for name, k in bem.TUNE_KEYS.items():
        assert bem.lib().bh_tune_name(k).decode() == name
    assert bem.lib().bh_tune_name(len(bem.TUNE_KEYS)) is None


def test_load_set_and_refuse(bem_defaults, tmp_path):
    bem = bem_defaults
    f = tmp_path / "t.txt"
    f.write_text("# comment line\nlanes 3   # trailing comment\n\ngemm_tiles_fp64 200\n")
    bem.tune_load(str(f))
    assert bem.tune_source() == str(f)
    assert bem.tune_get("lanes") == 3 and bem.tune_get("gemm_tiles_fp64") == 200
    # entries the file does not name take their defaults
    assert bem.tune_get("asm_batch") == 1 and bem.tune_get("gemm_bf6") == 1
    bem.tune_set("gemm_longk_k", 512)
    assert bem.tune_get("gemm_longk_k") == 512
    for bad in ("lanes 9\n", "lanes\n", "no_such_entry 1\n", "lanes 2 3\n", "gemm_longk_k 16\n"):
        f.write_text(bad)
        with pytest.raises(bem.BemError):
            bem.tune_load(str(f))
    # a refused file leaves the table as it was
    assert bem.tune_get("lanes") == 3
    with pytest.raises(bem.BemError):
        bem.tune_set("lanes", 0)
    with pytest.raises(bem.BemError):
        bem.tune_load(str(tmp_path / "missing.txt"))


def test_committed_table_loads(bem_defaults):
    bem = bem_defaults
    path = os.path.join(os.path.dirname(bem.LIB_PATH), "tune_gfx950.txt")
    if not os.path.exists(path):
        pytest.skip("no tuning table committed")
    bem.tune_load(path)
    for name in bem.TUNE_KEYS:
        bem.tune_get(name)


def test_autotuner_threshold_fit():
    """the fit of the dispatch rule on a synthetic grid: the 128x128 kernel
    wins from 64 tiles in FP64 and from 36 in FP32, and with k >= 1024
    already from 16 tiles in both; the fit finds exactly those thresholds,
    and keeps the defaults when they are already optimal"""
    _lib()
    from bem3_emf_b2_amd import autotune as at
    Ts, ks = [9, 16, 25, 36, 64, 100, 144], [64, 256, 1024, 2048]

    def grid(t1):
        return [(T, k, 1.0, 0.5 if (T >= t1 or (T >= 16 and k >= 1024)) else 2.0) for T in Ts for k in ks]

    rows = {"fp64": grid(64), "fp32": grid(36)}
    t1s, (t2, kl), c, c0 = at.fit_thresholds(rows, ({"fp64": 160, "fp32": 160}, (64, 2048)))
    assert t1s == {"fp64": 64, "fp32": 36} and t2 == 16 and kl == 1024
    assert c < c0
    t1s, longk, c, c0 = at.fit_thresholds(rows, ({"fp64": 64, "fp32": 36}, (16, 1024)))
    assert t1s == {"fp64": 64, "fp32": 36} and longk == (16, 1024) and c == c0
    assert at.rule_big(64, 32, 64, 16, 1024) and not at.rule_big(63, 1023, 64, 16, 1024)

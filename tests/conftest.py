import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


# long GPU runs append the coupling loop's progress lines to a file under
# gpurun_out/ (pytest holds back the test's own output until it ends)
PROGRESS = os.path.join(ROOT, "gpurun_out", "test_progress.log")


def progress(msg):
    """a line in the progress file (and on stdout)"""
    os.makedirs(os.path.dirname(PROGRESS), exist_ok=True)
    with open(PROGRESS, "a") as f:
        f.write(msg + "\n")
    print(msg, flush=True)


def pytest_configure(config):
    os.makedirs(os.path.dirname(PROGRESS), exist_ok=True)
    os.environ.setdefault("BEM3_HIP_PROGRESS", PROGRESS)
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def golden():
    return GOLDEN

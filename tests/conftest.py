import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
# test_oracle_mutants.py runs the oracle checks against planted-mistake copies
# of the oracle package: the directory that holds such a copy goes first
if os.environ.get("VLS_ORACLE_ROOT"):
    sys.path.insert(0, os.environ["VLS_ORACLE_ROOT"])


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on cuda:0)")

"""Shared fixtures: the CPU oracle (checker) and the product library."""
import ctypes
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
ORACLE_KB_LIB = os.path.join(ORACLE_DIR, "liboracle_kb.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_oracle(path=ORACLE_LIB):
    """The F65521 oracle (liboracle.so) or, with ORACLE_KB_LIB, the KoalaBear one."""
    if not os.path.exists(path):
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
    L = ctypes.CDLL(path)
    L.orc_pcs_new.restype = ctypes.c_void_p
    L.orc_pcs_free.argtypes = [ctypes.c_void_p]
    L.orc_last_error.restype = ctypes.c_char_p
    return L


@pytest.fixture(scope="session")
def oracle():
    return load_oracle()


@pytest.fixture(scope="session")
def oracle_kb():
    return load_oracle(ORACLE_KB_LIB)


@pytest.fixture(scope="session")
def product():
    import parcomm_amd
    return parcomm_amd.lib()

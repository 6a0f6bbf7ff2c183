"""Test configuration: the ``gpu`` marker and shared fixtures.

CPU tests (``-m "not gpu"``) pin the oracle against the golden vectors written by the
compiled reference, check invariants, and check that the native library loads and
exports its C ABI. GPU tests (``-m gpu``) compare the HIP path with the oracle.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle.oracle import Lib
    return Lib("oracle")


@pytest.fixture(scope="session")
def ref_lib():
    from oracle.oracle import Lib, have_ref
    if not have_ref():
        pytest.skip("oracle/_ref is not built (needs the reference tree)")
    return Lib("ref")


@pytest.fixture(scope="session")
def gpu():
    """The native module, after checking that a HIP device is present."""
    from ctf_amd import native
    native.lib()
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_int(0)
    if hip.hipGetDeviceCount(ctypes.byref(n)) != 0 or n.value < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return native

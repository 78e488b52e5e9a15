import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

REFERENCE_TXT = "/root/reference/cpp/tests/runtime/case_regression_reference.txt"
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP engine)")
    config.addinivalue_line("markers", "slow: long-running parity case")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.load("port")
    return O


@pytest.fixture(scope="session")
def reference_oracle():
    from oracle import oracle as O

    if not O.available("reference"):
        pytest.skip("oracle/_ref/libwaslref.so is not built here (no /root/reference)")
    O.load("reference")
    return O


@pytest.fixture(scope="session")
def wasl():
    import wasl_amd

    wasl_amd.library()
    return wasl_amd


@pytest.fixture(scope="session")
def gpu(wasl):
    import ctypes

    # Fail loudly rather than fall back: a GPU test without a device is an error
    # on the GPU box; here (no device) the gpu marker is deselected.
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        rc = hip.hipGetDeviceCount(ctypes.byref(n))
    except OSError as exc:
        pytest.fail(f"HIP runtime not loadable: {exc}")
    if rc != 0 or n.value < 1:
        pytest.fail("no HIP device visible for a gpu-marked test")
    return wasl

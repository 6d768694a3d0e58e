import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libplsqr_hip.so on the GPU)")
    config.addinivalue_line("markers", "slow: large inputs (67^3 grid, 2M rays)")


def load_pkg():
    name = "plsqr_amd"
    if name in sys.modules:
        return sys.modules[name]
    pkg = os.path.join(ROOT, "plsqr3.2013.09_amd")
    spec = importlib.util.spec_from_file_location(name, os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        import oracle  # noqa: E402
    finally:
        sys.path.pop(0)
    return oracle


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_cases():
    sys.path.insert(0, GOLDEN)
    try:
        import make_golden  # noqa: E402
    finally:
        sys.path.pop(0)
    return make_golden


@pytest.fixture(scope="session")
def P():
    return load_pkg()


@pytest.fixture(scope="session")
def O():
    return load_oracle()


@pytest.fixture(scope="session")
def MG():
    return golden_cases()


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm(a - b)
    nb = np.linalg.norm(b)
    return d / nb if nb > 0 else d


@pytest.fixture(scope="session")
def gpu_available():
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        cnt = ctypes.c_int(0)
        ok = hip.hipGetDeviceCount(ctypes.byref(cnt)) == 0 and cnt.value > 0
    except OSError:
        ok = False
    if not ok:
        pytest.fail("no GPU visible: the gpu-marked tests need an MI355X")
    return True

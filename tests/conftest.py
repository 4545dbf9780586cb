import importlib.util
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "scaling-subsequence-similarity-join-based-on-dynamic-time-warping-cpu-gpu_amd")
GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def load_pkg():
    if "dtwj_amd" in sys.modules:
        return sys.modules["dtwj_amd"]
    spec = importlib.util.spec_from_file_location("dtwj_amd", os.path.join(PKG_DIR, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["dtwj_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def golden_cases():
    path = os.path.join(GOLDEN_DIR, "golden.json")
    with open(path) as f:
        db = json.load(f)
    return db


def golden_series(name):
    return np.load(os.path.join(GOLDEN_DIR, name + ".npy"), allow_pickle=False)


@pytest.fixture(scope="session")
def dtwj():
    return load_pkg()


@pytest.fixture(scope="session")
def oracle():
    import oracle_ctypes
    oracle_ctypes.lib()
    return oracle_ctypes


@pytest.fixture(scope="session")
def gpu(dtwj):
    if dtwj.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests must run on the MI355X box")
    return dtwj

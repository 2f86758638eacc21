"""Shared fixtures: loads the product bindings (buff-em_amd/) and the oracle."""
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running")


def load_bfa():
    spec = importlib.util.spec_from_file_location("buffem_amd", os.path.join(ROOT, "buff-em_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def bfa():
    return load_bfa()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


SEED = 0x5EED


def kuhn(bfa, n, R=1.0):
    return bfa.kuhn_sphere(n, R, 0.05, SEED)

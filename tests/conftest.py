"""Shared test setup.

The package directory name contains hyphens, so it is registered under the
import name `qumasun_amd`.  The oracle (oracle/, CPU FP64 restatement of the
reference) is importable as `oracle` for checking only.
"""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "qumasun-workshop-template_amd")


def _load_pkg():
    if "qumasun_amd" in sys.modules:
        return sys.modules["qumasun_amd"]
    spec = importlib.util.spec_from_file_location("qumasun_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["qumasun_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


_load_pkg()
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def qm():
    return _load_pkg()


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch

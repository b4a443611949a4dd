import importlib.util
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(REPO, "raw2png-conversionthrugpu_amd")
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def load_pkg():
    """Imports the encoder package (its directory name has a hyphen)."""
    name = "raw2png_conversionthrugpu_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def r2p():
    return load_pkg()


@pytest.fixture(scope="session")
def synth():
    load_pkg()
    import importlib
    return importlib.import_module("raw2png_conversionthrugpu_amd.synth")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    return O


@pytest.fixture(scope="session")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a HIP device")
    return torch

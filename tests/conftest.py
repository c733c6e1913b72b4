import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (MI355X) and the built HIP library")
    # torch carries its own HIP runtime beside the engine's: let torch
    # enumerate the devices before the engine's runtime starts, or torch
    # finds no device afterwards (counting devices does not initialise one)
    try:
        import torch
        torch.cuda.device_count()
    except Exception:
        pass
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-f", "oracle/Makefile"], cwd=ROOT, check=True)


_CASES = {}


@pytest.fixture(scope="session")
def case_factory(tmp_path_factory):
    """case name -> (folder, parameter vectors), built once per session."""
    import cases

    def get(name):
        if name not in _CASES:
            folder = str(tmp_path_factory.mktemp(name))
            _CASES[name] = cases.build(name, folder)
        return _CASES[name]
    return get


def gpu_available():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False

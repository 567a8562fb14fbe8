"""Shared test fixtures: the product package (HIP library) and the CPU oracle (checker)."""
import importlib.util
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _load(name, path, search=None):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, path, submodule_search_locations=[str(search)] if search else None)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_package():
    return _load("pbrt_v2_amd", ROOT / "pbrt-v2_amd" / "__init__.py", ROOT / "pbrt-v2_amd")


def load_scenes():
    load_package()
    return _load("pbrt_v2_amd.scenes", ROOT / "pbrt-v2_amd" / "scenes.py")


def load_oracle():
    return _load("rt_oracle_py", ROOT / "oracle" / "oracle.py")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    # torch first: its HIP runtime must be the process's before the product libraries (and
    # the C++ test driver, which links them) load the system one, whatever the test order
    import torch  # noqa: F401


@pytest.fixture(scope="session")
def orc():
    o = load_oracle()
    o.lib()
    return o


@pytest.fixture(scope="session")
def scenes():
    return load_scenes()


@pytest.fixture(scope="session")
def rt():
    return load_package()


@pytest.fixture(scope="session")
def gpu_ctx(rt):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    ctx = rt.Context(0)
    yield ctx
    ctx.close()

import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_amaru():
    """Import the product package (directory ``amaru.jl_amd``) as ``amaru_amd``."""
    if "amaru_amd" in sys.modules:
        return sys.modules["amaru_amd"]
    pkg = os.path.join(ROOT, "amaru.jl_amd")
    spec = importlib.util.spec_from_file_location("amaru_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["amaru_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def amaru():
    return load_amaru()


@pytest.fixture(scope="session")
def pins():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "reference_pins.json")) as f:
        return json.load(f)

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "gradpower.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


_CASES = {}


def get_case(name, **kw):
    """Synthetic cases are deterministic; build each once per session."""
    from gradpower_hip import synth
    key = (name, tuple(sorted(kw.items())))
    if key not in _CASES:
        _CASES[key] = getattr(synth, name)(**kw)
    return _CASES[key]


@pytest.fixture(scope="session")
def case_getter():
    return get_case

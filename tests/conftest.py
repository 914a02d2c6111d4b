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


# Named variants of the synthetic builders.  case2_60bus: the 2,000-bus
# workload's generator at 60 buses, 12 machines (SEXS/ESDC1A exciters,
# TGOV1/IEESGO governors), 4 IEEEST stabilisers and 6 PV static generator
# stubs -- every device type in a case small enough for per-step checks.
_VARIANTS = {"case2_60bus": ("case2_grid", dict(M=8, nbus=60, n_genrou=12, n_pss=4, n_static=6, seed=77))}


def get_case(name, **kw):
    """Synthetic cases are deterministic; build each once per session."""
    from gradpower_hip import synth
    if name in _VARIANTS:
        name, base = _VARIANTS[name]
        kw = dict(base, **kw)
    key = (name, tuple(sorted(kw.items())))
    if key not in _CASES:
        _CASES[key] = getattr(synth, name)(**kw)
    return _CASES[key]


@pytest.fixture(scope="session")
def case_getter():
    return get_case

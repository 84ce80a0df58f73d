import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    try:
        import viss_recon_shade_amd as v
        return v.load().vrs_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o._load()
    return o


@pytest.fixture(scope="session")
def ctx():
    import viss_recon_shade_amd as v
    if not gpu_available():
        pytest.fail("GPU test collected but no HIP device is visible")
    c = v.Context(0)
    yield c
    c.close()


# Accuracy figures that GPU tests report, printed as one section at the end of
# the run (so they appear in the tail of a quiet `pytest -q` log as well).
_FIGURES = []


@pytest.fixture(scope="session")
def report():
    def add(name, **vals):
        _FIGURES.append((name, vals))
    return add


def pytest_terminal_summary(terminalreporter, exitstatus, config):
    if not _FIGURES:
        return
    terminalreporter.section("parity figures")
    for name, vals in _FIGURES:
        body = ", ".join(f"{k}={v:.3g}" if isinstance(v, float) else f"{k}={v}"
                         for k, v in vals.items())
        terminalreporter.write_line(f"{name}: {body}")

"""Test configuration: the `gpu` marker, import paths, and shared helpers.

CPU tests (-m "not gpu") check the oracle against the reference's known answers, closed
forms and brute force, the data generators, and that the C ABI library loads and exports
every symbol the public headers declare.  GPU tests (-m gpu) call the HIP engine through its C ABI and
compare with the oracle on the same inputs.
"""
import os
import sys
import pathlib

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "parallax-runtime_amd", ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def has_gpu():
    try:
        from parallax_amd.device import device_count
        return device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu_ctx():
    """One stream context for the whole GPU session (tests run in one process)."""
    if os.environ.get("PARALLAX_FORCE_CPU"):
        pytest.fail("GPU test running with PARALLAX_FORCE_CPU set (it would measure nothing)")
    from parallax_amd.device import Context, device_count
    if device_count() == 0:
        pytest.fail("GPU test collected on a machine without a GPU")
    ctx = Context()
    yield ctx
    ctx.close()

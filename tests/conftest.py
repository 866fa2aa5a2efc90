import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libygrid.so")


@pytest.fixture(scope="session")
def gpu_ctx():
    """A libygrid context on cuda:0.  GPU tests fail (not skip) when the library is missing."""
    import yandasoft_amd as ya
    ya.lib()
    ctx = ya.Context(0)
    yield ctx
    ctx.close()


def announce(msg: str):
    """Print the case name before a GPU step so that a failure is attributable."""
    print(f"[case] {msg}", flush=True)

"""Shared fixtures.  GPU tests carry @pytest.mark.gpu and only run on a box
with an MI355X; everything else runs on the CPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle_lib():
    """The CPU oracle, built on demand (test infrastructure only)."""
    import oracle
    if not os.path.exists(oracle.LIB):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def gpu_ctx():
    """One HIP context for the whole GPU session (fails loudly without the
    built library or a device)."""
    from seiscl_amd import Context
    ctx = Context(0)
    yield ctx
    ctx.close()

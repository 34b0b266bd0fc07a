import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# The C++ restatement runs OpenMP loops; with several test processes on the
# same cores (pytest -n), spinning idle threads slowed its tests forty-fold.
os.environ.setdefault("OMP_WAIT_POLICY", "PASSIVE")


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def ref_lib():
    """The C++ restatement, built on first use."""
    from oracle import ref_lib as rl
    if not os.path.exists(rl._LIB_PATH):
        rl.build()
    rl.load()
    return rl

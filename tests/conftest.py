import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from fs2_amd import build as _build  # noqa: E402

_build.ensure_built()   # a fresh checkout: compile once; otherwise load what is built


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def fs():
    import fs2_amd.fs as fs_mod
    return fs_mod

import os
import re
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


# The parity figures the GPU tests print ("[case] ... rel RMS ...") are
# captured by pytest for passing tests; they are collected here and listed in
# the terminal summary, so a quiet run (-q) still reports them.
_FIGURES = []
_FIGURE_KEYS = ("rel RMS", "error", "rms ", "gradient")
_NUMBER = re.compile(r"\d\.\d")


def pytest_runtest_logreport(report):
    if report.when == "call" and report.passed:
        for line in report.capstdout.splitlines():
            if (line.startswith("[case]") and any(k in line for k in _FIGURE_KEYS)
                    and _NUMBER.search(line)):
                _FIGURES.append(line[len("[case] "):])


def pytest_terminal_summary(terminalreporter):
    if _FIGURES:
        terminalreporter.write_sep("-", "parity figures of the passing tests")
        for line in _FIGURES:
            terminalreporter.write_line(line)

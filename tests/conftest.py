import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

# Measured figures that tests hand to the `figures` fixture; printed in the
# terminal summary, so that they appear under `pytest -q` as well.
_FIGURES = []


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture
def figures(request):
    """figures(text): record one line of measured figures for the summary."""
    def add(text):
        line = "%s: %s" % (request.node.name, text)
        _FIGURES.append(line)
        print(line, flush=True)
    return add


def pytest_terminal_summary(terminalreporter, exitstatus, config):
    if _FIGURES:
        terminalreporter.section("measured figures")
        for line in _FIGURES:
            terminalreporter.write_line(line)

import atexit
import os
import shutil
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle"), GOLDEN):
    if p not in sys.path:
        sys.path.insert(0, p)

import render_decks  # noqa: E402

# The seven regression cases are stored as data (tests/golden/regression_cases.json)
# and rendered into deck files once per test session.
DECKS = tempfile.mkdtemp(prefix="hlbm_regression_decks_")
REGRESSION_DECKS = render_decks.write_decks(DECKS)
atexit.register(shutil.rmtree, DECKS, True)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def decks_dir():
    return DECKS

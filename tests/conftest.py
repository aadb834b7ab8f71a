import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line(
        "markers", "slow: full-size cases (minutes)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture
def measured(request):
    """record(quantity, value, tolerance): prints the measured error beside
    its bound and, when PACMAN_AMD_MEASURED_LOG names a file, appends
    "test<TAB>quantity<TAB>value<TAB>tolerance" to it."""
    path = os.environ.get("PACMAN_AMD_MEASURED_LOG")

    def record(quantity, value, tolerance):
        line = f"{request.node.nodeid}\t{quantity}\t{value:.3e}\t{tolerance:.1e}"
        print(f"[measured] {line}", flush=True)
        if path:
            with open(path, "a") as fh:
                fh.write(line + "\n")
    return record

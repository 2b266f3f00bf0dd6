import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def gpu():
    """Skips when no HIP device is visible; otherwise names the device."""
    import comd_amd

    n, name = comd_amd.device_info()
    if n < 1:
        pytest.skip("no HIP device")
    print(f"[device] {name}", flush=True)
    return name


@pytest.fixture(scope="session")
def cu500():
    import comd_amd

    return comd_amd.EAMPotential("cu500")


@pytest.fixture(scope="session")
def ocu500():
    import oracle_py

    return oracle_py.OraclePot.eam(os.path.join(ROOT, "comd_amd", "pots"), "cu500")

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "gpu-comcot_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "slow: long CPU reference runs")


def bits(a):
    """Bit patterns of a float32 array, for exact comparisons (0.0 != -0.0)."""
    import numpy as np
    return np.ascontiguousarray(a, dtype=np.float32).view(np.int32)


@pytest.fixture(scope="session")
def oracle_built():
    from oracle import bind
    if not os.path.exists(bind.ORACLE_SO):
        bind.build(ref=False)
    return bind

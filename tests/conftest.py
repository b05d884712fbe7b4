import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "k-litho_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running CPU check")


def relerr(a, b) -> float:
    """Relative L2 (Frobenius) error of a against b.  Complex inputs keep their
    imaginary parts (the difference is taken in complex128)."""
    a = np.asarray(a)
    b = np.asarray(b)
    dt = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a = a.astype(dt)
    b = b.astype(dt)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="session")
def ctx():
    """One libklt context for all GPU tests of the session."""
    import klt
    if klt.lib().klt_device_count() < 1:
        pytest.fail("no GPU visible to libklt")
    c = klt.Context(0)
    yield c
    c.close()

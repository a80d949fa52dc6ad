import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(scope="session")
def H():
    import hw_amd
    return hw_amd


@pytest.fixture(scope="session")
def O():
    import oracle
    oracle.threads(min(8, os.cpu_count() or 1))
    return oracle


@pytest.fixture(scope="session")
def T():
    from hw_amd import types
    return types


@pytest.fixture(scope="session")
def gpu(H):
    if H.device_count() == 0:
        pytest.skip("no GPU")
    return True


def gen_traces(O, T, H, lanes, frames, chunks, seed=7, **kw):
    """Seeded synthetic traces from the oracle's generator: list of [lanes, frames, 64] int16."""
    p = H.tracegen_default(seed=seed, **kw)
    st = np.zeros(lanes, dtype=T.TRACEGEN_STATE)
    return [O.tracegen(p, lanes, frames, c * frames, st, c == 0) for c in range(chunks)]

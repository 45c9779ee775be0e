import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "bc1_golden.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (run on the MI355X box)")


@pytest.fixture(scope="session")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle, build
    from oracle.binding import ORACLE_SO
    if not os.path.exists(ORACLE_SO):
        build(with_reference=False)
    o = Oracle()
    o.init(0)
    return o


@pytest.fixture(scope="session")
def reference():
    from oracle import Reference, reference_available
    if not reference_available("scalar"):
        pytest.skip("reference build (oracle/_ref) not present")
    r = Reference("scalar")
    r.init(0)
    return r


@pytest.fixture(scope="session")
def gpu():
    """The encoder library with the tables initialised for the D3D10 decoder."""
    import icbc_amd
    icbc_amd.init(icbc_amd.Decoder.D3D10)
    return icbc_amd

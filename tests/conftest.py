import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")

# Plain-method workload on a 32^3 logical grid (BCC-seeded Pmmm fields).
CFG0 = dict(N=32, fA=0.25, xN=20.0, L=4.0, ds=0.01, acc=0.1, nstep=100)
# Plain method on the GPU: 64^3 logical grid, Ns = 200.
CFG1 = dict(N=64, fA=0.25, xN=20.0, L=4.0, ds=0.005, acc=0.1, nstep=200)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def golden():
    return {
        "transforms": np.load(os.path.join(GOLDEN, "transforms.npz")),
        "diffusion": np.load(os.path.join(GOLDEN, "diffusion.npz")),
    }


@pytest.fixture(scope="session")
def golden_scft():
    """The oracle's 200-iteration SCFT traces of the 64^3 workload
    (tests/golden/make_scft_traces.py)."""
    return np.load(os.path.join(GOLDEN, "scft_cfg1.npz"))


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oc():
    import ct_oracle
    ct_oracle.lib()
    return ct_oracle


@pytest.fixture(scope="session")
def ct():
    import cosmictiger_amd
    cosmictiger_amd.lib()
    return cosmictiger_amd


@pytest.fixture(scope="session")
def gpu(ct):
    """Initialise device 0; the product path must load and run the HIP library."""
    ct.init(0)
    return ct


def uniform_box(n, seed):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32) for _ in range(3)]


def clustered_box(n, seed, frac=0.7, nhalo=20):
    """Uniform background plus NFW-like clumps (test sizes)."""
    from cosmictiger_amd.workloads import nfw_box
    return nfw_box(n, seed, clustered_fraction=frac, nhalos=nhalo)

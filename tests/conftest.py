import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "palmtracer-cpp_amd"), GOLDEN, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def ref_inputs():
    return np.load(os.path.join(GOLDEN, "reference_inputs.npz"))


@pytest.fixture(scope="session")
def ref_csv():
    return np.load(os.path.join(GOLDEN, "reference_csv.npz"))


@pytest.fixture(scope="session")
def ref_vectors():
    return np.load(os.path.join(GOLDEN, "ref_vectors.npz"))


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def pt_gpu():
    """The HIP library on a GPU; loud failure when the build is missing."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import palmtracer_amd
    palmtracer_amd.lib()
    return palmtracer_amd

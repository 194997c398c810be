import json
import os
import sys

import pytest

# The reference build is bit-deterministic only single-threaded (its gradient
# scatter uses OpenMP atomics); it is loaded lazily, after this is set.
os.environ.setdefault("OMP_NUM_THREADS", "1")

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(HERE, "golden"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def golden(name):
    with open(os.path.join(HERE, "golden", "ref_%s.json" % name)) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def gpu():
    """Loads the HIP library on cuda:0; fails loudly if either is missing."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected on a machine without a GPU")
    import minim_amd

    return minim_amd

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libddx kernels)")


def golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def deck(name):
    from ddpcm_amd.inputs import read_input
    return read_input(os.path.join(GOLDEN, "inputs", name + ".txt"))


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    return oracle


def hsnorm_rms(x, lmax):
    """rms over spheres of sqrt(sum_l 1/(l+1) sum_m x_lm^2) (hnorm, matvec.f90:281-323)."""
    x = np.asarray(x)
    l = np.concatenate([[l] * (2 * l + 1) for l in range(lmax + 1)])
    u = np.sqrt((x * x / (1.0 + l)).sum(axis=1))
    return float(np.sqrt((u * u).mean()))

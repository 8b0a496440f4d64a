import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "sbd-pub_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
DATA = os.path.join(GOLDEN, "data")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long CPU test")


def molecule(name, subset):
    """(norb, I0, I1, I2, strings) of a reference data set, e.g. ('h2o', '1em3')."""
    from sbdx import io
    fd = io.read_fcidump(os.path.join(DATA, f"{name}_fcidump.txt.gz"))
    I0, I1, I2 = io.setup_integrals(fd)
    s = io.read_strings(os.path.join(DATA, f"{name}-{subset}-alpha.txt"), fd["norb"])
    return dict(norb=fd["norb"], nelec=fd["nelec"], I0=I0, I1=I1, I2=I2, astr=s, bstr=s.copy())


def random_vector(n, seed):
    v = np.random.default_rng(seed).standard_normal(n)
    return v / np.linalg.norm(v)


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    return oracle

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def rel_err(got, ref):
    """Per-element relative error; entries exactly zero in ref must be exactly zero in got."""
    import numpy as np
    got = np.asarray(got)
    ref = np.asarray(ref)
    zero = ref == 0.0
    assert np.all(got[zero] == 0.0), "entries exactly zero in the reference must be exactly zero"
    nz = ~zero
    if not nz.any():
        return 0.0
    return float(np.max(np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])))


@pytest.fixture(scope="session")
def wl():
    import wloop_amd
    return wloop_amd

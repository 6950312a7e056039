import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: full-size parity case (minutes of CPU oracle work)")


@pytest.fixture(scope="session")
def si8():
    import json

    import numpy as np
    here = os.path.join(ROOT, "tests", "golden")
    psi = np.load(os.path.join(here, "si8_psi.npy"))
    with open(os.path.join(here, "si8_reference.json")) as f:
        ref = json.load(f)
    return psi, ref

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libmdx kernels)")
    config.addinivalue_line("markers", "slow: longer CPU checks")


def load_golden(name):
    """Fixture written by tests/golden/make_golden.py from the compiled reference."""
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_state(fx):
    from mdacp_amd.workloads import State
    return State(fx["q0"].copy(), fx["p0"].copy(), fx["L"].copy(), float(fx["dt"]))


GOLDEN_CASES = ["fcc4000", "fcc4000_pert", "fcc500_rho0.7_pert", "fcc500_rho0.9_pert",
                "fcc864_rho0.3_pert"]

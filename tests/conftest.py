import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X / gfx950)")
    config.addinivalue_line("markers", "slow: large-grid property checks")


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def golden_meta():
    return json.loads((GOLDEN / "golden_meta.json").read_text())


def load_golden(name):
    with np.load(GOLDEN / name, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def dev():
    """cuda:0, plus the library loaded through torch's HIP runtime.  Fails
    loudly (not skip) when a device exists but the library is not built."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import stencil_solver_amd as ss

    ss.lib()
    return torch.device("cuda:0")

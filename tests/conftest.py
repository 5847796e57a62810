import importlib
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs through the HIP library)")


@pytest.fixture(scope="session")
def lc():
    """The product package (HIP library via the C ABI)."""
    return importlib.import_module("lion-codes_amd")


@pytest.fixture(scope="session")
def wl():
    return importlib.import_module("lion-codes_amd.workloads")


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle (checker only)."""
    from oracle import oracle
    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test without a visible GPU")
    return torch.device("cuda:0")


def golden(name: str) -> dict:
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def bits(a) -> np.ndarray:
    """Bit patterns of a float/complex array, for bit-exact comparisons."""
    a = np.ascontiguousarray(a)
    if a.dtype == np.complex128:
        a = a.view(np.float64)
    return a.view(np.uint64)

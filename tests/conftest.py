import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def golden_dir():
    return ROOT / "tests" / "golden"


@pytest.fixture(scope="session")
def elt4(golden_dir):
    """tests/datasets/graphs/4elt.npz of the reference (a data fixture)."""
    import numpy as np
    z = np.load(golden_dir / "4elt.npz")
    indptr = z["indptr"].astype(np.int64)
    dest = z["indices"]
    w = z["data"]
    fe = indptr[:-1]
    od = (indptr[1:] - indptr[:-1]).astype(np.int32)
    return {"indptr": indptr, "dest": dest, "weight": w, "fe": fe, "od": od,
            "n": int(z["shape"][0])}

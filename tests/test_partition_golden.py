"""The oracle's binary conversion (validation_oracle.to_bin, the checker of
the accuracy indices) against the reference's golden vectors for
convertClusterStringToBin (MyML/tests/MyML/utils/test_partition.py:46-80,
held in tests/golden/partition_formats.json)."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle.validation_oracle import to_bin

GOLD = json.loads((Path(__file__).parent / "golden" / "partition_formats.json").read_text())


@pytest.mark.parametrize("case", ["string_to_bin_min0", "string_to_bin_min1"])
def test_to_bin_golden(case):
    g = GOLD[case]
    out = to_bin(np.array(g["labels"], np.int64))
    assert np.array_equal(out, np.array(g["bin"], np.float32))


def test_to_bin_rejects_gaps():
    with pytest.raises(ValueError):
        to_bin(np.array([0, 3, 3, 5]))
    with pytest.raises(IndexError):
        to_bin(np.array([0, 0, 2, 2]))  # passes the count check, then out of range

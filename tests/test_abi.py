"""The C-ABI library loads and exports every symbol include/eac_hip.h declares
(no compute calls, so this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

from masters_final_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    text = (ROOT / "include" / "eac_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(eac_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "eac_boruvka_mst" in names and "eac_scan_exclusive_i32" in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load_library()
    for name in _declared():
        assert hasattr(lib, name), name
        assert isinstance(getattr(lib, name), ctypes._CFuncPtr)


def test_ctypes_signatures_cover_header():
    assert set(_declared()) == set(_lib.SIGNATURES)


def test_workspace_queries_and_errors_without_gpu():
    lib = _lib.load_library()
    assert lib.eac_abi_version() == 1
    assert lib.eac_scan_workspace_bytes(10_000_000, 4) > 0
    assert lib.eac_boruvka_workspace_bytes(1_000_000, 27_000_000) > 27_000_000 * 8
    # argument errors are reported before any device work
    rc = lib.eac_boruvka_mst(None, None, 99, None, None, 10, 10, 1, None, None, None, None,
                             None, 0, None)
    assert rc == _lib.EAC_ERR_ARG
    assert b"dtype" in lib.eac_last_error()


def test_no_cpu_fallback_without_gpu():
    """Product calls raise instead of computing on the CPU when no GPU is visible."""
    import numpy as np
    import pytest
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from masters_final_amd.graph.mst import boruvka_minho_seq
    with pytest.raises(_lib.EACUnavailable):
        boruvka_minho_seq(np.array([1, 0], np.int32), np.array([1, 1], np.int32),
                          np.array([0, 1], np.int32), np.array([1, 1], np.int32))


def test_tuning_knobs_roundtrip():
    from masters_final_amd import tune
    from masters_final_amd._lib import EACError
    before = tune.get_knobs()
    try:
        tune.set_knobs(light_x10=25, first_batch=9, filter_bpcu=3, co_window=4096,
                       level_step=8, co_unroll=8, co_ranges=1, km_waves=2)
        assert tune.get_knobs() == {"light_x10": 25, "first_batch": 9, "filter_bpcu": 3,
                                    "co_window": 4096, "level_step": 8, "co_unroll": 8,
                                    "co_ranges": 1, "km_waves": 2}
        tune.set_knobs(light_x10=-1)
        assert tune.get_knobs()["light_x10"] == -1
        lib = __import__("masters_final_amd._lib", fromlist=["x"]).load_library()
        assert lib.eac_tune_set(99, 1) != 0
    finally:
        tune.set_knobs(**before)


def test_round_bound_counts_weight_levels():
    """bits(n-1) + 3 rounds per weight level of the Borůvka split: one level
    for sparse graphs, up to four for dense ones, one split with the level
    step knob at -1, none with the split off."""
    from masters_final_amd import tune
    lib = _lib.load_library()
    before = tune.get_knobs()
    try:
        tune.set_knobs(light_x10=0, level_step=0)
        assert lib.eac_mst_round_bound(1 << 20, 1 << 21, 1) == 23  # 2 entries/vertex
        assert lib.eac_mst_round_bound(1 << 20, 64 << 20, 1) == 4 * 23
        # 32 entries/vertex: fractions 3/32, 12/32, 48/32 -> three levels
        assert lib.eac_mst_round_bound(1 << 20, 32 << 20, 1) == 3 * 23
        tune.set_knobs(level_step=-1)
        assert lib.eac_mst_round_bound(1 << 20, 40 << 20, 1) == 2 * 23
        tune.set_knobs(light_x10=-1, level_step=0)
        assert lib.eac_mst_round_bound(1 << 20, 40 << 20, 1) == 23
        assert lib.eac_mst_round_bound(0, 10, 1) == 0
    finally:
        tune.set_knobs(**before)


def test_one_hip_runtime_when_the_library_loads_first():
    """Loading the library before torch leaves one HIP runtime in the
    process (torch's), not a second copy of it."""
    import subprocess
    import sys
    code = ("from masters_final_amd import _lib; _lib.load_library(); import torch; "
            "print(sorted({l.split()[-1] for l in open('/proc/self/maps') "
            "if 'libamdhip64' in l}))")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    import ast
    libs = ast.literal_eval(out.stdout.strip().splitlines()[-1])
    assert len(libs) == 1, libs

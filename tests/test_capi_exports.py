"""The C-ABI library loads on a machine without a GPU and exports every
entry point that include/dqmc_amd.h declares (no compute calls here)."""
import ctypes as C
import os
import re

import pytest

import dqmc_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "dqmc_amd.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|void)\s+(dqmc_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert len(syms) >= 20
    assert set(syms) == set(dqmc_amd.EXPORTS)


def test_library_exports_all_declared_symbols():
    lib = C.CDLL(dqmc_amd.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_footprint_without_gpu():
    # pure host logic: validation and arena sizing
    p = dqmc_amd.make_params(12, 60, 6.0, pace=10, checkerboard=True, delay=16)
    nb = dqmc_amd.footprint([p])
    n2 = 144 * 144 * 8
    assert nb > 2 * 6 * 2 * n2  # two spins x six stack levels x (U and T)
    assert dqmc_amd.footprint([p, p]) > nb
    with pytest.raises(dqmc_amd.DqmcError):
        dqmc_amd.footprint([dqmc_amd.make_params(5, 10, 1.0, checkerboard=True)])  # odd side
    with pytest.raises(dqmc_amd.DqmcError):
        dqmc_amd.footprint([dqmc_amd.make_params(4, 10, 1.0, U=float("nan"))])
    with pytest.raises(dqmc_amd.DqmcError):
        dqmc_amd.footprint([dqmc_amd.make_params(4, 10, 1.0, U=-1.0)])       # repulsive needs U >= 0
    with pytest.raises(dqmc_amd.DqmcError):
        dqmc_amd.footprint([dqmc_amd.make_params(4, 10, 1.0, delay=33)])   # blocks take at most 32 sites

"""The C-ABI library loads on a machine without a GPU and exports every
symbol include/mvmc_hip.h declares; calls fail loudly instead of falling back."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mvmc_hip.h")).read()
    return sorted(set(re.findall(r"\b(mvmc_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    import mvmc_hip
    assert sorted(mvmc_hip.EXPORTED_SYMBOLS) == declared_symbols()


def test_library_exports_declared_symbols():
    import mvmc_hip
    if not os.path.exists(mvmc_hip.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(mvmc_hip.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert lib.mvmc_abi_version() == 1


def test_no_gpu_means_loud_failure():
    import mvmc_hip
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(mvmc_hip.BackendUnavailable):
        mvmc_hip.Ensemble(os.path.join(ROOT, "tests", "golden", "inputs", "hubbard_chain", "namelist.def"))

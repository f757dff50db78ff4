"""The C-ABI library loads and exports every entry point that include/pylians_hip.h declares
(no compute calls: this runs without a GPU)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "pylians_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pyl_[a-zA-Z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    assert "pyl_mas_deposit" in names and "pyl_pk" in names and "pyl_bk" in names


def test_library_exports_all_declared_symbols():
    from pylians_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libpylians_hip.so is not built: run `python -m pylians_amd.build`")
    L = _lib.lib()
    for name in declared():
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, name
    assert L.pyl_abi_version() == 1


def test_no_gpu_fails_loudly():
    import torch
    from pylians_amd import MAS_library
    import numpy as np
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(RuntimeError):
        MAS_library.MA(np.zeros((4, 3), np.float32), np.zeros((8, 8, 8), np.float32), 1.0, "CIC")

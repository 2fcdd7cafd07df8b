"""CPU-side checks of the product boundary: the C-ABI library loads, exports
every symbol include/plum_gpu.h declares, and fails loudly (no silent CPU
fallback) when no GPU is present."""
import ctypes as C
import os
import re

import pytest

from plum_amd import engine as E
from plum_amd import systems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "plum_gpu.h")).read()
    return sorted(set(re.findall(r"\b(plum_[a-z_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(E.ABI_SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(E.LIB_PATH)
    for s in declared_symbols():
        assert hasattr(lib, s), s


def test_struct_sizes_match_header():
    # plum_params: 3d+i+d+4i+d+i+32d+3d+i+2d+3d+32d+16i+3i, with padding
    assert C.sizeof(E.Terms) == 7 * 8
    assert C.sizeof(E.MCParams) == 5 * 8 + 8 + 4 + 4 + 8 + 4 + 4 + 8 + 4 + 4


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-GPU error path")
def test_create_fails_loudly_without_gpu():
    sd = systems.config("tiny")
    with pytest.raises(E.PlumError) as ei:
        E.Engine(sd)
    assert "HIP" in str(ei.value) or "device" in str(ei.value)


def test_invalid_input_rejected_before_device_use():
    sd = systems.config("tiny", npbc=4)
    with pytest.raises(E.PlumError) as ei:
        E.Engine(sd)
    assert "npbc" in str(ei.value)
    sd = systems.config("tiny")
    sd["mol_start"] = sd["mol_start"].copy()
    sd["mol_start"][-1] += 1
    with pytest.raises(E.PlumError):
        E.Engine(sd)

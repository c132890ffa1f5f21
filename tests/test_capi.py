"""The drop-in boundary: the C-ABI library loads without a GPU, exports every
entry point include/pic_hip.h declares, and refuses cleanly where there is no
device (the reference's CUDA backend raises RuntimeError the same way,
solver_cuda.py:152-172). No kernel runs here."""

import ctypes
import re
from pathlib import Path

import pytest

from pulsed_ion_chamber_hip import _lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "pic_hip.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*int\s+(pic_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_expected_entry_points():
    assert declared_symbols() == sorted(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = _lib.open_library()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    raw = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in declared_symbols():
        assert ctypes.cast(getattr(raw, name), ctypes.c_void_p).value


def test_params_struct_matches_the_header_layout():
    # 4 int64 + 14 double + 2 int32
    assert ctypes.sizeof(_lib.PicParams) == 4 * 8 + 14 * 8 + 2 * 4


def test_last_error_is_readable_without_a_device():
    assert isinstance(_lib.last_error(), str)


def test_invalid_params_are_rejected_before_touching_a_device():
    lib = _lib.open_library()
    p = _lib.PicParams(no_xy=2, nzb=2, nze=0, no_z=1, cutoff_voxels=1.0)
    size = ctypes.c_size_t()
    rc = lib.pic_mem_required(ctypes.byref(p), 0, ctypes.byref(size))
    assert rc == _lib.PIC_E_INVALID
    assert "too small" in _lib.last_error()


def test_mem_required_counts_four_carrier_arrays():
    lib = _lib.open_library()
    p = _lib.PicParams(no_xy=100, nzb=50, nze=5, no_z=40, cutoff_voxels=5.0)
    size64 = ctypes.c_size_t()
    size32 = ctypes.c_size_t()
    assert lib.pic_mem_required(ctypes.byref(p), _lib.PIC_FP64, ctypes.byref(size64)) == 0
    assert lib.pic_mem_required(ctypes.byref(p), _lib.PIC_FP32, ctypes.byref(size32)) == 0
    # 50 values per row: an even length, so FP32 rows need no padding
    assert size64.value - size32.value == 4 * 100 * 100 * 50 * 4


def test_no_device_raises_runtime_error_not_a_fallback():
    try:
        import torch

        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    from configs import fast_config
    from pulsed_ion_chamber_hip import run_simulation_hip

    with pytest.raises(RuntimeError, match="GPU"):
        run_simulation_hip(fast_config(), progress=False)

"""ctypes binding of ``libpic_hip.so`` (C ABI in include/pic_hip.h).

The library is built in-tree by :mod:`pulsed_ion_chamber_hip._build`. There is
no fallback: if the library or a GPU is missing, :func:`load` raises, so no
caller can silently run anything but the HIP kernels.
"""

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libpic_hip.so"

PIC_OK = 0
PIC_E_INVALID = -1
PIC_E_NOMEM = -2
PIC_E_HIP = -3
PIC_E_NODEVICE = -4

PIC_FP64 = 0
PIC_FP32 = 1
PRECISIONS = {"fp64": PIC_FP64, "fp32": PIC_FP32}

PIC_WALL_ABSORBING = 0
PIC_WALL_REFLECTING = 1

PIC_GATHER_AUTO = 0
PIC_GATHER_SCAN = 1
PIC_GATHER_BINNED = 2
GATHER_MODES = {"auto": PIC_GATHER_AUTO, "scan": PIC_GATHER_SCAN, "binned": PIC_GATHER_BINNED}

# every symbol include/pic_hip.h declares
EXPORTED = (
    "pic_last_error",
    "pic_device_info",
    "pic_mem_required",
    "pic_create",
    "pic_destroy",
    "pic_upload_tracks",
    "pic_run_steps",
    "pic_read_per_step",
    "pic_run_pulse",
    "pic_run_pulse_static",
    "pic_deposit",
    "pic_step",
    "pic_read_fields",
    "pic_write_fields",
    "pic_reset",
    "pic_synchronize",
    "pic_set_gather_mode",
    "pic_get_gather_mode",
    "pic_set_tiling",
    "pic_set_tile_group",
    "pic_set_steps_per_pass",
    "pic_get_tiling",
    "pic_enable_timing",
    "pic_get_timing",
    "pic_set_graphs",
    "pic_batch_create",
    "pic_batch_destroy",
    "pic_batch_set_tiling",
    "pic_batch_get_tiling",
    "pic_batch_run",
    "pic_batch_synchronize",
    "pic_batch_enable_timing",
    "pic_batch_get_timing",
    "pic_batch_set_graphs",
    "pic_adjoint_create",
    "pic_adjoint_create_levels",
    "pic_adjoint_destroy",
    "pic_adjoint_run",
    "pic_adjoint_get_info",
    "pic_adjoint_get_levels",
)


class PicParams(ctypes.Structure):
    _fields_ = [
        ("no_xy", ctypes.c_int64),
        ("nzb", ctypes.c_int64),
        ("nze", ctypes.c_int64),
        ("no_z", ctypes.c_int64),
        ("mid_xy", ctypes.c_double),
        ("scoring_radius_sq", ctypes.c_double),
        ("p_lat", ctypes.c_double),
        ("p_zm", ctypes.c_double),
        ("p_zp", ctypes.c_double),
        ("p_cen", ctypes.c_double),
        ("n_lat", ctypes.c_double),
        ("n_zm", ctypes.c_double),
        ("n_zp", ctypes.c_double),
        ("n_cen", ctypes.c_double),
        ("alpha_dt", ctypes.c_double),
        ("gaussian_factor", ctypes.c_double),
        ("h2_over_b2", ctypes.c_double),
        ("cutoff_voxels", ctypes.c_double),
        ("wall", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int64)
_VP = ctypes.c_void_p
_LIB = None


class HipError(RuntimeError):
    """A failed call into the HIP library; ``code`` is the PIC_E_* status."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


def open_library():
    """dlopen the library and declare signatures; no GPU needed."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH.name} is not built. Build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` from the repository root "
            "(hipcc --offload-arch=gfx950)."
        )
    L = ctypes.CDLL(str(LIB_PATH))
    sig = {
        "pic_last_error": [ctypes.c_char_p, ctypes.c_size_t],
        "pic_device_info": [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t),
                            ctypes.POINTER(ctypes.c_size_t)],
        "pic_mem_required": [ctypes.POINTER(PicParams), ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)],
        "pic_create": [ctypes.POINTER(PicParams), ctypes.c_int, ctypes.POINTER(_VP)],
        "pic_destroy": [_VP],
        "pic_upload_tracks": [_VP, _DP, _DP, ctypes.c_int64, _IP, ctypes.c_int64],
        "pic_run_steps": [_VP, ctypes.c_int64, ctypes.c_int64],
        "pic_read_per_step": [_VP, ctypes.c_int64, ctypes.c_int64, _DP],
        "pic_run_pulse": [_VP, _DP, _DP, ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, _DP],
        "pic_run_pulse_static": [_VP, _DP, _DP, ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, _DP],
        "pic_deposit": [_VP, _DP, _DP, ctypes.c_int64, _DP],
        "pic_step": [_VP, _DP],
        "pic_read_fields": [_VP, _DP, _DP],
        "pic_write_fields": [_VP, _DP, _DP],
        "pic_reset": [_VP],
        "pic_synchronize": [_VP],
        "pic_set_gather_mode": [_VP, ctypes.c_int],
        "pic_get_gather_mode": [_VP, ctypes.POINTER(ctypes.c_int)],
        "pic_set_tiling": [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int],
        "pic_set_tile_group": [_VP, ctypes.c_int],
        "pic_set_steps_per_pass": [_VP, ctypes.c_int, ctypes.c_int],
        "pic_get_tiling": [_VP] + [ctypes.POINTER(ctypes.c_int)] * 3,
        "pic_enable_timing": [_VP, ctypes.c_int],
        "pic_get_timing": [_VP, _DP, _IP, _DP, _IP, _DP],
        "pic_set_graphs": [_VP, ctypes.c_int],
        "pic_batch_create": [ctypes.POINTER(_VP), ctypes.c_int, ctypes.POINTER(_VP)],
        "pic_batch_destroy": [_VP],
        "pic_batch_set_tiling": [_VP, ctypes.c_int, ctypes.c_int],
        "pic_batch_get_tiling": [_VP] + [ctypes.POINTER(ctypes.c_int)] * 4,
        "pic_batch_run": [_VP, ctypes.c_int64, ctypes.c_int64],
        "pic_batch_synchronize": [_VP],
        "pic_batch_enable_timing": [_VP, ctypes.c_int],
        "pic_batch_get_timing": [_VP, _DP, _IP, _DP, _IP],
        "pic_batch_set_graphs": [_VP, ctypes.c_int],
        "pic_adjoint_create": [_VP, ctypes.c_int64, ctypes.POINTER(_VP)],
        "pic_adjoint_destroy": [_VP],
        "pic_adjoint_run": [_VP, _DP, _DP, _DP],
        "pic_adjoint_get_info": [_VP, _IP, ctypes.POINTER(ctypes.c_size_t), _DP, _DP],
        "pic_adjoint_create_levels": [_VP, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(_VP)],
        "pic_adjoint_get_levels": [_VP, _IP, _IP, _IP],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _LIB = L
    return L


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    open_library().pic_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int, what: str) -> None:
    if rc == PIC_OK:
        return
    msg = f"{what}: {last_error()}"
    if rc == PIC_E_NOMEM:
        raise MemoryError(msg)
    if rc == PIC_E_INVALID:
        raise ValueError(msg)
    raise HipError(rc, msg)


def device_info():
    """``(name, free_bytes, total_bytes)`` of the current GPU; raises
    RuntimeError when no GPU is visible."""
    L = open_library()
    name = ctypes.create_string_buffer(256)
    free = ctypes.c_size_t()
    total = ctypes.c_size_t()
    rc = L.pic_device_info(name, len(name), ctypes.byref(free), ctypes.byref(total))
    if rc == PIC_E_NODEVICE:
        raise RuntimeError(
            "run_simulation_hip needs an AMD GPU visible to HIP (none found). The CPU "
            "backends of the reference package need none."
        )
    check(rc, "pic_device_info")
    return name.value.decode(), free.value, total.value


_LIB_DIGEST = None


def library_digest() -> str:
    """First 16 hex digits of the SHA-256 of the built library: identifies the
    kernel set (a rebuilt library may accept other tilings, or run them at
    other speeds)."""
    global _LIB_DIGEST
    if _LIB_DIGEST is None:
        import hashlib

        _LIB_DIGEST = hashlib.sha256(LIB_PATH.read_bytes()).hexdigest()[:16]
    return _LIB_DIGEST


def dp(a: np.ndarray):
    return a.ctypes.data_as(_DP)


def ip(a: np.ndarray):
    return a.ctypes.data_as(_IP)

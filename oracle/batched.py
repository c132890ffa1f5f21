"""ctypes front end of the C/OpenMP oracle (``pic_oracle.c``). Test
infrastructure only: see the package docstring.

Restates pulsed_ion_chamber/solver_numba_parallel.py:526-723 end to end:
:func:`run_batched` takes a config and returns the same :class:`Result` the
reference would.
"""

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"


class OrcParams(ctypes.Structure):
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
        ("reflecting", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


_LIB = None
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int64)


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def build_native(dest: Path) -> Path:
    """Compile pic_oracle.c for the CPU this runs on (-march=native, the
    Makefile's other flags) into ``dest``; used only by the benchmark's CPU
    baseline, on the machine that times it. Returns the library's path."""
    dest = Path(dest)
    dest.mkdir(parents=True, exist_ok=True)
    out = dest / "liboracle_native.so"
    subprocess.run(["gcc", "-O3", "-march=native", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
                    "-std=c11", "-shared", "-o", str(out), str(HERE / "pic_oracle.c"), "-lm"], check=True)
    return out


def use_library(path) -> None:
    """Load the oracle from ``path`` instead of the in-tree build (before the
    first call)."""
    global LIB_PATH, _LIB
    LIB_PATH = Path(path)
    _LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        L.orc_lw_step.argtypes = [ctypes.POINTER(OrcParams), _DP, _DP, _DP, _DP, _DP]
        L.orc_lw_step.restype = None
        L.orc_deposit.argtypes = [ctypes.POINTER(OrcParams), _DP, _DP, _DP, _DP, ctypes.c_int64, _DP]
        L.orc_deposit.restype = ctypes.c_double
        L.orc_run.argtypes = [ctypes.POINTER(OrcParams), _DP, _DP, _IP, ctypes.c_int64, _DP, _DP, _DP, ctypes.c_int]
        L.orc_run.restype = ctypes.c_int
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_last_loop_seconds.restype = ctypes.c_double
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(_DP)


def params(config, alpha_cm3_s=1.60e-6) -> OrcParams:
    (p_lat, p_zm, p_zp, p_cen), (n_lat, n_zm, n_zp, n_cen) = config.scheme_coefficients()
    h2 = config.unit_length_cm**2
    b2 = config.track_radius_cm**2
    return OrcParams(
        no_xy=config.no_xy,
        nzb=config.no_z_with_buffer,
        nze=config.no_z_electrode,
        no_z=config.no_z,
        mid_xy=float(config.mid_xy),
        scoring_radius_sq=float(config.scoring_radius_sq),
        p_lat=p_lat, p_zm=p_zm, p_zp=p_zp, p_cen=p_cen,
        n_lat=n_lat, n_zm=n_zm, n_zp=n_zp, n_cen=n_cen,
        alpha_dt=alpha_cm3_s * config.dt,
        gaussian_factor=float(config.Gaussian_factor),
        h2_over_b2=h2 / b2,
        cutoff_voxels=float(config.track_cutoff_voxels),
        reflecting=1 if config.lateral_boundary == "reflecting" else 0,
        pad=0,
    )


def lw_step(config, pos, neg):
    """One sweep of ``pos``/``neg`` (read-only) into fresh next arrays;
    returns ``(pos_next, neg_next, (recombined, total+, total-))``."""
    q = params(config)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    neg = np.ascontiguousarray(neg, dtype=np.float64)
    pn = np.zeros_like(pos)
    nn = np.zeros_like(neg)
    out = np.zeros(3)
    lib().orc_lw_step(ctypes.byref(q), _dp(pos), _dp(neg), _dp(pn), _dp(nn), _dp(out))
    return pn, nn, tuple(out)


def deposit(config, pos, neg, xs, ys):
    """Batched deposition of tracks ``xs, ys`` into ``pos``/``neg`` in place;
    returns the scored injected charge."""
    q = params(config)
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    total = np.zeros(config.no_xy * config.no_xy)
    return lib().orc_deposit(ctypes.byref(q), _dp(pos), _dp(neg), _dp(xs), _dp(ys), len(xs), _dp(total))


def run_per_step(config, schedule, xs, ys, n_steps=None, nthreads=0, fields=True):
    """Run the loop for ``n_steps`` (default all); returns ``(per_step, pos, neg)``
    with ``per_step`` shaped ``(total_time_steps, 4)``."""
    q = params(config)
    n_total = config.total_time_steps
    n_steps = n_total if n_steps is None else min(int(n_steps), n_total)
    offsets = np.zeros(n_total + 1, dtype=np.int64)
    np.cumsum(schedule, out=offsets[1:])
    per_step = np.zeros((n_total, 4))
    shape = (config.no_xy, config.no_xy, config.no_z_with_buffer)
    pos = np.empty(shape) if fields else None
    neg = np.empty(shape) if fields else None
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    ys = np.ascontiguousarray(ys, dtype=np.float64)
    rc = lib().orc_run(
        ctypes.byref(q), _dp(xs), _dp(ys), offsets.ctypes.data_as(_IP), n_steps, _dp(per_step),
        _dp(pos) if fields else None, _dp(neg) if fields else None, int(nthreads),
    )
    if rc != 0:
        raise MemoryError("oracle could not allocate its four carrier arrays")
    return per_step, pos, neg


def run_batched(config, rng=None, n_steps=None, nthreads=0, fields=True):
    """Full run through the C oracle, returning a Result like the reference's
    ``run_simulation_numba_parallel``."""
    from pulsed_ion_chamber_hip.pulses import draw_tracks
    from pulsed_ion_chamber_hip.state import build_result, track_density_xy

    schedule, xs, ys = draw_tracks(config, rng)
    n_total = config.total_time_steps
    done = n_total if n_steps is None else min(int(n_steps), n_total)
    per_step, pos, neg = run_per_step(config, schedule, xs, ys, done, nthreads, fields)
    if not fields:
        pos = neg = np.zeros((0, 0, 0))
    used = int(schedule[:done].sum())
    return build_result(config, per_step, done, pos, neg, track_density_xy(config, xs[:used], ys[:used]),
                        truncated=done < n_total)


def max_threads() -> int:
    return int(lib().orc_max_threads())


def last_loop_seconds() -> float:
    """Wall time of the last run's step loop (allocation and first touch of
    the arrays excluded)."""
    return float(lib().orc_last_loop_seconds())

"""NumPy restatement of the reference's serial backend. Test infrastructure
only: see the package docstring.

Follows pulsed_ion_chamber/solver_numba.py:
  deposit_track  <- _insert_track_numba       solver_numba.py:54-127
  lw_step        <- _lax_wendroff_step_numba  solver_numba.py:130-213
  run_serial     <- run_simulation_numba      solver_numba.py:237-370
with state.py:113-154 for the wall. Each voxel value is formed with the same
operations in the same order as the reference (the Gaussian factors with
``math.exp`` per element, as Numba's scalar ``exp``), and every scored sum is a
sequential left-to-right sum in the reference's loop order (``np.cumsum``), so
fields and per-step sums match the serial reference bit for bit where the
Numba build does not fuse multiply-adds. Meant for small grids: it is
vectorised per step but deposits track by track.
"""

from math import ceil, exp, floor

import numpy as np


def _seqsum(values: np.ndarray) -> float:
    """Left-to-right FP64 sum, the order of a plain ``for`` loop."""
    if values.size == 0:
        return 0.0
    return float(np.cumsum(values, dtype=np.float64)[-1])


def deposit_track(pos, neg, x, y, config) -> float:
    """One track into both arrays; returns the scored charge (solver_numba.py:54-127)."""
    no_xy, no_z, nze = config.no_xy, config.no_z, config.no_z_electrode
    cut = config.track_cutoff_voxels
    h2 = config.unit_length_cm**2
    b2 = config.track_radius_cm**2
    gf = config.Gaussian_factor
    i_lo = max(int(ceil(x - cut)), 0)
    i_hi = min(int(floor(x + cut)) + 1, no_xy)
    j_lo = max(int(ceil(y - cut)), 0)
    j_hi = min(int(floor(y + cut)) + 1, no_xy)
    gi = np.array([exp(-((i - x) ** 2) * h2 / b2) for i in range(i_lo, i_hi)])
    gj = np.array([exp(-((j - y) ** 2) * h2 / b2) for j in range(j_lo, j_hi)])
    v = (gf * gi)[:, None] * gj[None, :]
    pos[i_lo:i_hi, j_lo:j_hi, nze : nze + no_z] += v[:, :, None]
    neg[i_lo:i_hi, j_lo:j_hi, nze : nze + no_z] += v[:, :, None]
    ii = np.arange(i_lo, i_hi)[:, None]
    jj = np.arange(j_lo, j_hi)[None, :]
    scored = (
        (ii >= 1) & (ii <= no_xy - 2) & (jj >= 1) & (jj <= no_xy - 2)
        & ((ii - config.mid_xy) ** 2 + (jj - config.mid_xy) ** 2 < config.scoring_radius_sq)
    )
    return _seqsum((v * no_z)[scored])


def scored_columns(config) -> np.ndarray:
    """(no_xy-2, no_xy-2) mask of interior columns inside the scored disc."""
    idx = np.arange(1, config.no_xy - 1)
    d2 = (idx[:, None] - config.mid_xy) ** 2 + (idx[None, :] - config.mid_xy) ** 2
    return d2 < config.scoring_radius_sq


def lw_step(pos, neg, pos_next, neg_next, config, alpha_dt):
    """Sweep the interior of ``pos``/``neg`` into ``*_next`` (solver_numba.py:130-213).
    Returns ``(recombined, total+, total-)`` over the scored voxels."""
    (pl, pzm, pzp, pc), (nl, nzm, nzp, nc) = config.scheme_coefficients()
    P, N = pos, neg
    c = (slice(1, -1), slice(1, -1), slice(1, -1))
    p = P[c]
    n = N[c]
    p_new = (
        pzm * P[1:-1, 1:-1, :-2] + pzp * P[1:-1, 1:-1, 2:]
        + pl * P[1:-1, :-2, 1:-1] + pl * P[1:-1, 2:, 1:-1]
        + pl * P[:-2, 1:-1, 1:-1] + pl * P[2:, 1:-1, 1:-1]
        + pc * p
    )
    n_new = (
        nzm * N[1:-1, 1:-1, 2:] + nzp * N[1:-1, 1:-1, :-2]
        + nl * N[1:-1, :-2, 1:-1] + nl * N[1:-1, 2:, 1:-1]
        + nl * N[:-2, 1:-1, 1:-1] + nl * N[2:, 1:-1, 1:-1]
        + nc * n
    )
    r = alpha_dt * p * n
    po = p_new - r
    no = n_new - r
    pos_next[c] = po
    neg_next[c] = no
    nze, no_z = config.no_z_electrode, config.no_z
    k = np.arange(1, config.no_z_with_buffer - 1)
    mask = scored_columns(config)[:, :, None] & ((k >= nze) & (k < nze + no_z))[None, None, :]
    return _seqsum(r[mask]), _seqsum(po[mask]), _seqsum(no[mask])


def apply_lateral_boundary(a, mode):
    """state.py:113-154."""
    if mode != "reflecting":
        return
    a[0] = a[1]
    a[-1] = a[-2]
    a[:, 0] = a[:, 1]
    a[:, -1] = a[:, -2]


def run_serial(config, rng=None, n_steps=None):
    """The whole loop (solver_numba.py:237-370); returns a Result."""
    from pulsed_ion_chamber_hip.pulses import draw_tracks
    from pulsed_ion_chamber_hip.state import build_result, track_density_xy

    schedule, xs, ys = draw_tracks(config, rng)
    offsets = np.concatenate([[0], np.cumsum(schedule)])
    shape = (config.no_xy, config.no_xy, config.no_z_with_buffer)
    pos, neg, pos_next, neg_next = (np.zeros(shape) for _ in range(4))
    alpha_dt = 1.60e-6 * config.dt
    total = config.total_time_steps
    done = total if n_steps is None else min(int(n_steps), total)
    per_step = np.zeros((total, 4))
    for step in range(done):
        injected = 0.0
        for t in range(offsets[step], offsets[step + 1]):
            injected += deposit_track(pos, neg, xs[t], ys[t], config)
        rec, tp, tn = lw_step(pos, neg, pos_next, neg_next, config, alpha_dt)
        per_step[step] = (injected, rec, tp, tn)
        pos[1:-1, 1:-1, 1:-1] = pos_next[1:-1, 1:-1, 1:-1]
        neg[1:-1, 1:-1, 1:-1] = neg_next[1:-1, 1:-1, 1:-1]
        apply_lateral_boundary(pos, config.lateral_boundary)
        apply_lateral_boundary(neg, config.lateral_boundary)
    used = int(offsets[done])
    return build_result(config, per_step, done, pos, neg, track_density_xy(config, xs[:used], ys[:used]),
                        truncated=done < total)

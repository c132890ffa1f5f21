"""CPU reference for the derivative of k_s with respect to the physical
parameters. Test infrastructure only (see the package docstring).

The reference has no derivative of any kind; what it fixes is the forward
map, and that is what this module differentiates. The forward model is the
reference's loop (pulsed_ion_chamber/solver_numba.py:237-370: deposition
:54-127, sweep :130-213, wall state.py:113-176, f_t state.py:51-110) with its
physical inputs made explicit,

    theta = (alpha, mu+, mu-, D+, D-, b)      [cm^3/s, cm^2/(V s), cm^2/s, cm]

and everything else frozen at the configuration's values: dt and the step
count (so a change of mu or D does not re-run the stability loop of
config.py:43-70), the track schedule and centres, the grid, the deposition
box half-width (config.py track_cutoff_voxels, taken at the base radius) and
the line charge N0 = Gaussian_factor * pi * b0^2. The stencil weights follow
config.py:466-482: s = D dt / h^2, c = mu E dt / h, lateral s, z_minus
s + c(c+1)/2, z_plus s + c(c-1)/2, centre 1 - c^2 - 6s; the recombination
product is alpha dt p n (solver_numba.py:273-274); the deposition density of a
track is N0/(pi b^2) exp(-((i-x)^2 + (j-y)^2) h^2/b^2).

With the per-step injected charge I_t and recombined charge R_t summed over
the scored voxels, f = (sum I - sum R)/sum I and k_s = 1/f, so

    dk_s = (sum I * d(sum R) - sum R * d(sum I)) / (sum I - sum R)^2.

d(sum R) comes from one reverse sweep (the adjoint): with S_t the state at
the start of step t, X_t = S_t + D_t the swept input and S_{t+1} = H(X_t),

    mu_t = grad_X r(X_t) + (dH/dX)^T mu_{t+1},     mu_T = 0,

where r(X) = alpha dt sum_scored p n, H is the interior sweep plus the wall
(absorbing: the ring is carried, so its adjoint is carried back; reflecting:
the ring is a copy of its neighbour, so the transposed copy folds the ring
adjoint into the interior first). The transposed stencil is the same stencil
with the z_minus / z_plus roles swapped (the negative species' are already
swapped in the forward, so they swap back). Parameter sensitivities collect
mu_{t+1} . dH/dtheta and mu_t . dD_t/db at every step.

Pinned by FP64 central finite differences of this same model (relative step
1e-6; tests/test_adjoint_oracle.py), and the model's forward k_s by the
reference restatement in oracle/batched.py.
"""

import math
from dataclasses import dataclass

import numpy as np

PARAMS = ("alpha", "mu_positive", "mu_negative", "D_positive", "D_negative", "track_radius")
RECOMBINATION_ALPHA_CM3_S = 1.60e-6  # solver_numba.py:273


@dataclass
class Frozen:
    """Everything the model keeps fixed while theta varies."""

    no_xy: int
    nzb: int
    nze: int
    no_z: int
    dt: float
    h: float
    efield: float
    cutoff: float
    n0: float
    mid: float
    rsq: float
    reflecting: bool
    schedule: np.ndarray
    xs: np.ndarray
    ys: np.ndarray


def base_theta(config) -> np.ndarray:
    """theta at the configuration's own values."""
    return np.array([RECOMBINATION_ALPHA_CM3_S, config.mu_positive, config.mu_negative, config.D_positive,
                     config.D_negative, config.track_radius_cm], dtype=np.float64)


def freeze(config, rng=None) -> Frozen:
    from pulsed_ion_chamber_hip.pulses import draw_tracks

    schedule, xs, ys = draw_tracks(config, rng)
    b0 = config.track_radius_cm
    return Frozen(
        no_xy=config.no_xy, nzb=config.no_z_with_buffer, nze=config.no_z_electrode, no_z=config.no_z,
        dt=config.dt, h=config.unit_length_cm, efield=config.Efield_V_cm, cutoff=config.track_cutoff_voxels,
        n0=config.Gaussian_factor * math.pi * b0 * b0, mid=config.mid_xy, rsq=config.scoring_radius_sq,
        reflecting=config.lateral_boundary == "reflecting", schedule=np.asarray(schedule), xs=np.asarray(xs),
        ys=np.asarray(ys),
    )


def weights(fz: Frozen, diffusion: float, mobility: float):
    """(lateral, z_minus, z_plus, centre) and (s, c) of one species (config.py:466-482)."""
    s = diffusion * fz.dt / fz.h**2
    c = mobility * fz.efield * fz.dt / fz.h
    return (s, s + c * (c + 1.0) / 2.0, s + c * (c - 1.0) / 2.0, 1.0 - c * c - 6.0 * s), (s, c)


def _masks(fz: Frozen):
    i = np.arange(fz.no_xy)[:, None]
    j = np.arange(fz.no_xy)[None, :]
    interior = (i >= 1) & (i <= fz.no_xy - 2) & (j >= 1) & (j <= fz.no_xy - 2)
    cols = interior & ((i - fz.mid) ** 2 + (j - fz.mid) ** 2 < fz.rsq)
    k = np.arange(fz.nzb)
    gap = (k >= fz.nze) & (k < fz.nze + fz.no_z)
    scored = cols[:, :, None] & gap[None, None, :]
    return cols, gap, scored


def deposition(fz: Frozen, b: float, step: int, with_derivative=False):
    """2D density d_t(i, j) of one step's tracks (and dd_t/db), the box of
    each track fixed at the frozen half-width."""
    off0 = int(fz.schedule[:step].sum())
    m = int(fz.schedule[step])
    nx = fz.no_xy
    d = np.zeros((nx, nx))
    dd = np.zeros((nx, nx)) if with_derivative else None
    if m == 0:
        return d, dd
    s = fz.h**2 / b**2
    gf = fz.n0 / (math.pi * b * b)
    idx = np.arange(nx, dtype=np.float64)
    for t in range(off0, off0 + m):
        x, y = fz.xs[t], fz.ys[t]
        li, hi = max(int(math.ceil(x - fz.cutoff)), 0), min(int(math.floor(x + fz.cutoff)) + 1, nx)
        lj, hj = max(int(math.ceil(y - fz.cutoff)), 0), min(int(math.floor(y + fz.cutoff)) + 1, nx)
        di2 = (idx[li:hi] - x) ** 2
        dj2 = (idx[lj:hj] - y) ** 2
        gi = np.exp(-di2 * s)
        gj = np.exp(-dj2 * s)
        v = (gf * gi)[:, None] * gj[None, :]
        d[li:hi, lj:hj] += v
        if with_derivative:
            # d/db [N0/(pi b^2) gi gj] = v (-2/b + 2 h^2/b^3 (di^2 + dj^2))
            dd[li:hi, lj:hj] += v * (-2.0 / b + 2.0 * fz.h**2 / b**3 * (di2[:, None] + dj2[None, :]))
    return d, dd


def _mirror(a):
    a[0] = a[1]
    a[-1] = a[-2]
    a[:, 0] = a[:, 1]
    a[:, -1] = a[:, -2]


def _mirror_transpose(m):
    """Transpose of _mirror: the j copies undone first, then the i copies."""
    m[:, 1] += m[:, 0]
    m[:, 0] = 0.0
    m[:, -2] += m[:, -1]
    m[:, -1] = 0.0
    m[1] += m[0]
    m[0] = 0.0
    m[-2] += m[-1]
    m[-1] = 0.0


def forward(fz: Frozen, theta, store=False):
    """Run the model. Returns (sum I, sum R, k_s, states) where states[t] =
    (X_p, X_n, d_t) when store is set."""
    alpha, mup, mun, dp_, dn_, b = (float(v) for v in theta)
    (pl, pzm, pzp, pc), _ = weights(fz, dp_, mup)
    (nl, nzm, nzp, nc), _ = weights(fz, dn_, mun)
    adt = alpha * fz.dt
    cols, gap, scored = _masks(fz)
    shape = (fz.no_xy, fz.no_xy, fz.nzb)
    P = np.zeros(shape)
    N = np.zeros(shape)
    sum_i = 0.0
    sum_r = 0.0
    states = []
    c = (slice(1, -1),) * 3
    g0, g1 = fz.nze, fz.nze + fz.no_z
    for t in range(len(fz.schedule)):
        d, _ = deposition(fz, b, t)
        P[:, :, g0:g1] += d[:, :, None]
        N[:, :, g0:g1] += d[:, :, None]
        sum_i += float(np.sum(d[cols])) * fz.no_z
        if store:
            states.append((P.copy(), N.copy(), d))
        p = P[c]
        n = N[c]
        r = adt * p * n
        p_new = (pzm * P[1:-1, 1:-1, :-2] + pzp * P[1:-1, 1:-1, 2:] + pl * P[1:-1, :-2, 1:-1]
                 + pl * P[1:-1, 2:, 1:-1] + pl * P[:-2, 1:-1, 1:-1] + pl * P[2:, 1:-1, 1:-1] + pc * p)
        n_new = (nzm * N[1:-1, 1:-1, 2:] + nzp * N[1:-1, 1:-1, :-2] + nl * N[1:-1, :-2, 1:-1]
                 + nl * N[1:-1, 2:, 1:-1] + nl * N[:-2, 1:-1, 1:-1] + nl * N[2:, 1:-1, 1:-1] + nc * n)
        sum_r += float(np.sum(r[scored[c]]))
        P[c] = p_new - r
        N[c] = n_new - r
        if fz.reflecting:
            _mirror(P)
            _mirror(N)
    ks = sum_i / (sum_i - sum_r)
    return sum_i, sum_r, ks, states


def gradient(fz: Frozen, theta):
    """k_s and dk_s/dtheta by one forward pass (states kept) and one reverse
    sweep. Also returns the pieces: sum I, sum R, d(sum R)/dtheta, d(sum I)/db."""
    alpha, mup, mun, dp_, dn_, b = (float(v) for v in theta)
    (pl, pzm, pzp, pc), (sp, cp) = weights(fz, dp_, mup)
    (nl, nzm, nzp, nc), (sn, cn) = weights(fz, dn_, mun)
    adt = alpha * fz.dt
    sum_i, sum_r, ks, states = forward(fz, theta, store=True)
    cols, gap, scored = _masks(fz)
    c = (slice(1, -1),) * 3
    shape = (fz.no_xy, fz.no_xy, fz.nzb)
    mu_p = np.zeros(shape)
    mu_n = np.zeros(shape)
    g = dict(alpha=0.0, pl=0.0, pzm=0.0, pzp=0.0, pc=0.0, nl=0.0, nzm=0.0, nzp=0.0, nc=0.0, b=0.0)
    di_db = 0.0
    g0, g1 = fz.nze, fz.nze + fz.no_z
    sc = scored[c]
    for t in range(len(fz.schedule) - 1, -1, -1):
        P, N, _ = states[t]
        # adjoint of the swept interior at t+1 (reflecting: ring folded in)
        mt_p, mt_n = mu_p.copy(), mu_n.copy()
        if fz.reflecting:
            _mirror_transpose(mt_p)
            _mirror_transpose(mt_n)
        a = mt_p[c]
        e = mt_n[c]
        p = P[c]
        n = N[c]
        pn = p * n
        # parameter sensitivities through H and r
        g["alpha"] += fz.dt * float(np.sum(pn[sc])) - fz.dt * float(np.sum((a + e) * pn))
        g["pzm"] += float(np.sum(a * P[1:-1, 1:-1, :-2]))
        g["pzp"] += float(np.sum(a * P[1:-1, 1:-1, 2:]))
        g["pl"] += float(np.sum(a * (P[1:-1, :-2, 1:-1] + P[1:-1, 2:, 1:-1] + P[:-2, 1:-1, 1:-1] + P[2:, 1:-1, 1:-1])))
        g["pc"] += float(np.sum(a * p))
        g["nzm"] += float(np.sum(e * N[1:-1, 1:-1, 2:]))
        g["nzp"] += float(np.sum(e * N[1:-1, 1:-1, :-2]))
        g["nl"] += float(np.sum(e * (N[1:-1, :-2, 1:-1] + N[1:-1, 2:, 1:-1] + N[:-2, 1:-1, 1:-1] + N[2:, 1:-1, 1:-1])))
        g["nc"] += float(np.sum(e * n))
        # mu_t = grad r + (dH/dX)^T mu_{t+1}
        new_p = np.zeros(shape)
        new_n = np.zeros(shape)
        if not fz.reflecting:  # carried ring (and the untouched k end planes)
            ring = np.ones(shape, dtype=bool)
            ring[c] = False
            new_p[ring] = mu_p[ring]
            new_n[ring] = mu_n[ring]
        new_p[1:-1, 1:-1, :-2] += pzm * a
        new_p[1:-1, 1:-1, 2:] += pzp * a
        new_p[1:-1, :-2, 1:-1] += pl * a
        new_p[1:-1, 2:, 1:-1] += pl * a
        new_p[:-2, 1:-1, 1:-1] += pl * a
        new_p[2:, 1:-1, 1:-1] += pl * a
        new_p[c] += pc * a
        new_n[1:-1, 1:-1, 2:] += nzm * e
        new_n[1:-1, 1:-1, :-2] += nzp * e
        new_n[1:-1, :-2, 1:-1] += nl * e
        new_n[1:-1, 2:, 1:-1] += nl * e
        new_n[:-2, 1:-1, 1:-1] += nl * e
        new_n[2:, 1:-1, 1:-1] += nl * e
        new_n[c] += nc * e
        new_p[c] -= adt * n * (a + e)
        new_n[c] -= adt * p * (a + e)
        new_p[c] += np.where(sc, adt * n, 0.0)
        new_n[c] += np.where(sc, adt * p, 0.0)
        mu_p, mu_n = new_p, new_n
        # deposition of step t, through mu_t
        _, dd = deposition(fz, b, t, with_derivative=True)
        colsum = np.sum(mu_p[:, :, g0:g1] + mu_n[:, :, g0:g1], axis=2)
        g["b"] += float(np.sum(dd * colsum))
        di_db += float(np.sum(dd[cols])) * fz.no_z
    # weights -> (mu, D): s = D dt/h^2, c = mu E dt/h
    ds = fz.dt / fz.h**2
    dc = fz.efield * fz.dt / fz.h
    dr = np.array([
        g["alpha"],
        (g["pzm"] * (cp + 0.5) + g["pzp"] * (cp - 0.5) - 2.0 * cp * g["pc"]) * dc,
        (g["nzm"] * (cn + 0.5) + g["nzp"] * (cn - 0.5) - 2.0 * cn * g["nc"]) * dc,
        (g["pl"] + g["pzm"] + g["pzp"] - 6.0 * g["pc"]) * ds,
        (g["nl"] + g["nzm"] + g["nzp"] - 6.0 * g["nc"]) * ds,
        g["b"],
    ])
    di = np.array([0.0, 0.0, 0.0, 0.0, 0.0, di_db])
    dks = (sum_i * dr - sum_r * di) / (sum_i - sum_r) ** 2
    return ks, dks, dict(sum_i=sum_i, sum_r=sum_r, d_sum_r=dr, d_sum_i=di, weights=g)


def finite_difference(fz: Frozen, theta, rel=1e-6):
    """Central differences of k_s, one parameter at a time."""
    theta = np.asarray(theta, dtype=np.float64)
    out = np.zeros(len(theta))
    for q in range(len(theta)):
        h = rel * theta[q]
        up = theta.copy()
        dn = theta.copy()
        up[q] += h
        dn[q] -= h
        out[q] = (forward(fz, up)[2] - forward(fz, dn)[2]) / (up[q] - dn[q])
    return out

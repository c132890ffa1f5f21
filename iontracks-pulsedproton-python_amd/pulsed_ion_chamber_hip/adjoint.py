"""dk_s/d(alpha, mu+, mu-, D+, D-, track radius) on the GPU by reverse mode.

The reference computes k_s only; this module differentiates the same loop
(pulsed_ion_chamber/solver_numba.py:237-370) with respect to its physical
inputs, for fitting them to measured k_s values. The model freezes what those
inputs would otherwise move discontinuously: dt and the step count (the
stability loop of config.py:43-70 and the clearance time), the schedule and
track centres, the grid and the deposition box half-width (taken at the base
track radius). theta enters through

    stencil weights (config.py:466-482)   s = D dt/h^2, c = mu E dt/h
    recombination (solver_numba.py:273)   alpha dt
    deposition density                    N0/(pi b^2) exp(-r^2 h^2/b^2)

so a run at theta0 is the configuration's own run, and k_s(theta) is smooth.

The library's reverse pass (pic_adjoint_run) returns the derivatives of the
summed recombined and injected charge with respect to the pic_params fields;
the chain rule to theta and k_s = I/(I - R) is applied here. Memory: the
state at the start of every ``segments_per_outer``-th segment of
``checkpoint_every`` steps, the segment starts of one such span (recomputed
by one extra forward pass over the run) and one segment of recomputed
states; ``segments_per_outer`` = 0 takes about the square root of the
segment count on long runs and keeps every segment start on short ones.
"""

import ctypes
import math
from typing import Optional

import numpy as np

from . import _lib
from .constants import RECOMBINATION_ALPHA_CM3_S
from .pulses import draw_tracks
from .solver_hip import HipSimulation, params_from_config
from .state import collection_efficiency

PARAMS = ("alpha", "mu_positive", "mu_negative", "D_positive", "D_negative", "track_radius")


class KsModel:
    """k_s of one configuration as a function of theta (see the module).

    ``tiling`` is the forward sweep's tile shape, as in run_simulation_hip:
    None for the library's rule or ``(rows_per_tile, planes_per_chunk,
    values_per_lane)``; every tiling gives the same forward fields."""

    def __init__(self, config, rng: Optional[np.random.Generator] = None, checkpoint_every: int = 64,
                 segments_per_outer: int = 0, tiling: Optional[tuple] = None):
        self.config = config
        self.tiling = tuple(int(v) for v in tiling) if tiling is not None else None
        self.schedule, self.xs, self.ys = draw_tracks(config, rng)
        self.theta0 = np.array([RECOMBINATION_ALPHA_CM3_S, config.mu_positive, config.mu_negative,
                                config.D_positive, config.D_negative, config.track_radius_cm], dtype=np.float64)
        b0 = config.track_radius_cm
        self.line_charge = config.Gaussian_factor * math.pi * b0 * b0  # N0
        self.checkpoint_every = int(checkpoint_every)
        self.segments_per_outer = int(segments_per_outer)
        self.last_info = {}

    # --- theta -> the C struct ------------------------------------------
    def _species(self, diffusion, mobility):
        c = self.config
        s = diffusion * c.dt / c.unit_length_cm**2
        cc = mobility * c.Efield_V_cm * c.dt / c.unit_length_cm
        return (s, s + cc * (cc + 1.0) / 2.0, s + cc * (cc - 1.0) / 2.0, 1.0 - cc * cc - 6.0 * s), cc

    def params(self, theta) -> "_lib.PicParams":
        alpha, mup, mun, dp, dn, b = (float(v) for v in theta)
        p = params_from_config(self.config)
        (p.p_lat, p.p_zm, p.p_zp, p.p_cen), _ = self._species(dp, mup)
        (p.n_lat, p.n_zm, p.n_zp, p.n_cen), _ = self._species(dn, mun)
        p.alpha_dt = alpha * self.config.dt
        p.gaussian_factor = self.line_charge / (math.pi * b * b)
        p.h2_over_b2 = self.config.unit_length_cm**2 / (b * b)
        return p  # cutoff_voxels stays the configuration's

    # --- forward only ---------------------------------------------------
    def ks(self, theta) -> float:
        """k_s at theta (one forward run, FP64 fields)."""
        T = self.config.total_time_steps
        with HipSimulation(self.config, "fp64", params=self.params(theta)) as sim:
            if self.tiling:
                sim.set_tiling(*self.tiling)
            sim.upload_tracks(self.schedule, self.xs, self.ys)
            sim.run_steps(0, T)
            per = sim.read_per_step(0, T)
        return 1.0 / collection_efficiency(per[:, 0], per[:, 1], T)[-1]

    # --- forward + reverse ----------------------------------------------
    def ks_and_gradient(self, theta=None):
        """(k_s, dk_s/dtheta) at theta (default: the configuration's values)."""
        theta = self.theta0 if theta is None else np.asarray(theta, dtype=np.float64)
        L = _lib.open_library()
        sums = np.zeros(2)
        d_rec = np.zeros(11)
        d_inj = np.zeros(2)
        with HipSimulation(self.config, "fp64", params=self.params(theta)) as sim:
            if self.tiling:
                sim.set_tiling(*self.tiling)
            sim.upload_tracks(self.schedule, self.xs, self.ys)
            adj = ctypes.c_void_p()
            _lib.check(L.pic_adjoint_create_levels(sim.handle, self.checkpoint_every, self.segments_per_outer,
                                                   ctypes.byref(adj)), "pic_adjoint_create_levels")
            try:
                _lib.check(L.pic_adjoint_run(adj, _lib.dp(sums), _lib.dp(d_rec), _lib.dp(d_inj)), "pic_adjoint_run")
                k = ctypes.c_int64()
                nbytes = ctypes.c_size_t()
                fwd = ctypes.c_double()
                bwd = ctypes.c_double()
                _lib.check(L.pic_adjoint_get_info(adj, ctypes.byref(k), ctypes.byref(nbytes), ctypes.byref(fwd),
                                                  ctypes.byref(bwd)), "pic_adjoint_get_info")
                m, outer, nseg = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
                _lib.check(L.pic_adjoint_get_levels(adj, ctypes.byref(m), ctypes.byref(outer), ctypes.byref(nseg)),
                           "pic_adjoint_get_levels")
                self.last_info = dict(checkpoint_every=k.value, device_bytes=nbytes.value, forward_ms=fwd.value,
                                      backward_ms=bwd.value, segments_per_outer=m.value,
                                      outer_checkpoints=outer.value, segments=nseg.value)
            finally:
                L.pic_adjoint_destroy(adj)
        return self._chain(theta, sums, d_rec, d_inj)

    def _chain(self, theta, sums, d_rec, d_inj):
        c = self.config
        alpha, mup, mun, dp, dn, b = (float(v) for v in theta)
        _, cp = self._species(dp, mup)
        _, cn = self._species(dn, mun)
        ds = c.dt / c.unit_length_cm**2
        dc = c.Efield_V_cm * c.dt / c.unit_length_cm
        gf = self.line_charge / (math.pi * b * b)
        h2b2 = c.unit_length_cm**2 / (b * b)
        (p_lat, p_zm, p_zp, p_cen, n_lat, n_zm, n_zp, n_cen, g_adt, g_gf, g_h) = d_rec
        d_r = np.array([
            g_adt * c.dt,
            (p_zm * (cp + 0.5) + p_zp * (cp - 0.5) - 2.0 * cp * p_cen) * dc,
            (n_zm * (cn + 0.5) + n_zp * (cn - 0.5) - 2.0 * cn * n_cen) * dc,
            (p_lat + p_zm + p_zp - 6.0 * p_cen) * ds,
            (n_lat + n_zm + n_zp - 6.0 * n_cen) * ds,
            g_gf * (-2.0 * gf / b) + g_h * (-2.0 * h2b2 / b),
        ])
        d_i = np.zeros(6)
        d_i[5] = d_inj[0] * (-2.0 * gf / b) + d_inj[1] * (-2.0 * h2b2 / b)
        s_i, s_r = sums
        ks = s_i / (s_i - s_r)
        grad = (s_i * d_r - s_r * d_i) / (s_i - s_r) ** 2
        self.last_info.update(sum_injected=s_i, sum_recombined=s_r)
        return ks, grad

    def finite_difference(self, theta=None, rel: float = 1e-6) -> np.ndarray:
        """Central differences of :meth:`ks`, one parameter at a time."""
        theta = self.theta0 if theta is None else np.asarray(theta, dtype=np.float64)
        out = np.zeros(len(theta))
        for q in range(len(theta)):
            up, dn = theta.copy(), theta.copy()
            up[q] += rel * theta[q]
            dn[q] -= rel * theta[q]
            out[q] = (self.ks(up) - self.ks(dn)) / (up[q] - dn[q])
        return out

    def finite_difference_richardson(self, theta=None, rel: float = 1e-3) -> np.ndarray:
        """Richardson-extrapolated central differences: with D(h) the central
        quotient at step h = rel * theta_q, (4 D(h) - D(2h)) / 3 cancels the
        h^2 term of the truncation error, leaving O(h^4). A step of 1e-3
        keeps the rounding noise of k_s (about 1e-14 relative after 1e5
        FP64 steps) near 1e-11 in the quotient while h^4 stays ~1e-12, so
        the result resolves the derivative far below 1e-6 relative even for
        the weakly coupled diffusion coefficients, where a plain quotient at
        1e-6 is limited by that noise to ~1e-4."""
        theta = self.theta0 if theta is None else np.asarray(theta, dtype=np.float64)
        out = np.zeros(len(theta))
        for q in range(len(theta)):
            quot = []
            for mult in (1.0, 2.0):
                h = mult * rel * theta[q]
                up, dn = theta.copy(), theta.copy()
                up[q] += h
                dn[q] -= h
                quot.append((self.ks(up) - self.ks(dn)) / (up[q] - dn[q]))
            out[q] = (4.0 * quot[0] - quot[1]) / 3.0
        return out


def ks_function():
    """``torch.autograd.Function`` mapping a float64 theta tensor of length 6
    to k_s, with the adjoint as its backward (imported lazily: torch is only
    needed for this wrapper)."""
    import torch

    class KsFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, model: KsModel):
            th = theta.detach().cpu().numpy().astype(np.float64)
            if theta.requires_grad:
                ks, grad = model.ks_and_gradient(th)
                ctx.save_for_backward(torch.as_tensor(grad, dtype=theta.dtype))
            else:
                ks = model.ks(th)
            return theta.new_tensor(ks)

        @staticmethod
        def backward(ctx, grad_out):
            (grad,) = ctx.saved_tensors
            return grad_out * grad.to(grad_out.device), None

    return KsFunction


def ks_torch(model: KsModel, theta):
    """k_s(theta) as a differentiable torch scalar."""
    return ks_function().apply(theta, model)

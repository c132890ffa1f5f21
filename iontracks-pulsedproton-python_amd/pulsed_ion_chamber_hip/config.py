"""Run configuration: the physical inputs and every integer and stencil weight
derived from them.

This is the input contract of the time-stepping path. It restates
pulsed_ion_chamber/config.py:43-70 (stable dt), :180-401 (derived grid, timing,
track structure and count) and :466-482 (stencil weights) so that a run built
here has the same grid, the same dt, the same number of steps and the same
number of tracks as one built by the reference. Those quantities come out of
``int(round(...))`` of floating-point quotients and out of a dt found by
repeated division, so the order of every operation matters and is kept.

Any object exposing the same attributes (the reference's own
``SimulationConfig`` included) can be handed to :func:`run_simulation_hip`;
this class exists so the backend does not need the reference installed.
"""

import warnings
from dataclasses import dataclass
from math import exp, pi, sqrt
from typing import Optional

import numpy as np

from .constants import (
    AIR_DENSITY_KG_M3,
    ION_DIFFUSION_CM2_S,
    ION_MOBILITY_CM2_VS,
    W_EV_PER_ION_PAIR,
)
from .stopping_power import E_MeV_u_to_LET_keV_um, calc_track_radius_cm, dose_rate_to_fluence_rate

LATERAL_BOUNDARIES = ("absorbing", "reflecting")
SCORING_REGIONS = ("track_disc", "full_grid")


def von_neumann_dt(ion_diff, grid_spacing_cm, ion_mobility, Efield_V_cm):
    """Largest dt of the sequence 1/1.01^n satisfying ``6 s + c^2 <= 1`` for every
    species (config.py:43-70). A loop, not a closed form: the reference's dt is
    whatever this exact sequence of divisions lands on."""
    diffusion = np.atleast_1d(ion_diff)
    mobility = np.atleast_1d(ion_mobility)
    if len(diffusion) != len(mobility):
        raise ValueError(
            f"ion_diff has {len(diffusion)} entries and ion_mobility {len(mobility)}: "
            "give one diffusion coefficient per mobility, species by species."
        )
    dt = 1.0
    while True:
        dt /= 1.01
        s = diffusion * dt / grid_spacing_cm**2
        c = mobility * Efield_V_cm * dt / grid_spacing_cm
        if np.all(6 * s + c**2 <= 1):
            return dt


def _available_host_bytes() -> Optional[int]:
    """MemAvailable, clamped by a cgroup v2 ``memory.max`` if one is set."""
    available = None
    try:
        with open("/proc/meminfo") as meminfo:
            for line in meminfo:
                if line.startswith("MemAvailable:"):
                    available = int(line.split()[1]) * 1024
                    break
    except (OSError, ValueError, IndexError):
        return None
    try:
        with open("/sys/fs/cgroup/memory.max") as handle:
            raw = handle.read().strip()
        if raw != "max":
            limit = int(raw)
            with open("/sys/fs/cgroup/memory.current") as handle:
                used = int(handle.read().strip())
            headroom = max(0, limit - used)
            available = headroom if available is None else min(available, headroom)
    except (OSError, ValueError):
        pass
    return available


def _format_bytes(n: float) -> str:
    for unit in ("B", "KiB", "MiB", "GiB", "TiB"):
        if abs(n) < 1024.0 or unit == "TiB":
            return f"{n:.2f} {unit}" if unit != "B" else f"{n:.0f} B"
        n /= 1024.0
    return f"{n:.2f} TiB"


@dataclass
class SimulationConfig:
    """Field names, defaults and meaning follow config.py:73-178."""

    E_MeV_u: float = 150.0
    particle: str = "proton"
    voltage_V: float = 200.0
    electrode_gap_cm: float = 0.2

    pulse_duration_s: float = 540e-6
    repetition_rate_hz: float = 50.0
    dose_rate_Gy_s: float = 60.0
    n_pulses: int = 1
    n_clearance_separation_times: float = 2.0
    rf_frequency_hz: Optional[float] = None

    grid_size_um: float = 40.0
    sampled_radius_cm: float = 0.002
    buffer_radius: int = 4
    no_z_electrode: int = 3
    max_voxels: float = 1e8
    memory_budget_fraction: Optional[float] = 0.8

    mu_positive_cm2_Vs: Optional[float] = None
    mu_negative_cm2_Vs: Optional[float] = None
    D_positive_cm2_s: Optional[float] = None
    D_negative_cm2_s: Optional[float] = None
    W_eV: float = W_EV_PER_ION_PAIR
    air_density_kg_m3: float = AIR_DENSITY_KG_M3

    chamber_fill_fraction: float = 1.0
    lateral_boundary: str = "absorbing"
    scoring_region: str = "track_disc"
    track_cutoff_sigmas: Optional[float] = 10.0
    dt_seconds: Optional[float] = None
    seed: Optional[int] = None

    def __post_init__(self):
        self.unit_length_cm = self.grid_size_um * 1e-4
        self.LET_keV_um = E_MeV_u_to_LET_keV_um(self.E_MeV_u, self.particle)
        self.track_radius_cm = calc_track_radius_cm(self.LET_keV_um)
        self.Efield_V_cm = self.voltage_V / self.electrode_gap_cm
        self.area_cm2 = pi * self.sampled_radius_cm**2

        self._resolve_carrier_constants()
        self._validate_options()
        self._derive_grid()
        self._derive_timing()
        self._derive_track_structure()
        self._estimate_and_check_memory()

    # --- validation -------------------------------------------------------
    def _resolve_carrier_constants(self):
        """Per-species constants, defaulting to the averaged pair (config.py:438-464)."""
        pick = lambda override, default: default if override is None else override  # noqa: E731
        self.mu_positive = pick(self.mu_positive_cm2_Vs, ION_MOBILITY_CM2_VS)
        self.mu_negative = pick(self.mu_negative_cm2_Vs, ION_MOBILITY_CM2_VS)
        self.D_positive = pick(self.D_positive_cm2_s, ION_DIFFUSION_CM2_S)
        self.D_negative = pick(self.D_negative_cm2_s, ION_DIFFUSION_CM2_S)
        for name, value in (
            ("mu_positive_cm2_Vs", self.mu_positive),
            ("mu_negative_cm2_Vs", self.mu_negative),
            ("D_positive_cm2_s", self.D_positive),
            ("D_negative_cm2_s", self.D_negative),
        ):
            if value <= 0:
                raise ValueError(f"{name} is {value!r}: transport coefficients are magnitudes and must be > 0.")
        self.two_carrier_species = not (
            self.mu_positive == self.mu_negative and self.D_positive == self.D_negative
        )

    def _validate_options(self):
        if self.lateral_boundary not in LATERAL_BOUNDARIES:
            raise ValueError(
                f"unknown lateral_boundary {self.lateral_boundary!r} (choose from {LATERAL_BOUNDARIES})."
            )
        if self.scoring_region not in SCORING_REGIONS:
            raise ValueError(f"unknown scoring_region {self.scoring_region!r} (choose from {SCORING_REGIONS}).")
        fill = self.chamber_fill_fraction
        if isinstance(fill, bool) or not 0 < fill <= 1:
            raise ValueError(f"chamber_fill_fraction {fill!r} is outside (0, 1]: it is the filled share of the chamber.")

    # --- grid (config.py:200-284) ------------------------------------------
    def _derive_grid(self):
        h = self.unit_length_cm
        self.no_xy = int(round(2 * self.sampled_radius_cm / h)) + 2 * self.buffer_radius
        self.no_z = int(round(self.electrode_gap_cm / h))
        self.no_z_with_buffer = 2 * self.no_z_electrode + self.no_z
        self.effective_gap_cm = self.no_z * h
        mismatch = abs(self.effective_gap_cm - self.electrode_gap_cm)
        self._gap_mismatch_relative = mismatch / self.electrode_gap_cm if self.electrode_gap_cm > 0 else 0.0
        if self._gap_mismatch_relative > 1e-3:
            warnings.warn(
                f"electrode_gap_cm={self.electrode_gap_cm:.6g} cm rounds to no_z={self.no_z} voxels of "
                f"{self.grid_size_um:.4g} um = {self.effective_gap_cm:.6g} cm; the field Efield_V_cm still "
                "uses the requested gap.",
                stacklevel=3,
            )
        if self.no_xy * self.no_xy * self.no_z > self.max_voxels:
            raise ValueError(
                f"{self.no_xy}x{self.no_xy}x{self.no_z} = {self.no_xy * self.no_xy * self.no_z:.3g} voxels "
                f"exceeds max_voxels={self.max_voxels:.3g}; a coarser grid_size_um or a smaller "
                "sampled_radius_cm brings it down."
            )
        # The centre is the continuous midpoint no_xy/2.0, kept as a float.
        self.outer_radius = self.no_xy / 2.0
        self.mid_xy = self.outer_radius
        self.inner_radius = self.outer_radius - self.buffer_radius
        if self.inner_radius <= 0:
            raise ValueError(
                f"no sampled disc is left: inner_radius {self.inner_radius:.3g} voxels once the "
                f"{self.buffer_radius}-voxel buffer is taken from no_xy={self.no_xy}; raise sampled_radius_cm "
                "or lower grid_size_um or buffer_radius."
            )
        self.inner_radius_sq = self.inner_radius**2
        self.sampling_radius = self.inner_radius * self.chamber_fill_fraction
        self.sampling_radius_sq = self.sampling_radius**2
        self.scoring_radius_sq = (
            self.inner_radius_sq if self.scoring_region == "track_disc" else 2.0 * self.no_xy**2
        )

    # --- dt and pulse timing (config.py:286-361) ----------------------------
    def _derive_timing(self):
        self.von_neumann_dt = von_neumann_dt(
            (self.D_positive, self.D_negative),
            self.unit_length_cm,
            (self.mu_positive, self.mu_negative),
            self.Efield_V_cm,
        )
        if self.dt_seconds is None:
            self.dt = self.von_neumann_dt
        elif self.dt_seconds > self.von_neumann_dt:
            raise ValueError(
                f"dt_seconds={self.dt_seconds:.4g} s is above {self.von_neumann_dt:.4g} s, the largest step "
                "the explicit scheme stays stable at on this grid and field."
            )
        else:
            self.dt = float(self.dt_seconds)
        self.courant_number = (
            max(self.mu_positive, self.mu_negative) * self.Efield_V_cm * self.dt / self.unit_length_cm
        )
        self.rf_cycles_per_time_step = self.dt * self.rf_frequency_hz if self.rf_frequency_hz else None
        if self.rf_cycles_per_time_step is not None and self.rf_cycles_per_time_step < 1.0:
            warnings.warn(
                f"dt = {self.dt:.3g} s resolves the RF period (1/{self.rf_frequency_hz:.4g} Hz), so the "
                "time-averaged beam may misrepresent the bucket structure.",
                stacklevel=3,
            )
        self.slowest_mobility_cm2_Vs = min(self.mu_positive, self.mu_negative)
        self.separation_time_steps = int(
            self.electrode_gap_cm / (2.0 * self.slowest_mobility_cm2_Vs * self.Efield_V_cm * self.dt)
        )
        self.clearance_time_steps = int(round(self.n_clearance_separation_times * self.separation_time_steps))
        self.pulse_time_bins = np.arange(0.0, self.pulse_duration_s + self.dt, self.dt)
        self.pulse_time_steps = len(self.pulse_time_bins) - 1
        if self.repetition_rate_hz > 0:
            self.pulse_period_steps = int(round(1.0 / self.repetition_rate_hz / self.dt))
        else:
            self.pulse_period_steps = self.pulse_time_steps
        if self.n_pulses > 1 and self.pulse_period_steps < self.pulse_time_steps:
            raise ValueError(
                f"one pulse takes {self.pulse_time_steps} steps but pulses repeat every "
                f"{self.pulse_period_steps} at this dt: shorten pulse_duration_s or lower repetition_rate_hz."
            )
        self.total_time_steps = (
            (self.n_pulses - 1) * self.pulse_period_steps + self.pulse_time_steps + self.clearance_time_steps
        )

    # --- Gaussian tracks and their number (config.py:363-399) ---------------
    def _derive_track_structure(self):
        self.N0 = self.LET_keV_um * 1e7 / self.W_eV
        self.Gaussian_factor = self.N0 / (pi * self.track_radius_cm**2)
        self.track_sigma_cm = self.track_radius_cm / sqrt(2.0)
        if self.track_cutoff_sigmas is None:
            self.track_cutoff_voxels = float(self.no_xy)
            self.track_cutoff_cm = self.track_cutoff_voxels * self.unit_length_cm
            self.truncated_charge_fraction = 0.0
        else:
            if self.track_cutoff_sigmas <= 0:
                raise ValueError(
                    f"track_cutoff_sigmas {self.track_cutoff_sigmas!r} is not a positive width "
                    "(None keeps every Gaussian tail)."
                )
            self.track_cutoff_cm = self.track_cutoff_sigmas * self.track_sigma_cm
            self.track_cutoff_voxels = self.track_cutoff_cm / self.unit_length_cm
            self.truncated_charge_fraction = exp(-((self.track_cutoff_cm / self.track_radius_cm) ** 2))
        dose_per_pulse_Gy = self.dose_rate_Gy_s / self.repetition_rate_hz
        instantaneous_Gy_s = dose_per_pulse_Gy / self.pulse_duration_s
        fluence_rate = dose_rate_to_fluence_rate(
            instantaneous_Gy_s, self.E_MeV_u, self.particle, air_density_kg_m3=self.air_density_kg_m3
        )
        self.number_of_tracks_per_pulse = max(1, int(round(fluence_rate * self.pulse_duration_s * self.area_cm2)))

    def _estimate_and_check_memory(self):
        """Host-side sizing as in config.py:403-436: four FP64 carrier arrays, the
        arrival-time draw and the 2D scratch; the larger phase is the peak."""
        voxels = self.no_xy**2 * self.no_z_with_buffer
        self.carrier_array_bytes = 4 * voxels * 8
        self.track_schedule_bytes = 2 * self.number_of_tracks_per_pulse * 8
        self.scratch_bytes = self.no_xy**2 * 8
        self.estimated_memory_bytes = max(self.track_schedule_bytes, self.carrier_array_bytes + self.scratch_bytes)
        fraction = self.memory_budget_fraction
        if fraction is None:
            return
        if not 0 < fraction <= 1:
            raise ValueError(f"budget_fraction {fraction!r} is outside (0, 1].")
        available = _available_host_bytes()
        if available is None or self.estimated_memory_bytes <= available * fraction:
            return
        raise MemoryError(
            f"A {self.no_xy}x{self.no_xy}x{self.no_z_with_buffer} grid with "
            f"{self.number_of_tracks_per_pulse:,} tracks/pulse needs about "
            f"{_format_bytes(self.estimated_memory_bytes)}, but only {_format_bytes(available)} is "
            f"available and the budget is {fraction:.0%} of that. Coarsen grid_size_um or reduce "
            "sampled_radius_cm, or set memory_budget_fraction=None."
        )

    def scheme_coefficients(self):
        """Lax-Wendroff weights ``(lateral, z_minus, z_plus, centre)`` per species
        (config.py:466-482): ``s = D dt/h^2``, ``c = mu E dt/h``."""

        def weights(diffusion, mobility):
            s = diffusion * self.dt / self.unit_length_cm**2
            c = mobility * self.Efield_V_cm * self.dt / self.unit_length_cm
            return (s, s + c * (c + 1.0) / 2.0, s + c * (c - 1.0) / 2.0, 1.0 - c * c - 6.0 * s)

        return weights(self.D_positive, self.mu_positive), weights(self.D_negative, self.mu_negative)

    def summary(self) -> str:
        species = (
            f"two species, mu+ = {self.mu_positive:.3g}, mu- = {self.mu_negative:.3g} cm^2/(V s); "
            f"D+ = {self.D_positive:.3g}, D- = {self.D_negative:.3g} cm^2/s"
            if self.two_carrier_species
            else f"single averaged species, mu = {self.mu_positive:.3g} cm^2/(V s), D = {self.D_positive:.3g} cm^2/s"
        )
        return (
            f"Particle   : {self.particle} @ {self.E_MeV_u:.1f} MeV/u (LET = {self.LET_keV_um:.3g} keV/um, "
            f"b = {self.track_radius_cm * 1e4:.3g} um)\n"
            f"Chamber    : gap = {self.electrode_gap_cm} cm, V = {self.voltage_V} V\n"
            f"Carriers   : {species}\n"
            f"Boundary   : {self.lateral_boundary}, scoring over {self.scoring_region}\n"
            f"Grid       : {self.no_xy} x {self.no_xy} x {self.no_z_with_buffer} voxels "
            f"({self.unit_length_cm * 1e4:.3g} um/voxel)\n"
            f"Time step  : {self.dt * 1e9:.1f} ns, {self.total_time_steps} steps\n"
            f"Pulse      : {self.pulse_time_steps} steps, {self.number_of_tracks_per_pulse} tracks, "
            f"{self.n_pulses} pulse(s)"
        )

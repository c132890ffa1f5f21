"""Host-side track arrivals: how many tracks start in each time step, and where.

Both are drawn on the host from ``numpy.random.Generator`` and uploaded, so the
device sees exactly the track centres the reference's CPU and CUDA backends
deposit for the same seed.

* Arrival counts follow pulses.py:39-61: a cumulative sum of uniforms rescaled
  onto ``[0, pulse_duration_s]`` and histogrammed over ``pulse_time_bins``,
  repeated every ``pulse_period_steps``.
* Positions follow pulses.py:64-171: rejection sampling in ``[0, no_xy)^2``
  against the sampling disc, two doubles per attempt taken from one flat stream
  in draw order. How the stream is fetched from the generator (block size) does
  not change a single accepted point, only how far ahead the generator ends.
"""

import numpy as np


def pulse_arrival_counts(config, rng: np.random.Generator) -> np.ndarray:
    """Tracks arriving in each step of one pulse (pulses.py:52-61)."""
    arrivals = np.cumsum(rng.random(config.number_of_tracks_per_pulse))
    arrivals /= arrivals[-1]
    arrivals *= config.pulse_duration_s
    counts, _ = np.histogram(arrivals, config.pulse_time_bins)
    return counts.astype(np.int64)


def build_track_schedule(config, rng: np.random.Generator) -> np.ndarray:
    """int64 ``schedule[total_time_steps]``: new tracks per step (pulses.py:39-49)."""
    schedule = np.zeros(config.total_time_steps, dtype=np.int64)
    counts = pulse_arrival_counts(config, rng)
    for pulse in range(config.n_pulses):
        first = pulse * config.pulse_period_steps
        schedule[first : first + len(counts)] += counts
    return schedule


def sample_xy_inside_cylinder(rng, mid_xy, inner_radius, no_xy):
    """One accepted point, one attempt at a time (pulses.py:64-81). The readable
    definition :class:`CylinderSampler` must reproduce."""
    r_sq = inner_radius * inner_radius
    while True:
        x = rng.uniform(0.0, 1.0) * no_xy
        y = rng.uniform(0.0, 1.0) * no_xy
        if (x - mid_xy) ** 2 + (y - mid_xy) ** 2 <= r_sq:
            return x, y


class CylinderSampler:
    """Block rejection sampler drawing the same points as repeated
    :func:`sample_xy_inside_cylinder` calls (pulses.py:84-171).

    Candidate pairs are evaluated a block at a time; unconsumed candidates stay
    queued for the next call, so uneven request sizes see one continuous
    stream.
    """

    def __init__(self, rng: np.random.Generator, mid_xy, inner_radius: float, no_xy: int):
        self.rng = rng
        self.mid_xy = mid_xy
        self.r_sq = inner_radius * inner_radius
        self.no_xy = no_xy
        acceptance = np.pi * self.r_sq / (no_xy * no_xy)
        self._acceptance = min(1.0, max(0.05, acceptance))
        # queued candidates in draw order, already scaled to grid units
        self._x = np.empty(0)
        self._y = np.empty(0)
        self._ok = np.empty(0, dtype=bool)
        self._next = 0

    def _refill(self, wanted: int) -> None:
        pairs = max(4096, int(wanted / self._acceptance * 1.2) + 1)
        stream = self.rng.random(2 * pairs)
        self._x = stream[0::2] * self.no_xy
        self._y = stream[1::2] * self.no_xy
        self._ok = (self._x - self.mid_xy) ** 2 + (self._y - self.mid_xy) ** 2 <= self.r_sq
        self._next = 0

    def sample(self, n_tracks: int):
        """``(xs, ys)``: the next ``n_tracks`` accepted points."""
        n_tracks = int(n_tracks)
        if n_tracks == 0:
            return np.empty(0), np.empty(0)
        xs, ys = [], []
        still = n_tracks
        while still > 0:
            if self._next >= len(self._ok):
                self._refill(still)
            ok = self._ok[self._next :]
            hits = np.flatnonzero(ok)
            if len(hits) >= still:
                end = self._next + int(hits[still - 1]) + 1
                take = hits[:still] + self._next
            else:
                end = len(self._ok)
                take = hits + self._next
            xs.append(self._x[take])
            ys.append(self._y[take])
            still -= len(take)
            self._next = end
        if len(xs) == 1:
            return xs[0], ys[0]
        return np.concatenate(xs), np.concatenate(ys)


def draw_tracks(config, rng: np.random.Generator = None):
    """Schedule plus every track centre of the run, in arrival order.

    Returns ``(schedule, xs, ys)`` with ``xs[offsets[s]:offsets[s+1]]`` the
    tracks of step ``s`` where ``offsets = concat([0], cumsum(schedule))``.
    Sampling the whole run at once yields the same points as sampling step by
    step, because the sampler consumes one continuous stream.
    """
    rng = rng if rng is not None else np.random.default_rng(config.seed)
    schedule = build_track_schedule(config, rng)
    sampler = CylinderSampler(rng, config.mid_xy, config.sampling_radius, config.no_xy)
    xs, ys = sampler.sample(int(schedule.sum()))
    return schedule, np.ascontiguousarray(xs), np.ascontiguousarray(ys)

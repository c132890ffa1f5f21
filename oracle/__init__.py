"""CPU ORACLE -- test infrastructure only.

Checkers for the HIP backend, never part of it. Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this
package; the shipped library neither imports nor links anything here.

* :mod:`oracle.batched` -- ctypes front end of ``pic_oracle.c``, a C/OpenMP
  restatement of the reference's batched backend
  (pulsed_ion_chamber/solver_numba_parallel.py). Fast enough for the
  full-precision published k_s fixtures and the CPU baseline timing.
* :mod:`oracle.serial` -- a NumPy restatement of the reference's serial
  backend (pulsed_ion_chamber/solver_numba.py), track by track, for small
  grids. It shares no code with the C restatement, so the two check each other.
* :mod:`oracle.theory` -- Jaffe's single-track k_s (pulsed_ion_chamber/theory.py:43-76).

How each part is pinned is listed in tests/test_oracle_pins.py.
"""

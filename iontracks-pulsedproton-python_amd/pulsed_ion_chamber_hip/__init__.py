"""MI355X-native backend for the pulsed ion chamber's time-stepping loop.

``run_simulation_hip(config)`` returns the same ``Result`` as the reference
package's ``run_simulation_numba`` / ``run_simulation_numba_parallel`` /
``run_simulation_cuda`` (pulsed_ion_chamber/__init__.py:26-38), computed by
hand-written gfx950 HIP kernels behind a ctypes C ABI (include/pic_hip.h).
``SimulationConfig`` restates the reference's input contract so the backend
runs without the reference installed; the reference's own config object works
too. ``run_batch_hip(configs)`` runs a whole parameter sweep (uneven grids
and step counts) in one device stream; ``KsModel(config).ks_and_gradient()``
gives dk_s/d(alpha, mu+, mu-, D+, D-, track radius) by reverse mode.
"""

from .adjoint import KsModel, ks_torch
from .batch import HipBatch, run_batch_hip
from .config import SimulationConfig
from .solver_hip import HipSimulation, run_simulation_hip, warmup_hip
from .state import Result

__all__ = ["SimulationConfig", "Result", "run_simulation_hip", "warmup_hip", "HipSimulation", "run_batch_hip",
           "HipBatch", "KsModel", "ks_torch"]

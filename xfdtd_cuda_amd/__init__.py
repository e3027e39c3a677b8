"""xfdtd_cuda_amd -- an MI355X (gfx950) FDTD time-stepping engine behind the
xfdtd::cuda::SimulationHD interface of XFDTD_CUDA.

Python entry points:
  Simulation     scene model + set-up (objects, PML, TFSF, NF2FF, ADE media)
  SimulationHD   the reference's drop-in API: setGridDim, setBlockDim, run(n)
  Engine         direct handle on the engine C ABI (precision, graphs, timing)
  Batch          many 3-D scenes stepped together by the batched kernels
  scenes         the reference drivers' scenes as builders
"""
from ._native import XFError, device_count
from .simulation import (C_0, EPSILON_0, ETA_0, FACES, MU_0, Batch, Engine, Material,
                         Simulation, SimulationHD, Waveform, scene_array)

__all__ = [
    "C_0", "EPSILON_0", "ETA_0", "FACES", "MU_0", "Batch", "Engine", "Material", "Simulation",
    "SimulationHD", "Waveform", "XFError", "device_count", "scene_array",
]

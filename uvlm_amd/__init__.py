"""uvlm_amd -- the UVLM per-timestep hot path on AMD Instinct MI355X (gfx950).

Layers:
  include/uvlm_gpu.h   C-ABI of the HIP kernels (uvlm_amd/csrc, libuvlm_gpu.so)
  include/uvlm_host.h  C entry points of the reference-shaped C++ host API
                       (uvlm_amd/host: namespace UVLM, libuvlm.so)
  uvlm_amd.gpu         Python view of both (numpy arrays)
  uvlm_amd.workloads   the measured configurations
"""
from . import _native  # noqa: F401
from .gpu import Context, Simulation, UVLMError, case_spec  # noqa: F401

__all__ = ["Context", "Simulation", "UVLMError", "case_spec"]

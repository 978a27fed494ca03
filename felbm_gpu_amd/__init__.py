"""felbm_gpu_amd: MI355X (gfx950) engine for the felbm Lee-Liu free-energy
multiphase D3Q19 step, driven through the C ABI of include/felbm_mi355x.h."""
from ._lib import FelbmError, load  # noqa: F401
from .engine import MultiPhaseGPU, Batch, mrt_matrices, forcing_factor_at  # noqa: F401

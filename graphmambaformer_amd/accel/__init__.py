"""Accelerator tier of the path: the HIP kernels behind the reference's tier
function signatures, and the per-stage backend seam."""

from .hip_kernels import banded_sw, chain_dp, kernels_available  # noqa: F401
from .backend import kernel_backend  # noqa: F401

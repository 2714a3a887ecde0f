"""Accelerator tier of the path: the HIP kernels behind the reference's tier
function signatures, and the per-stage backend seam."""

from .hip_kernels import (  # noqa: F401
    banded_sw,
    chain_dp,
    kernels_available,
    kmer_lookup,
    wfa_distance,
)
from .backend import kernel_backend  # noqa: F401

"""Per-stage backend seam (graphmambaformer/accel/backend.py:571-583).

The reference resolves ``AccelConfig.stage_backends[stage]`` in
{"auto", "torch", "cuda_rawkernel"}; this library adds "hip" and resolves
"auto" to it whenever libgmfa.so loads and a HIP device is visible.
"""

from __future__ import annotations

from .hip_kernels import kernels_available


def kernel_backend(stage: str, stage_backends: dict | None = None) -> str:
    override = (stage_backends or {}).get(stage)
    if override and override != "auto":
        return override
    if stage in ("seeding", "chaining", "extension") and kernels_available():
        return "hip"
    return "torch"

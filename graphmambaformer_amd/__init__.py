"""MI355X-native banded affine-gap Smith-Waterman + traceback and anchor chaining.

Drop-in for the classical alignment core of sarakh1999/GraphMambaFormer
(graphmambaformer/alignment/extension.py and chaining.py): hand-written HIP
kernels for gfx950 behind a ctypes C ABI (include/gmfa.h, libgmfa.so), with
host code on PyTorch-ROCm tensors.
"""

__version__ = "0.1.0"

"""MI355X-native alignment core: k-mer seeding, anchor chaining, banded
affine-gap Smith-Waterman with traceback, wavefront alignment, BAM CIGAR
records and continuous serving.

Drop-in for the classical alignment core of sarakh1999/GraphMambaFormer
(graphmambaformer/alignment/seeding.py, chaining.py, extension.py):
hand-written HIP kernels for gfx950 behind a ctypes C ABI (include/gmfa.h,
libgmfa.so), with host code on PyTorch-ROCm tensors.
"""

__version__ = "0.1.0"

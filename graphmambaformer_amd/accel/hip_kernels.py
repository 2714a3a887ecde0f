"""Tier functions with the signatures of graphmambaformer/accel/cuda_kernels.py.

``banded_sw`` (cuda_kernels.py:504-566) and ``chain_dp`` (cuda_kernels.py:
440-501) take and return torch tensors on the current device.  Unlike the
NVRTC tier they have no band-width cap below 2047, keep exact int32 scores and
integer end coordinates, and follow the portable reference semantics (best cell
= first row, first column; chaining float64 with numpy's log2), so an engine
that cross-checks tier against portable results always accepts them.
``RuntimeError`` still means "tier unavailable" to callers.
"""

from __future__ import annotations

import numpy as np
import torch


def kernels_available() -> bool:
    try:
        from .._lib import lib

        lib()
        return torch.cuda.is_available()
    except Exception:
        return False


class _Cfg:
    def __init__(self, **kw):
        self.__dict__.update(kw)


def banded_sw(query, target, query_len, target_len, half_band, match_score, mismatch_penalty,
              gap_open, gap_extend, band_offset=None):
    """(score float32, query_end int64, target_end int64) -- no x-drop."""
    from ..alignment.extension import banded_affine_sw_batch

    cfg = _Cfg(match_score=match_score, mismatch_penalty=mismatch_penalty, gap_open=gap_open,
               gap_extend=gap_extend, x_drop=0.0, half_band=half_band)
    r = banded_affine_sw_batch(query, target, query_len, target_len, cfg, half_band=half_band,
                               band_offset=band_offset, return_matrices=False)
    return r.score, r.query_end, r.target_end


def chain_dp(anchors, n_anchors, lookback, max_gap, gap_open, gap_extend, log_coeff, bonus=None,
             weights=None):
    """anchors (B, A, 3) int32 (read_end, ref_end, reserved) -> (f float32, p int64)."""
    from ..alignment.chaining import chain_dp_batched

    B, A, _ = anchors.shape
    if weights is None:
        weights = anchors[..., 2].to(torch.float64)
    if tuple(weights.shape) != (B, A):
        raise ValueError("weights must have shape (B, A)")
    cfg = _Cfg(max_lookback=lookback, max_gap=max_gap, gap_open=gap_open, gap_extend=gap_extend,
               log_coeff=log_coeff)
    return chain_dp_batched(anchors[..., 0], anchors[..., 1], weights, n_anchors, cfg, bonus=bonus)


__all__ = ["banded_sw", "chain_dp", "kernels_available", "np"]

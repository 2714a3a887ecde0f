"""Tier functions with the signatures of graphmambaformer/accel/cuda_kernels.py.

``banded_sw`` (cuda_kernels.py:504-566), ``chain_dp`` (cuda_kernels.py:
440-501), ``kmer_lookup`` (405-437) and ``wfa_distance`` (569-606) take and
return torch tensors on the current device.  Unlike the
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


def kmer_lookup(table: torch.Tensor, offsets: torch.Tensor, queries: torch.Tensor):
    """(start int32, count int32) of each query in a sorted k-mer table."""
    import ctypes

    from .._lib import KmerIndex, check, lib, ptr, stream_of

    dev = table.device
    keys = table.to(torch.int64).contiguous()
    offs = offsets.to(device=dev, dtype=torch.int64).contiguous()
    q = queries.to(device=dev, dtype=torch.int64).contiguous()
    n = q.numel()
    start = torch.zeros(n, dtype=torch.int64, device=dev)
    count = torch.zeros(n, dtype=torch.int32, device=dev)
    if keys.numel() == 0 or n == 0:
        return start.to(torch.int32), count
    # one bucket (dir_bits 0): a plain binary search over the whole table
    dir_ = torch.tensor([0, keys.numel()], dtype=torch.int32, device=dev)
    ix = KmerIndex()
    ix.keys, ix.offsets, ix.positions, ix.dir = ptr(keys), ptr(offs), ptr(offs), ptr(dir_)
    ix.n_keys, ix.k, ix.dir_bits = int(keys.numel()), 31, 0
    check(lib().gmfa_kmer_lookup(ctypes.byref(ix), ptr(q), n, ptr(start), ptr(count),
                                 stream_of(dev)), "gmfa_kmer_lookup")
    return start.to(torch.int32), count


def wfa_distance(query: torch.Tensor, target: torch.Tensor, query_len: torch.Tensor,
                 target_len: torch.Tensor, max_distance: int) -> torch.Tensor:
    """Batched unit-cost WFA distance, -1 above max_distance."""
    from .. import kernels as K

    B = query.shape[0]
    ql = torch.as_tensor(query_len).detach().cpu().numpy().astype(np.int64).reshape(B)
    tl = torch.as_tensor(target_len).detach().cpu().numpy().astype(np.int64).reshape(B)
    q = query.detach().cpu().numpy()
    t = target.detach().cpu().numpy()
    inp = K.upload_pairs([q[b, :ql[b]].astype(np.int16) for b in range(B)],
                         [t[b, :tl[b]].astype(np.int16) for b in range(B)], np.zeros(B, np.int32),
                         query.device if query.device.type == "cuda" else "cuda", alphabet="ref")
    dist = _wfa_distances(inp, int(max_distance))
    return torch.from_numpy(dist.astype(np.int32)).to(query.device)


def _wfa_distances(inp, max_distance):
    import ctypes

    from .._lib import check, lib, ptr, stream_of

    n = inp.n_pairs
    dev = inp.device
    cap = int(min(max_distance, int(np.maximum(inp.qlen_host, inp.tlen_host).max(initial=0))))
    scratch = torch.empty(max(n * 2 * (2 * cap + 3), 1), dtype=torch.int32, device=dev)
    dist = torch.full((max(n, 1),), -1, dtype=torch.int32, device=dev)
    b = inp.batch_struct()
    check(lib().gmfa_wfa_distance(ctypes.byref(b), int(max_distance), cap, ptr(scratch), ptr(dist),
                                  stream_of(dev)), "gmfa_wfa_distance")
    return dist.cpu().numpy()[:n]


__all__ = ["banded_sw", "chain_dp", "kmer_lookup", "wfa_distance", "kernels_available", "np"]

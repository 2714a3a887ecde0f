"""Stage configs of the path, field-compatible with the reference.

Mirrors ``ExtensionConfig`` and ``ChainingConfig`` of
graphmambaformer/config.py:661-767 (same field names, same defaults) for the
fields the SW and chaining stages read.  The drop-in functions accept either
these objects or the reference's own config instances (duck typed).
"""

from __future__ import annotations

from dataclasses import dataclass


@dataclass
class ExtensionConfig:
    """graphmambaformer/config.py:739-767."""

    algorithm: str = "banded_sw"
    match_score: float = 2.0
    mismatch_penalty: float = 4.0
    gap_open: float = 6.0
    gap_extend: float = 2.0
    x_drop: float = 600.0
    half_band: int = 64
    max_half_band: int = 512
    flank: int = 100
    max_window: int = 32_768
    wfa_max_distance: int = 4_096


@dataclass
class ChainingConfig:
    """graphmambaformer/config.py:661-736 (DP and selection fields; the
    AGNES / graph-bonus fields are carried for compatibility)."""

    max_lookback: int = 64
    max_gap: int = 5_000
    gap_open: float = 6.0
    gap_extend: float = 0.05
    log_coeff: float = 0.5
    graph_bonus: float = 4.0
    graph_max_hops: int = 3
    ref_path_bias: float = 1.5
    adaptive_seed_scoring: bool = True
    confidence_threshold: float = 0.7
    high_confidence_prob: float = 0.7
    low_confidence_prob: float = 0.3
    min_confidence_anchors: int = 5
    max_confidence_anchors: int = 1000
    logit_gate_gain: float = 0.25
    logit_gate_min: float = 0.1
    logit_gate_max: float = 3.0
    gnn_transition_bonus: float = 2.0
    gnn_transition_logit_clip: float = 3.0
    min_chain_score: float = 20.0
    min_chain_anchors: int = 1
    max_chains: int = 8
    secondary_overlap: float = 0.5
    secondary_score_ratio: float = 0.6


def integer_scoring(cfg) -> tuple[int, int, int, int]:
    """(match, mismatch, open, extend) as ints; the int32 DP is exact only for
    integer-valued scores with open >= extend (extension.py:31-33)."""
    vals = (cfg.match_score, cfg.mismatch_penalty, cfg.gap_open, cfg.gap_extend)
    out = []
    for v in vals:
        f = float(v)
        if not f.is_integer() or f < 0:
            raise ValueError(f"SW scoring must be non-negative integers, got {vals}")
        out.append(int(f))
    if out[2] < out[3]:
        raise ValueError("gap_open must be >= gap_extend")
    return tuple(out)

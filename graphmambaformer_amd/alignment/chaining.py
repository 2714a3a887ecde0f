"""Stage 2 drop-in: anchor-chaining DP on the MI355X kernel.

Same call signatures as graphmambaformer/alignment/chaining.py of
sarakh1999/GraphMambaFormer:

  chain_dp_numpy(read_end, ref_end, weight, cfg, bonus=None)
      chaining.py:129-180, bit for bit: float64 scores, the penalty of every
      gap taken from a table evaluated with the reference's own numpy
      expression (chaining.py:120-126), (f + advance) - penalty in IEEE
      double, ties to the smallest predecessor.
  chain_dp_batched(read_end, ref_end, weight, n_anchors, cfg, bonus=None)
      chaining.py:183-258 shape contract ((B, A) padded, f float32, parent
      int64, padded slots 0 / -1); each row equals chain_dp_numpy on it (the
      reference's own contract, tests/test_alignment_stages.py:502-517), with
      integer coordinates so references beyond 2^24 bp stay exact.
  AffineChainer(cfg, backend="hip").chain / chain_batch / select
      chaining.py:292-745: per strand stable (ref_end, read_end) order, anchor
      weights (length, backbone bias, AGNES seed-score gate), device DP (all
      reads and strands of a batch in one launch), host backtrack (libgmfa)
      and primary/secondary selection.

Coordinates must be integer valued; ``max_lookback`` <= 256 (up to 4 windows
of 64 predecessors, one per lane of a wavefront).  A context carrying a pangenome distance oracle or learned
transitions raises NotImplementedError (its bonus can still be passed to
chain_dp_numpy directly).
"""

from __future__ import annotations

from typing import Optional, Sequence

import numpy as np
import torch

from .. import kernels as K
from ..config import ChainingConfig
from .types import AnchorSet, Chain

NO_PREDECESSOR = -1

_LUT_CACHE: dict = {}


def _device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("graphmambaformer_amd needs a HIP device (none visible)")
    return torch.device("cuda", torch.cuda.current_device())


def penalty_table(cfg, device) -> torch.Tensor:
    """pen(g) for g in [0, max_gap): chaining.py:120-126 evaluated by numpy on
    the host (so the table carries exactly numpy's log2), copied once."""
    key = (float(cfg.gap_open), float(cfg.gap_extend), float(cfg.log_coeff), int(cfg.max_gap),
           str(device))
    t = _LUT_CACHE.get(key)
    if t is None:
        g = np.arange(max(int(cfg.max_gap), 1), dtype=np.int64)
        lut = np.where(g > 0, cfg.gap_open + cfg.gap_extend * g + cfg.log_coeff * np.log2(g + 1.0),
                       0.0)
        t = torch.from_numpy(lut.astype(np.float64)).to(device)
        _LUT_CACHE[key] = t
    return t


MAX_LOOKBACK = 256  # up to 4 windows of 64 predecessors (one per lane)


def _lookback(cfg) -> int:
    """The reference takes any max_lookback (chaining.py:154; its default is
    64, config.py:681); the chaining kernels hold up to 256 predecessors (one
    per lane in up to 4 windows of a 64-lane wavefront).  Larger windows are
    rejected, before any device work."""
    L = max(1, int(cfg.max_lookback))
    if L > MAX_LOOKBACK:
        raise ValueError(f"max_lookback {L} above {MAX_LOOKBACK}: the chaining kernels hold up "
                         "to 4 windows of 64 predecessors")
    return L


def _as_int64(a, name) -> np.ndarray:
    x = np.asarray(a)
    if x.dtype.kind == "f":
        if not np.all(np.isfinite(x)) or not np.all(x == np.round(x)):
            raise ValueError(f"{name} must hold integer coordinates")
    return x.astype(np.int64)


def _run_segments(read_end: np.ndarray, ref_end: np.ndarray, weight: np.ndarray,
                  seg_off: np.ndarray, cfg, bonus: Optional[np.ndarray] = None):
    L = _lookback(cfg)
    dev = _device()
    if len(read_end) == 0:
        return np.zeros(0, np.float64), np.zeros(0, np.int64)
    re_ = torch.from_numpy(np.ascontiguousarray(read_end, np.int64)).to(dev)
    fe = torch.from_numpy(np.ascontiguousarray(ref_end, np.int64)).to(dev)
    w = torch.from_numpy(np.ascontiguousarray(weight, np.float64)).to(dev)
    so = torch.from_numpy(np.ascontiguousarray(seg_off, np.int64)).to(dev)
    b = None
    if bonus is not None:
        b = torch.from_numpy(np.ascontiguousarray(bonus, np.float64)).to(dev)
    f, par = K.chain_dp(re_, fe, w, so, L, int(cfg.max_gap), pen_lut=penalty_table(cfg, dev),
                        bonus=b)
    return f.cpu().numpy(), par.cpu().numpy()


def chain_dp_numpy(read_end, ref_end, weight, cfg: ChainingConfig,
                   bonus: Optional[np.ndarray] = None):
    """(f, predecessor) for one read sorted by reference end (chaining.py:129-180)."""
    re_ = _as_int64(read_end, "read_end")
    fe = _as_int64(ref_end, "ref_end")
    w = np.asarray(weight, dtype=np.float64)
    n = len(re_)
    if n == 0:
        return np.zeros(0, np.float64), np.full(0, NO_PREDECESSOR, np.int64)
    if bonus is not None:
        bonus = np.asarray(bonus, np.float64).reshape(n, _lookback(cfg))
    return _run_segments(re_, fe, w, np.array([0, n], np.int64), cfg, bonus)


def _int_coords(x: torch.Tensor, name: str) -> torch.Tensor:
    """Integer-valued coordinates as int64 on their device (ValueError otherwise)."""
    if x.is_floating_point():
        if not bool(torch.all(torch.isfinite(x) & (x == torch.trunc(x)))):
            raise ValueError(f"{name} must hold integer coordinates")
    return x.to(torch.int64)


def chain_dp_batched(read_end: torch.Tensor, ref_end: torch.Tensor, weight: torch.Tensor,
                     n_anchors: torch.Tensor, cfg: ChainingConfig,
                     bonus: Optional[torch.Tensor] = None):
    """Padded (B, A) batch; returns (f float32, parent int64) of shape (B, A).

    The batch stays on the device: the valid slots of every row (the first
    n_anchors[b], row-major) are gathered into back-to-back segments, one DP
    launch runs them all, and the results are scattered back (padded slots 0
    / -1); each row equals chain_dp_numpy on it (chaining.py:183-258)."""
    B, A = read_end.shape
    out_dev = read_end.device
    L = _lookback(cfg)
    dev = _device()
    counts = torch.as_tensor(n_anchors).to(dev).reshape(B).to(torch.int64).clamp(0, A)
    mask = torch.arange(A, device=dev)[None, :] < counts[:, None]
    re_ = _int_coords(read_end.to(dev)[mask], "read_end")
    fe = _int_coords(ref_end.to(dev)[mask], "ref_end")
    w = weight.to(dev, torch.float64)[mask]
    seg = torch.zeros(B + 1, dtype=torch.int64, device=dev)
    torch.cumsum(counts, 0, out=seg[1:])
    fo = torch.zeros((B, A), dtype=torch.float32, device=dev)
    po = torch.full((B, A), NO_PREDECESSOR, dtype=torch.int64, device=dev)
    if re_.numel():
        b = None
        if bonus is not None:
            b = bonus.to(dev, torch.float64)[mask].reshape(-1, L).contiguous()
        f, par = K.chain_dp(re_.contiguous(), fe.contiguous(), w.contiguous(), seg, L,
                            int(cfg.max_gap), pen_lut=penalty_table(cfg, dev), bonus=b)
        # parents are segment-local already (indices within the read)
        fo[mask] = f.to(torch.float32)
        po[mask] = par
    return fo.to(out_dev), po.to(out_dev)


class AffineChainer:
    """Stage 2 entry point (chaining.py:292-745) on the HIP chaining kernel."""

    def __init__(self, cfg: ChainingConfig | None = None, backend: str = "hip",
                 workers: int | None = None):
        self.cfg = cfg or ChainingConfig()
        _lookback(self.cfg)
        self.backend = backend
        self._workers = workers

    # ---- weights ------------------------------------------------------------
    def trust_neural_scores(self, anchors: AnchorSet) -> bool:
        """AGNES confidence decision (chaining.py:307-322)."""
        if not self.cfg.adaptive_seed_scoring:
            return False
        finite = np.isfinite(anchors.score)
        n = int(finite.sum())
        if not (self.cfg.min_confidence_anchors <= n <= self.cfg.max_confidence_anchors):
            return False
        return self._seed_confidence(anchors.score[finite]) > self.cfg.confidence_threshold

    def _seed_confidence(self, scores: np.ndarray) -> float:
        """(mean of confident-high - mean of confident-low) / std (chaining.py:359-381)."""
        if scores.size == 0:
            return 0.0
        s = np.asarray(scores, dtype=np.float64)
        hi = s[s > self.cfg.high_confidence_prob]
        lo = s[s < self.cfg.low_confidence_prob]
        sigma = float(s.std())
        if sigma <= 1e-6:
            return 0.0
        return ((float(hi.mean()) if hi.size else 0.0) - (float(lo.mean()) if lo.size else 0.0)) / sigma

    def _logit_gate(self, score: np.ndarray, finite: np.ndarray) -> np.ndarray:
        """clip(1 + gain * logit(p)) on scored anchors (chaining.py:383-400)."""
        gate = np.ones(len(score), dtype=np.float64)
        p = np.clip(score[finite].astype(np.float64), 1e-4, 1.0 - 1e-4)
        gate[finite] = np.clip(1.0 + self.cfg.logit_gate_gain * np.log(p / (1.0 - p)),
                               self.cfg.logit_gate_min, self.cfg.logit_gate_max)
        return gate

    def _weights(self, anchors: AnchorSet, ctx) -> np.ndarray:
        """Per-anchor weight: length, backbone bias, AGNES gate (chaining.py:324-357)."""
        weight = anchors.length.astype(np.float64)
        backbone = getattr(ctx, "backbone", None) if ctx is not None else None
        if backbone is not None:
            node = anchors.node_id
            on_path = (node >= 0) & backbone[np.clip(node, 0, len(backbone) - 1)]
            weight = weight + on_path * self.cfg.ref_path_bias
        finite = np.isfinite(anchors.score)
        if not finite.any():
            return weight
        trust = getattr(ctx, "trust_neural", None) if ctx is not None else None
        if trust is False:
            return weight
        if self.cfg.adaptive_seed_scoring:
            if trust is True or self.trust_neural_scores(anchors):
                return weight * self._logit_gate(anchors.score, finite)
            return weight
        return weight * (0.5 + np.where(finite, anchors.score, 1.0))

    @staticmethod
    def _check_context(ctx) -> None:
        if ctx is None:
            return
        if getattr(ctx, "oracle", None) is not None or getattr(ctx, "learned_transitions", None):
            raise NotImplementedError("pangenome graph bonus / learned transitions are not part "
                                      "of this library (pass the bonus to chain_dp_numpy)")

    # ---- backtrack / select -------------------------------------------------
    def _backtrack(self, anchors: AnchorSet, f: np.ndarray, parent: np.ndarray, strand: int,
                   global_idx: np.ndarray) -> list:
        """chaining.py:530-576 (host, through libgmfa)."""
        chains = []
        for members, score in K.chain_backtrack(f, parent, self.cfg.min_chain_anchors,
                                                self.cfg.min_chain_score):
            chains.append(Chain(anchor_idx=global_idx[members], score=score, strand=strand,
                                read_start=int(anchors.read_pos[members].min()),
                                read_end=int(anchors.read_end[members].max()),
                                ref_start=int(anchors.ref_pos[members].min()),
                                ref_end=int(anchors.ref_end[members].max())))
        return chains

    def select(self, chains: list) -> list:
        """chaining.py:579-608."""
        if not chains:
            return []
        chains = sorted(chains, key=lambda c: -c.score)
        cutoff = chains[0].score * self.cfg.secondary_score_ratio
        kept: list = []
        for chain in chains:
            if chain.score < cutoff and kept:
                break
            overlap = max((self._read_overlap(chain, o) for o in kept if o.is_primary),
                          default=0.0)
            if overlap > self.cfg.secondary_overlap:
                continue
            chain.is_primary = not kept
            kept.append(chain)
            if len(kept) >= self.cfg.max_chains:
                break
        return kept

    @staticmethod
    def _read_overlap(a: Chain, b: Chain) -> float:
        lo = max(a.read_start, b.read_start)
        hi = min(a.read_end, b.read_end)
        if hi <= lo:
            return 0.0
        return (hi - lo) / max(min(a.read_span, b.read_span), 1)

    # ---- entry points -------------------------------------------------------
    def _segments(self, anchor_sets: Sequence[AnchorSet], contexts):
        """(read, strand) segments in DP order, concatenated."""
        segs = []
        for r, (anchors, ctx) in enumerate(zip(anchor_sets, contexts)):
            if len(anchors) == 0:
                continue
            for strand in (1, -1):
                subset, indices = anchors.for_strand(strand)
                if len(subset) < self.cfg.min_chain_anchors or len(subset) == 0:
                    continue
                ordered, order = subset.sorted_by_ref()
                self._check_context(ctx)
                segs.append((r, strand, ordered, indices[order], self._weights(ordered, ctx)))
        return segs

    def chain_batch(self, anchor_sets: Sequence[AnchorSet], contexts=None, device=None) -> list:
        """Chains per read (chaining.py:644-745): the strand split and weights on
        the host, then sort into DP order, DP, backtrack and select for every
        (read, strand) segment of the batch on the device, one readback."""
        contexts = list(contexts or [None] * len(anchor_sets))
        nr = len(anchor_sets)
        cols = {k: [] for k in ("re", "fe", "rp", "fp", "w", "id")}
        seg_off, seg_strand = [0], []
        read_seg = np.full((nr, 2), -1, np.int32)
        read_base = np.zeros(nr, np.int64)
        for r, (anchors, ctx) in enumerate(zip(anchor_sets, contexts)):
            read_base[r] = seg_off[-1]
            if len(anchors) == 0:
                continue
            self._check_context(ctx)
            for h, strand in enumerate((1, -1)):
                subset, indices = anchors.for_strand(strand)
                if len(subset) == 0 or len(subset) < self.cfg.min_chain_anchors:
                    continue
                read_seg[r, h] = len(seg_strand)
                seg_strand.append(strand)
                cols["re"].append(subset.read_end)
                cols["fe"].append(subset.ref_end)
                cols["rp"].append(subset.read_pos)
                cols["fp"].append(subset.ref_pos)
                cols["w"].append(self._weights(subset, ctx))
                cols["id"].append(indices)
                seg_off.append(seg_off[-1] + len(subset))
        if not seg_strand:
            return [[] for _ in anchor_sets]
        cat = {k: np.concatenate(v) for k, v in cols.items()}
        dev = _device()
        res = K.chain_pipeline(_as_int64(cat["re"], "read_end"), _as_int64(cat["fe"], "ref_end"),
                               _as_int64(cat["rp"], "read_pos"), _as_int64(cat["fp"], "ref_pos"),
                               cat["w"].astype(np.float64), cat["id"].astype(np.int64),
                               np.array(seg_off, np.int64), np.array(seg_strand, np.int32),
                               read_seg, read_base, _lookback(self.cfg), int(self.cfg.max_gap),
                               penalty_table(self.cfg, dev), self.cfg.min_chain_anchors,
                               self.cfg.min_chain_score, self.cfg.secondary_score_ratio,
                               self.cfg.secondary_overlap, self.cfg.max_chains, dev)
        out = []
        for r in range(nr):
            chains = []
            for x in range(int(res.n_kept[r])):
                c = int(res.kept[r, x])
                o, m = int(res.mem[c, 0]), int(res.mem[c, 1])
                sp = res.span[c]
                ch = Chain(anchor_idx=res.members[o:o + m].copy(), score=float(res.score[c]),
                           strand=int(res.strand[c]), read_start=int(sp[0]), read_end=int(sp[1]),
                           ref_start=int(sp[2]), ref_end=int(sp[3]))
                ch.is_primary = x == 0
                chains.append(ch)
            out.append(chains)
        return out

    def chain(self, anchors: AnchorSet, ctx=None, device=None) -> list:
        """chaining.py:620-642."""
        if len(anchors) == 0:
            return []
        return self.chain_batch([anchors], [ctx], device)[0]

"""Stage 3 drop-in: banded affine SW + traceback on the MI355X kernels.

Same call signatures and results as graphmambaformer/alignment/extension.py of
sarakh1999/GraphMambaFormer:

  banded_affine_sw_batch(query, target, query_len, target_len, cfg,
                         half_band=None, band_offset=None, return_matrices=False)
      extension.py:78-204.  Scores, 1-based ends and the batch-wide x-drop are
      identical; instead of float32 H/E/F matrices (12 B/cell) the result keeps
      4-bit traceback planes on the device (0.5 B/cell).
  traceback_banded(result, index, query, target, cfg, tolerance=1e-3)
      extension.py:207-302.  The first call walks every pair of the result on
      the device (one launch) and caches the CIGARs; ``query`` / ``target`` are
      taken to be the sequences the forward pass saw, as at every reference
      call site.
  ExtensionEngine(cfg, device, backend="hip").extend_chains(...)
      extension.py:469-641: per-chain windows on the chain diagonal, one batch
      per read with a shared half-band, soft clips, coordinate lift.

Scores must be integer valued with gap_open >= gap_extend (the int32 DP is
exact there, and the reference's closed-form E assumes it).  Base codes are the
reference alphabet (1..4 = ACGT, 5 = N, negative = padding); other codes are
rejected.  There is no CPU path: without a HIP device and libgmfa.so every
call raises.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

from .. import kernels as K
from ..config import ExtensionConfig, integer_scoring
from .seeding import encode_bases, reverse_complement_codes
from .types import AnchorSet, Chain, ExtensionResult, merge_cigar


def _hip_device(t: torch.Tensor) -> torch.device:
    if t.device.type == "cuda":
        return t.device
    if not torch.cuda.is_available():
        raise RuntimeError("graphmambaformer_amd needs a HIP device (none visible)")
    return torch.device("cuda", torch.cuda.current_device())


@dataclass
class _Walk:
    cigars: list
    read_start: np.ndarray
    ref_start: np.ndarray
    counts: np.ndarray


@dataclass
class TracebackPlanes:
    """Device-resident traceback state of one forward batch."""

    inputs: K.SwInputs
    forward: K.SwForward
    score: torch.Tensor
    query_end: torch.Tensor
    target_end: torch.Tensor
    _walk: Optional[_Walk] = field(default=None, repr=False)

    def walk(self) -> _Walk:
        if self._walk is None:
            tbk = K.sw_traceback(self.inputs, self.forward, self.score, self.query_end,
                                 self.target_end)
            self._walk = _Walk(cigars=K.decode_cigars(tbk),
                               read_start=tbk.read_start.cpu().numpy(),
                               ref_start=tbk.ref_start.cpu().numpy(),
                               counts=tbk.counts.cpu().numpy())
        return self._walk


@dataclass
class BandedDPResult:
    """extension.py:59-76; ``traceback`` replaces the H/E/F matrices."""

    score: torch.Tensor
    query_end: torch.Tensor
    target_end: torch.Tensor
    half_band: int
    band_offset: torch.Tensor
    H: Optional[torch.Tensor] = None
    E: Optional[torch.Tensor] = None
    F: Optional[torch.Tensor] = None
    traceback: Optional[TracebackPlanes] = None


def _inputs_from_codes(query: torch.Tensor, target: torch.Tensor, qlen_h: np.ndarray,
                       tlen_h: np.ndarray, off_h: np.ndarray, dev) -> K.SwInputs:
    q = query.to(device=dev, dtype=torch.int16)
    t = target.to(device=dev, dtype=torch.int16)
    ql = torch.from_numpy(qlen_h.astype(np.int32)).to(dev)
    tl = torch.from_numpy(tlen_h.astype(np.int32)).to(dev)
    qseq, qmask, qoff, badq = K.pack_device(q, ql, qlen_h)
    tseq, tmask, toff, badt = K.pack_device(t, tl, tlen_h)
    if int((badq + badt).item()) != 0:
        raise ValueError("base codes must be negative (padding), 1..4 (ACGT) or 5 (N)")
    return K.SwInputs(qseq=qseq, tseq=tseq, qmask=qmask, tmask=tmask, q_off=qoff, t_off=toff,
                      qlen=ql, tlen=tl,
                      band_offset=torch.from_numpy(off_h.astype(np.int32)).to(dev),
                      qlen_host=qlen_h.astype(np.int32), tlen_host=tlen_h.astype(np.int32),
                      boff_host=off_h.astype(np.int32), has_n=True)


def banded_affine_sw_batch(query: torch.Tensor, target: torch.Tensor, query_len: torch.Tensor,
                           target_len: torch.Tensor, cfg: ExtensionConfig,
                           half_band: Optional[int] = None,
                           band_offset: Optional[torch.Tensor] = None,
                           return_matrices: bool = False) -> BandedDPResult:
    """Batched banded affine-gap local alignment (extension.py:78-204)."""
    ma, mi, go, ge = integer_scoring(cfg)
    B, M = query.shape
    N = target.shape[1]
    hb = int(cfg.half_band if half_band is None else half_band)
    out_dev = query.device
    qlen_h = torch.as_tensor(query_len).detach().cpu().numpy().astype(np.int64).reshape(B)
    tlen_h = torch.as_tensor(target_len).detach().cpu().numpy().astype(np.int64).reshape(B)
    off_h = (np.zeros(B, np.int64) if band_offset is None else
             torch.as_tensor(band_offset).detach().cpu().numpy().astype(np.int64).reshape(B))
    offset_t = torch.as_tensor(off_h, dtype=torch.long, device=out_dev)
    if B == 0:
        z = torch.zeros(0, device=out_dev)
        return BandedDPResult(score=z.float(), query_end=z.long(), target_end=z.long(),
                              half_band=hb, band_offset=offset_t)
    if (qlen_h < 0).any() or (qlen_h > M).any() or (tlen_h < 0).any() or (tlen_h > N).any():
        raise ValueError("query_len / target_len must lie within the padded shapes")
    if np.abs(off_h).max() > (1 << 30):
        raise ValueError("band_offset out of range")
    dev = _hip_device(query)
    inp = _inputs_from_codes(query, target, qlen_h, tlen_h, off_h, dev)
    xdrop = float(cfg.x_drop) > 0
    fwd = K.sw_forward(inp, ma, mi, go, ge, hb, traceback=return_matrices, rowkeys=xdrop)
    if xdrop:
        sc, qe, te = K.xdrop_best(fwd, inp.qlen_host, float(cfg.x_drop), hb, inp.boff_host)
        sc_d = torch.from_numpy(sc).to(dev)
        qe_d = torch.from_numpy(qe).to(dev)
        te_d = torch.from_numpy(te).to(dev)
    else:
        sc_d, qe_d, te_d = fwd.score, fwd.query_end, fwd.target_end
    planes = None
    if return_matrices:
        planes = TracebackPlanes(inputs=inp, forward=fwd, score=sc_d, query_end=qe_d,
                                 target_end=te_d)
    return BandedDPResult(score=sc_d.to(device=out_dev, dtype=torch.float32),
                          query_end=qe_d.to(device=out_dev, dtype=torch.long),
                          target_end=te_d.to(device=out_dev, dtype=torch.long),
                          half_band=hb, band_offset=offset_t, traceback=planes)


def traceback_banded(result: BandedDPResult, index: int, query: np.ndarray, target: np.ndarray,
                     cfg: ExtensionConfig, tolerance: float = 1e-3) -> ExtensionResult:
    """CIGAR of pair ``index`` (extension.py:207-302)."""
    if result.traceback is None:
        raise ValueError("traceback needs banded_affine_sw_batch(return_matrices=True)")
    w = result.traceback.walk()
    cig = w.cigars[index]
    c = w.counts[index]
    return ExtensionResult(score=float(result.score[index]), cigar=list(cig),
                           read_start=int(w.read_start[index]),
                           read_end=int(result.query_end[index]),
                           ref_start=int(w.ref_start[index]),
                           ref_end=int(result.target_end[index]), n_match=int(c[0]),
                           n_mismatch=int(c[1]), n_insertion=int(c[2]), n_deletion=int(c[3]))


class ExtensionEngine:
    """Stage 3 entry point (extension.py:469-641) on the HIP kernels."""

    def __init__(self, cfg: ExtensionConfig | None = None, device=None, backend: str = "hip"):
        self.cfg = cfg or ExtensionConfig()
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.backend = backend

    def _window(self, chain: Chain, anchors: AnchorSet, read_len: int, ref_len: int):
        """(target_start, target_end, half_band) -- extension.py:489-509."""
        diagonal = chain.ref_start - chain.read_start
        start = max(0, diagonal - self.cfg.flank)
        end = min(ref_len, diagonal + read_len + self.cfg.flank)
        if end - start > self.cfg.max_window:
            end = start + self.cfg.max_window
        idx = np.asarray(chain.anchor_idx)
        spread = 0
        if len(idx):
            dg = anchors.diagonal[idx]
            spread = int(dg.max() - dg.min())
        return start, end, min(self.cfg.half_band + spread, self.cfg.max_half_band)

    def _prepare(self, read_seq: str, chains: Sequence[Chain], anchors: AnchorSet,
                 ref_codes: np.ndarray):
        fwd = encode_bases(read_seq)
        rev = reverse_complement_codes(fwd)
        read_len, ref_len = len(fwd), len(ref_codes)
        wins = [self._window(c, anchors, read_len, ref_len) for c in chains]
        hb = max(w[2] for w in wins)
        width = max(w[1] - w[0] for w in wins)
        B = len(chains)
        q = np.full((B, read_len), -1, np.int16)
        t = np.full((B, max(width, 1)), -1, np.int16)
        tlen = np.zeros(B, np.int64)
        off = np.zeros(B, np.int64)
        for b, (c, (start, end, _)) in enumerate(zip(chains, wins)):
            codes = fwd if c.strand > 0 else rev
            q[b, :len(codes)] = codes
            w = ref_codes[start:end]
            t[b, :len(w)] = w
            tlen[b] = len(w)
            off[b] = (c.ref_start - c.read_start) - start
        return q, t, np.full(B, read_len, np.int64), tlen, off, hb, wins, read_len

    def extend_chains(self, read_seq: str, chains: Sequence[Chain], anchors: AnchorSet,
                      ref_seq: str) -> list:
        """One ExtensionResult per chain (extension.py:511-641)."""
        if not chains:
            return []
        if self.cfg.algorithm == "wfa":
            raise NotImplementedError("the wavefront aligner is not part of this library")
        ref_codes = encode_bases(ref_seq)
        q, t, ql, tl, off, hb, wins, read_len = self._prepare(read_seq, chains, anchors,
                                                               ref_codes)
        dev = self.device
        res = banded_affine_sw_batch(torch.from_numpy(q).to(dev), torch.from_numpy(t).to(dev),
                                     torch.from_numpy(ql), torch.from_numpy(tl), self.cfg,
                                     half_band=hb, band_offset=torch.from_numpy(off),
                                     return_matrices=True)
        return self._finish(res, chains, wins, read_len)

    def _finish(self, res: BandedDPResult, chains, wins, read_len):
        out = []
        for b, (chain, (start, _e, _)) in enumerate(zip(chains, wins)):
            ext = traceback_banded(res, b, None, None, self.cfg)
            ext.ref_start += start
            ext.ref_end += start
            ext.cigar = merge_cigar([("S", ext.read_start)] + ext.cigar +
                                    [("S", read_len - ext.read_end)])
            out.append(ext)
        return out

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
  WavefrontAligner(cfg).align(query, target)
      extension.py:308-463: global unit-cost alignment, distance + CIGAR,
      RuntimeError past wfa_max_distance; align_batch for many pairs.
  ExtensionEngine(cfg, device, backend="hip").extend_chains(...)
      extension.py:469-730: per-chain windows on the chain diagonal, one batch
      per read with a shared half-band, soft clips, coordinate lift; the WFA
      branch with its per-chain banded fallback; extend_batch for many reads.

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


MAX_HALF_BAND = 1023  # the forward kernels' tiles hold bands up to 2047 cells wide


def _check_half_band(hb: int) -> int:
    """The reference takes any half-band (extension.py:105; its default
    max_half_band is 512, config.py:760); the kernels here hold up to 1023.
    Wider bands are rejected, before any device work."""
    if not 0 <= int(hb) <= MAX_HALF_BAND:
        raise ValueError(f"half_band {int(hb)} outside [0, {MAX_HALF_BAND}]: the banded SW "
                         f"kernels hold bands up to {2 * MAX_HALF_BAND + 1} cells wide")
    return int(hb)


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
    hb = _check_half_band(cfg.half_band if half_band is None else half_band)
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
        sc_d, qe_d, te_d = K.sw_xdrop(fwd, inp, np.array([0, B]), float(cfg.x_drop), hb)
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


class WavefrontAligner:
    """Global unit-cost alignment (extension.py:308-463) on the device.

    ``align(query, target)`` returns ``(edit_distance, cigar)`` and raises
    RuntimeError past ``cfg.wfa_max_distance`` (the caller's cue to fall back
    to banded DP), as the reference does; ``align_batch`` aligns many pairs in
    two launches and returns None for the pairs over the budget."""

    def __init__(self, cfg: ExtensionConfig | None = None, device=None):
        self.cfg = cfg or ExtensionConfig()
        self.device = torch.device(device) if device is not None else torch.device("cuda")

    def align_batch(self, queries: Sequence[np.ndarray], targets: Sequence[np.ndarray]) -> list:
        if len(queries) != len(targets):
            raise ValueError("one target per query")
        if not queries:
            return []
        inp = K.upload_pairs([np.asarray(q) for q in queries], [np.asarray(t) for t in targets],
                             np.zeros(len(queries), np.int32), self.device, alphabet="ref")
        dist, cigs, _ = K.wfa_align(inp, int(self.cfg.wfa_max_distance))
        return [None if d < 0 else (int(d), list(c)) for d, c in zip(dist, cigs)]

    def align(self, query: np.ndarray, target: np.ndarray):
        res = self.align_batch([query], [target])[0]
        if res is None:
            raise RuntimeError(f"WFA exceeded wfa_max_distance={self.cfg.wfa_max_distance}")
        return res


class ExtensionEngine:
    """Stage 3 entry point (extension.py:469-641) on the HIP kernels."""

    def __init__(self, cfg: ExtensionConfig | None = None, device=None, backend: str = "hip"):
        self.cfg = cfg or ExtensionConfig()
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.backend = backend
        self._ref_cache = None

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

    def _reference(self, ref_seq: str):
        """The reference as a resident 2-bit stream + N mask (cached for the
        last reference seen); windows are then offsets, never copies."""
        key = (len(ref_seq), hash(ref_seq))
        if self._ref_cache is None or self._ref_cache[0] != key:
            codes = encode_bases(ref_seq)
            w, m, _, has_n = K.pack_host([codes], alphabet="ref")
            dev = self.device
            self._ref_cache = (key, torch.from_numpy(w.view(np.int32)).to(dev),
                               torch.from_numpy(m.view(np.int32)).to(dev), has_n, len(codes))
        return self._ref_cache[1:]

    def _extend_wfa_batch(self, reads, chains, anchors, ref_seq):
        """The WFA branch (extension.py:643-730) for every chain of every read:
        global alignment of the read against the window at the chain's
        read-zero diagonal, length read_len + the chain's net indel; chains over
        wfa_max_distance are extended with banded SW alone, as the reference
        falls back one chain at a time."""
        import dataclasses

        cfg = self.cfg
        ref_words, ref_masks, ref_has_n, ref_len = self._reference(ref_seq)
        dev = self.device
        seqs, lens = [], []
        for r in range(len(reads)):
            f = encode_bases(reads[r])
            lens.append(len(f))
            seqs.append(f)
            seqs.append(reverse_complement_codes(f))
        qw, qm, qoff, q_has_n = K.pack_host(seqs, alphabet="ref") if seqs else (
            np.zeros(1, np.uint32), np.zeros(1, np.uint32), np.zeros(0, np.int64), False)
        has_n = bool(q_has_n or ref_has_n)
        owner, q_off, t_off, ql, tl = [], [], [], [], []
        for r in range(len(reads)):
            L = lens[r]
            for k, c in enumerate(chains[r]):
                start = int(np.clip(c.ref_start - c.read_start, 0, max(ref_len - 1, 0)))
                net = (c.ref_end - c.ref_start) - (c.read_end - c.read_start)
                end = min(ref_len, start + max(1, L + net))
                owner.append((r, k, start))
                q_off.append(int(qoff[2 * r + (0 if c.strand > 0 else 1)]))
                t_off.append(start)
                ql.append(L)
                tl.append(max(0, end - start))
        out: list = [[None] * len(chains[r]) for r in range(len(reads))]
        if not owner:
            return out

        def up(a):
            return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

        ql_h, tl_h = np.array(ql, np.int32), np.array(tl, np.int32)
        inp = K.SwInputs(qseq=up(qw.view(np.int32)), tseq=ref_words,
                         qmask=up(qm.view(np.int32)) if has_n else None,
                         tmask=ref_masks if has_n else None, q_off=up(np.array(q_off, np.int64)),
                         t_off=up(np.array(t_off, np.int64)), qlen=up(ql_h), tlen=up(tl_h),
                         band_offset=up(np.zeros(len(ql), np.int32)), qlen_host=ql_h,
                         tlen_host=tl_h, boff_host=np.zeros(len(ql), np.int32), has_n=has_n)
        dist, cigs, cnt = K.wfa_align(inp, int(cfg.wfa_max_distance))
        fallback = []
        for p, (r, k, start) in enumerate(owner):
            if dist[p] < 0:
                fallback.append((r, k))
                continue
            cig = cigs[p]
            gap = sum(cfg.gap_open + max(n - 1, 0) * cfg.gap_extend for op, n in cig
                      if op in ("I", "D"))
            score = (int(cnt[p, 0]) * cfg.match_score - int(cnt[p, 1]) * cfg.mismatch_penalty
                     - gap)
            out[r][k] = ExtensionResult(score=float(score), cigar=list(cig), read_start=0,
                                        read_end=int(ql[p]), ref_start=start,
                                        ref_end=start + int(tl[p]), n_match=int(cnt[p, 0]),
                                        n_mismatch=int(cnt[p, 1]), n_insertion=int(cnt[p, 2]),
                                        n_deletion=int(cnt[p, 3]))
        if fallback:
            sw = ExtensionEngine(dataclasses.replace(cfg, algorithm="banded_sw"), device=dev,
                                 backend=self.backend)
            sw._ref_cache = self._ref_cache
            res = sw.extend_batch([reads[r] for r, _ in fallback],
                                  [[chains[r][k]] for r, k in fallback],
                                  [anchors[r] for r, _ in fallback], ref_seq)
            for (r, k), one in zip(fallback, res):
                out[r][k] = one[0]
        return out

    def extend_chains(self, read_seq: str, chains: Sequence[Chain], anchors: AnchorSet,
                      ref_seq: str) -> list:
        """One ExtensionResult per chain (extension.py:511-641)."""
        if not chains:
            return []
        return self.extend_batch([read_seq], [chains], [anchors], ref_seq)[0]

    def extend_batch(self, reads: Sequence[str], chains: Sequence[Sequence[Chain]],
                     anchors: Sequence[AnchorSet], ref_seq: str) -> list:
        """extend_chains over many reads with one forward launch per distinct
        half-band (not one per read).

        Each read's chains form one reference batch (extension.py:534-585):
        they share the read's half-band (the max over its chains) and the
        batch-wide x-drop, applied per read on the device.  Reads are packed
        once (forward + reverse complement), windows are offsets into the
        resident reference, so no per-chain copy is made; soft clips and the
        coordinate lift (extension.py:628-640) are applied to the results."""
        cfg = self.cfg
        if not (len(reads) == len(chains) == len(anchors)):
            raise ValueError("reads, chains and anchors must have one entry per read")
        if cfg.algorithm == "wfa":
            return self._extend_wfa_batch(reads, chains, anchors, ref_seq)
        ma, mi, go, ge = integer_scoring(cfg)
        out: list = [[] for _ in reads]
        live = [r for r in range(len(reads)) if len(chains[r])]
        if not live:
            return out
        ref_words, ref_masks, ref_has_n, ref_len = self._reference(ref_seq)
        dev = self.device
        # queries: forward and reverse complement of every read, packed once
        seqs, lens = [], np.zeros(len(reads), np.int64)
        for r in range(len(reads)):
            f = encode_bases(reads[r])
            lens[r] = len(f)
            seqs.append(f)
            seqs.append(reverse_complement_codes(f))
        qw, qm, qoff, q_has_n = K.pack_host(seqs, alphabet="ref")
        has_n = bool(q_has_n or ref_has_n)
        qseq = torch.from_numpy(qw.view(np.int32)).to(dev)
        qmask = torch.from_numpy(qm.view(np.int32)).to(dev) if has_n else None
        # windows and per-read half-bands
        by_hb: dict = {}
        geo = {}
        for r in live:
            L = int(lens[r])
            wins = [self._window(c, anchors[r], L, ref_len) for c in chains[r]]
            hb = _check_half_band(max(w[2] for w in wins))
            geo[r] = wins
            by_hb.setdefault(hb, []).append(r)
        xd = float(cfg.x_drop)
        for hb, rs in sorted(by_hb.items()):
            q_off, t_off, ql, tl, bo, gofs, owner = [], [], [], [], [], [0], []
            for r in rs:
                L = int(lens[r])
                for k, (c, (start, end, _)) in enumerate(zip(chains[r], geo[r])):
                    q_off.append(int(qoff[2 * r + (0 if c.strand > 0 else 1)]))
                    t_off.append(min(int(start), ref_len))
                    tl.append(max(0, min(int(end), ref_len) - int(start)))
                    ql.append(L)
                    bo.append(int(c.ref_start - c.read_start) - int(start))
                    owner.append((r, k, int(start)))
                gofs.append(len(owner))
            ql_h = np.array(ql, np.int32)
            tl_h = np.array(tl, np.int32)
            bo_h = np.array(bo, np.int32)

            def up(a):
                return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

            inp = K.SwInputs(qseq=qseq, tseq=ref_words, qmask=qmask,
                             tmask=ref_masks if has_n else None,
                             q_off=up(np.array(q_off, np.int64)),
                             t_off=up(np.array(t_off, np.int64)), qlen=up(ql_h), tlen=up(tl_h),
                             band_offset=up(bo_h), qlen_host=ql_h, tlen_host=tl_h,
                             boff_host=bo_h, has_n=has_n)
            gofs = np.array(gofs, np.int64)
            if xd > 0:
                fwd = K.sw_forward(inp, ma, mi, go, ge, hb, traceback=True, rowkeys=True)
                sc, qe, te = K.sw_xdrop(fwd, inp, gofs, xd, hb)
                tbk = K.sw_traceback(inp, fwd, sc, qe, te)
            else:
                fwd = K.sw_forward(inp, ma, mi, go, ge, hb, traceback=True, cigar=True)
                sc, qe, te, tbk = fwd.score, fwd.query_end, fwd.target_end, fwd.walk
            sc_h, qe_h, te_h = sc.cpu().numpy(), qe.cpu().numpy(), te.cpu().numpy()
            rs_h, ts_h = tbk.read_start.cpu().numpy(), tbk.ref_start.cpu().numpy()
            cnt = tbk.counts.cpu().numpy()
            cigs = K.decode_cigars(tbk)
            for p, (r, k, start) in enumerate(owner):
                L = int(lens[r])
                rs0, qe0 = int(rs_h[p]), int(qe_h[p])
                ext = ExtensionResult(
                    score=float(sc_h[p]),
                    cigar=merge_cigar([("S", rs0)] + list(cigs[p]) + [("S", L - qe0)]),
                    read_start=rs0, read_end=qe0, ref_start=int(ts_h[p]) + start,
                    ref_end=int(te_h[p]) + start, n_match=int(cnt[p, 0]),
                    n_mismatch=int(cnt[p, 1]), n_insertion=int(cnt[p, 2]),
                    n_deletion=int(cnt[p, 3]))
                out[r].append((k, ext))
        return [[e for _, e in sorted(lst, key=lambda x: x[0])] for lst in out]

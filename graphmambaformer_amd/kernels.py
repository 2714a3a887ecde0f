"""Low-level host API over libgmfa: packing, SW forward, traceback walk, chaining DP.

Everything here works on torch tensors resident on one HIP device and calls
the C ABI through ctypes; outputs are allocated by the caller side (torch), so
the library owns no device memory.  The drop-in modules in
``graphmambaformer_amd.alignment`` are built on these functions.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import SwBatch, SwOut, SwParams, SwTile, check, lib, ptr, stream_of

# ----------------------------------------------------------------------------
# 2-bit packing
# ----------------------------------------------------------------------------


def _pack_rows(codes2: np.ndarray, nmask: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """codes2 / nmask: (words*16,) uint8 -> (words,) uint32 each."""
    c = codes2.reshape(-1, 16).astype(np.uint32)
    m = nmask.reshape(-1, 16).astype(np.uint32)
    sh = (np.arange(16, dtype=np.uint32) * 2)[None, :]
    return (c << sh).sum(axis=1, dtype=np.uint64).astype(np.uint32), \
        (m << sh).sum(axis=1, dtype=np.uint64).astype(np.uint32)


def pack_host(seqs: Sequence[np.ndarray], alphabet: str = "ref"):
    """Pack sequences back to back, each starting on a 16-base word.

    alphabet "ref": codes of the reference encoder (A,C,G,T = 1..4, N = 5,
    negative = padding; seeding.py:46-80); "2bit": codes 0..3, no N.
    Returns (words uint32, masks uint32, base offsets int64, has_n).
    """
    lens = np.array([len(s) for s in seqs], dtype=np.int64)
    nwords = (lens + 15) // 16
    woff = np.zeros(len(seqs) + 1, dtype=np.int64)
    np.cumsum(nwords, out=woff[1:])
    total = int(woff[-1]) if len(seqs) else 0
    total = max(total, 1)
    c2 = np.zeros(total * 16, dtype=np.uint8)
    nm = np.ones(total * 16, dtype=np.uint8)  # padding never matches
    has_n = False
    for k, s in enumerate(seqs):
        s = np.asarray(s)
        o = int(woff[k]) * 16
        if alphabet == "2bit":
            c2[o:o + len(s)] = s
            nm[o:o + len(s)] = 0
        else:
            s16 = s.astype(np.int16)
            bad = (s16 == 0) | (s16 > 5)
            if bad.any():
                raise ValueError("base codes must be negative (padding), 1..4 (ACGT) or 5 (N)")
            ok = (s16 >= 1) & (s16 <= 4)
            c2[o:o + len(s)] = np.where(ok, s16 - 1, 0)
            nm[o:o + len(s)] = ~ok
            has_n = has_n or bool((~ok).any())
    words, masks = _pack_rows(c2, nm)
    return words, masks, woff[:-1] * 16, has_n


def _as_i32(a: np.ndarray) -> np.ndarray:
    return a.view(np.int32) if a.dtype == np.uint32 else a.astype(np.int32)


@dataclass
class SwInputs:
    """A batch of pairs resident on the device (2-bit streams + per-pair geometry)."""

    qseq: torch.Tensor
    tseq: torch.Tensor
    qmask: Optional[torch.Tensor]
    tmask: Optional[torch.Tensor]
    q_off: torch.Tensor
    t_off: torch.Tensor
    qlen: torch.Tensor
    tlen: torch.Tensor
    band_offset: torch.Tensor
    qlen_host: np.ndarray
    tlen_host: np.ndarray
    boff_host: np.ndarray
    has_n: bool

    @property
    def n_pairs(self) -> int:
        return int(len(self.qlen_host))

    @property
    def device(self) -> torch.device:
        return self.qseq.device

    def batch_struct(self, order: Optional[torch.Tensor] = None) -> SwBatch:
        b = SwBatch()
        b.n_pairs = self.n_pairs
        b.max_qlen = int(self.qlen_host.max()) if self.n_pairs else 0
        b.qseq, b.tseq = ptr(self.qseq), ptr(self.tseq)
        b.qmask = ptr(self.qmask) if self.has_n else None
        b.tmask = ptr(self.tmask) if self.has_n else None
        b.q_words, b.t_words = self.qseq.numel(), self.tseq.numel()
        b.q_off, b.t_off = ptr(self.q_off), ptr(self.t_off)
        b.qlen, b.tlen, b.band_offset = ptr(self.qlen), ptr(self.tlen), ptr(self.band_offset)
        b.order = ptr(order)
        return b


def pack_stream(seqs: Sequence[np.ndarray], device, alphabet: str = "ref"):
    """Sequences back to back (no word alignment) as one byte stream, copied
    once and packed on the device (gmfa_pack_flat).  Returns (words, masks,
    base offsets int64, has_n); codes as in pack_host, negative codes are not
    accepted here."""
    dev = torch.device(device)
    lens = np.array([len(x) for x in seqs], dtype=np.int64)
    offs = np.zeros(len(seqs) + 1, dtype=np.int64)
    np.cumsum(lens, out=offs[1:])
    n = int(offs[-1])
    flat = np.concatenate([np.asarray(x).astype(np.uint8, copy=False) for x in seqs]) \
        if n else np.zeros(0, np.uint8)
    nw = max((n + 15) // 16, 1)
    seq = torch.zeros(nw, dtype=torch.int32, device=dev)
    mask = torch.full((nw,), -1, dtype=torch.int32, device=dev)
    flags = torch.zeros(1, dtype=torch.int32, device=dev)
    if n:
        codes = torch.from_numpy(flat).to(dev)
        check(lib().gmfa_pack_flat(ptr(codes), n, 1 if alphabet == "2bit" else 0, ptr(seq),
                                   ptr(mask), ptr(flags), stream_of(dev)), "gmfa_pack_flat")
    f = int(flags.item())
    if f & 1:
        raise ValueError("base codes must be 1..4 (ACGT) or 5 (N)" if alphabet != "2bit" else
                         "2-bit codes must be 0..3")
    return seq, mask, offs[:-1], bool(f & 2)


def upload_pairs(queries: Sequence[np.ndarray], targets: Sequence[np.ndarray], band_offset,
                 device, alphabet: str = "ref") -> SwInputs:
    """Copy a batch once to the device and pack it there (queries and targets
    each one stream, back to back)."""
    ql = np.array([len(q) for q in queries], dtype=np.int32)
    tl = np.array([len(t) for t in targets], dtype=np.int32)
    neg = any(len(x) and np.asarray(x).min() < 0 for x in list(queries) + list(targets))
    if neg:  # negative padding codes inside a sequence: the host packer masks them
        return _upload_pairs_host(queries, targets, band_offset, device, alphabet)
    dev = torch.device(device)
    qseq, qmask, qo, qn = pack_stream(queries, dev, alphabet)
    tseq, tmask, to, tn = pack_stream(targets, dev, alphabet)
    has_n = qn or tn
    bo = np.asarray(band_offset, dtype=np.int32).reshape(len(queries))

    def up(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

    return SwInputs(qseq=qseq, tseq=tseq, qmask=qmask if has_n else None,
                    tmask=tmask if has_n else None, q_off=up(qo), t_off=up(to), qlen=up(ql),
                    tlen=up(tl), band_offset=up(bo), qlen_host=ql, tlen_host=tl, boff_host=bo,
                    has_n=has_n)


def _upload_pairs_host(queries, targets, band_offset, device, alphabet):
    qw, qm, qo, qn = pack_host(queries, alphabet)
    tw, tm, to, tn = pack_host(targets, alphabet)
    has_n = qn or tn
    dev = torch.device(device)
    ql = np.array([len(q) for q in queries], dtype=np.int32)
    tl = np.array([len(t) for t in targets], dtype=np.int32)
    bo = np.asarray(band_offset, dtype=np.int32).reshape(len(queries))

    def up(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

    return SwInputs(qseq=up(_as_i32(qw)), tseq=up(_as_i32(tw)),
                    qmask=up(_as_i32(qm)) if has_n else None,
                    tmask=up(_as_i32(tm)) if has_n else None,
                    q_off=up(qo), t_off=up(to), qlen=up(ql), tlen=up(tl), band_offset=up(bo),
                    qlen_host=ql, tlen_host=tl, boff_host=bo, has_n=has_n)


def pack_device(codes: torch.Tensor, lens: torch.Tensor, lens_host: np.ndarray):
    """Pack a (B, L) int16 code matrix already on the device (drop-in path)."""
    B, L = codes.shape
    nwords = (lens_host.astype(np.int64) + 15) // 16
    woff = np.zeros(B + 1, dtype=np.int64)
    np.cumsum(nwords, out=woff[1:])
    total = max(int(woff[-1]), 1)
    dev = codes.device
    seq = torch.zeros(total, dtype=torch.int32, device=dev)
    mask = torch.full((total,), -1, dtype=torch.int32, device=dev)
    bad = torch.zeros(1, dtype=torch.int32, device=dev)
    woff_d = torch.from_numpy(woff[:-1]).to(dev)
    lens32 = lens.to(device=dev, dtype=torch.int32).contiguous()
    codes = codes.to(torch.int16).contiguous()
    check(lib().gmfa_pack_codes(ptr(codes), B, L, ptr(lens32), ptr(woff_d), total, ptr(seq),
                                ptr(mask), ptr(bad), stream_of(dev)), "gmfa_pack_codes")
    return seq, mask, woff_d * 16, bad


# ----------------------------------------------------------------------------
# SW forward + traceback
# ----------------------------------------------------------------------------


def resolve_tile(half_band: int, cells_per_lane: int = 0, lanes_per_pair: int = 0) -> SwTile:
    t = SwTile()
    check(lib().gmfa_sw_resolve_tile(int(half_band), int(cells_per_lane), int(lanes_per_pair),
                                     ctypes.byref(t)), "gmfa_sw_resolve_tile")
    return t


def narrow_tile(half_band: int) -> Optional[SwTile]:
    """The 64-lane tile with the fewest cells per lane that holds the band
    (one pair per wavefront): for W = 64*C + 1, C cells per lane plus the
    scalar cell (W = 65: one cell per lane), else the smallest general
    64-lane tile; None if none holds it."""
    W = 2 * int(half_band) + 1
    shapes = sorted(C for C, G in list_tiles() if G == 64)
    cands = ([(W - 1) // 64] if (W - 1) % 64 == 0 else []) + [C for C in shapes if C * 64 >= W]
    for C in cands:
        if C not in shapes:
            continue
        try:  # a scalar-cell variant exists for some (C, W = 64*C + 1) only
            return resolve_tile(half_band, C, 64)
        except (RuntimeError, ValueError):
            continue
    return None


SKEW_LANES = 63  # lanes_per_pair of the skewed tile (two diagonals per lane)


def latency_tile(half_band: int) -> Optional[SwTile]:
    """The tile that finishes one pair soonest when each pair has a SIMD to
    itself: for bands up to 126 diagonals the skewed tile (2 x 63: two
    diagonals per lane, cells along anti-diagonals, no prefix scan per row),
    else ``narrow_tile``."""
    if 2 * int(half_band) + 1 <= 2 * SKEW_LANES:
        try:
            return resolve_tile(half_band, 2, SKEW_LANES)
        except (RuntimeError, ValueError):
            pass
    return narrow_tile(half_band)


def batch_tile(half_band: int, n_pairs: int, device=None, rowkeys: bool = False) -> SwTile:
    """Default tile for a batch of ``n_pairs``.  The library's default packs
    several short-band pairs into one wavefront (16- or 32-lane groups), the
    least padding per cell; when that leaves fewer than 4 wavefronts per SIMD,
    a latency tile (one pair per wavefront: ``latency_tile``) spreads the
    pairs over the idle SIMDs and finishes sooner (1,024 pairs of 150 bp at
    half_band 32: 0.078 ms on 2x64 and 0.069 ms on 1x64 + the scalar cell
    against 0.13 ms on 5x16).  ``rowkeys``: the launch writes per-row keys,
    which the skewed tile has no form for (a 64-lane tile then)."""
    t = resolve_tile(half_band)
    if t.lanes_per_pair == 64:
        return t
    cus = torch.cuda.get_device_properties(device).multi_processor_count if device is not None else 256
    waves = -(-int(n_pairs) // (64 // t.lanes_per_pair))
    if waves >= 4 * 4 * cus:
        return t
    return (narrow_tile(half_band) if rowkeys else latency_tile(half_band)) or t


def list_tiles() -> list[tuple[int, int]]:
    """Compiled (cells_per_lane, lanes_per_pair) tiles of the forward kernel."""
    n = lib().gmfa_sw_list_tiles(None, None, 0)
    c = (ctypes.c_int32 * n)()
    g = (ctypes.c_int32 * n)()
    lib().gmfa_sw_list_tiles(c, g, n)
    return [(int(c[k]), int(g[k])) for k in range(n)]


def tb_bytes(qlen: np.ndarray, tile: SwTile) -> np.ndarray:
    """Traceback plane bytes per pair (same formula as gmfa_sw_tb_bytes)."""
    nblk = (np.asarray(qlen, dtype=np.int64) + 7) // 8
    return nblk * tile.lanes_stored * tile.cells_per_lane * 4


def make_workspace(inp: SwInputs, tile: SwTile, traceback: bool = True,
                   cigar: bool = True) -> dict:
    """Preallocate every buffer ``sw_forward`` / ``sw_traceback`` would
    allocate for this batch (work counter, traceback planes, CIGAR runs), so
    repeated passes over a resident batch allocate nothing."""
    dev = inp.device
    n = inp.n_pairs
    ws = {"counter": torch.zeros(4, dtype=torch.int32, device=dev)}  # 4 words: split launches
    if traceback:
        nb = tb_bytes(inp.qlen_host, tile)
        offs = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(nb // 4, out=offs[1:])
        ws["tb"] = torch.empty(max(int(offs[-1]), 1), dtype=torch.int32, device=dev)
        ws["tb_off"] = torch.from_numpy(offs[:-1]).to(dev)
        ws["tb_bytes"] = int(offs[-1]) * 4
    if cigar:
        cap = inp.qlen_host.astype(np.int64) + inp.tlen_host.astype(np.int64) + 1
        coffs = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(cap, out=coffs[1:])
        ws["cigar"] = torch.empty(max(int(coffs[-1]), 1), dtype=torch.int32, device=dev)
        ws["cigar_off"] = torch.from_numpy(coffs[:-1]).to(dev)
    return ws


@dataclass
class SwForward:
    score: torch.Tensor
    query_end: torch.Tensor
    target_end: torch.Tensor
    tb: Optional[torch.Tensor]
    tb_off: Optional[torch.Tensor]
    rowkey: Optional[torch.Tensor]
    rowkey_off_host: Optional[np.ndarray]
    tile: SwTile
    params: SwParams
    walk: Optional["SwTraceback"] = None  # CIGARs when walked inside the forward launch


def make_params(ma, mi, go, ge, half_band, flags, tile: SwTile, split_rows: int = 0,
                split_slices: int = 0, prio_rows: int = 0, max_waves: int = 0) -> SwParams:
    p = SwParams()
    p.match, p.mismatch, p.gap_open, p.gap_extend = int(ma), int(mi), int(go), int(ge)
    p.half_band = int(half_band)
    p.flags = int(flags)
    p.cells_per_lane = tile.cells_per_lane
    p.lanes_per_pair = tile.lanes_per_pair
    p.split_rows = int(split_rows)
    p.split_slices = int(split_slices)
    p.prio_rows = int(prio_rows)
    p.max_waves = int(max_waves)
    return p


def split_variants(half_band: int, cells_per_lane: int = 0) -> list[tuple[int, int]]:
    """Compiled split variants for the band: (cells per lane of a slice,
    slices per pair), over the regular tile of ``cells_per_lane`` (0: the
    default tile)."""
    n = lib().gmfa_sw_split_variants(int(half_band), int(cells_per_lane), None, None, 0)
    if n <= 0:
        return []
    c = (ctypes.c_int32 * n)()
    k = (ctypes.c_int32 * n)()
    lib().gmfa_sw_split_variants(int(half_band), int(cells_per_lane), c, k, n)
    return [(int(c[x]), int(k[x])) for x in range(n)]


def resident_waves(half_band: int, tile: SwTile, cigar: bool = True) -> int:
    """Wavefronts of the forward kernel (traceback planes, fused walk when
    ``cigar``) for this band and tile that the current device holds at once."""
    flags = _lib.SW_TRACEBACK | (_lib.SW_CIGAR if cigar else 0)
    p = make_params(0, 0, 0, 0, half_band, flags, tile)
    w = ctypes.c_int32()
    check(lib().gmfa_sw_resident_waves(ctypes.byref(p), ctypes.byref(w)), "gmfa_sw_resident_waves")
    return int(w.value)


SPLIT_MAX_PAIRS = 512  # split pairs of one launch, at most (the library's bound)
SPLIT_SHARE = 0.7      # split pairs hold at least this multiple of a resident wavefront's share
SPLIT_WAVE_FRAC = 0.75  # split wavefronts (pairs x slices) within this share of the resident ones


def default_launch(inp: SwInputs, half_band: int, tile: Optional[SwTile] = None) -> dict:
    """Launch options (sw_forward keywords) for a batch of uneven lengths:
    split the pairs whose cells (query rows x band width) reach SPLIT_SHARE of
    the batch's mean cells per resident wavefront, at most SPLIT_MAX_PAIRS,
    over the most slices compiled for the band whose wavefronts (pairs x
    slices) stay within SPLIT_WAVE_FRAC of the resident ones (the fewest
    slices when none do); the plain launch when no pair qualifies, the band
    has no split variant or the batch holds N masks.  One wavefront saturates
    its SIMD, so a pair far above the share sets the batch's end, and the
    split pairs, started first, must leave room for the rest.  Measured on the
    noisy 8,192-pair batch at half_band 751 (resident wavefronts 2,048): the
    512 longest (37-100 kb) over 3 wavefronts each 107.2 ms, over 4 (all
    2,048 wavefronts) 110.7 ms, the 256 longest over 4 110.2 ms, plain
    121 ms; splitting only the 40 longest (all ~100 kb) no faster than plain."""
    tile = tile or batch_tile(half_band, inp.n_pairs, inp.device)
    variants = split_variants(half_band, tile.cells_per_lane) if not inp.has_n else []
    plain = {"split_rows": 0, "split_slices": 0}
    if not variants or tile.lanes_per_pair != 64 or inp.n_pairs == 0:
        return plain
    rows = np.sort(np.asarray(inp.qlen_host, np.int64))[::-1]
    waves = max(1, resident_waves(half_band, tile))
    share = rows.sum() / waves
    k = min(int((rows >= SPLIT_SHARE * share).sum()), SPLIT_MAX_PAIRS)
    if k == 0:
        return plain
    counts = sorted({sl for _, sl in variants})
    fit = [sl for sl in counts if k * sl <= SPLIT_WAVE_FRAC * waves]
    return {"split_rows": int(rows[k - 1]), "split_slices": max(fit) if fit else counts[0]}


def _cigar_buffers(inp: SwInputs, ws: dict):
    n = inp.n_pairs
    dev = inp.device
    if "cigar" in ws:
        cig, cig_off = ws["cigar"], ws["cigar_off"]
    else:
        cap = inp.qlen_host.astype(np.int64) + inp.tlen_host.astype(np.int64) + 1
        offs = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(cap, out=offs[1:])
        cig = torch.empty(max(int(offs[-1]), 1), dtype=torch.int32, device=dev)
        cig_off = torch.from_numpy(offs[:-1]).to(dev)
    return SwTraceback(cigar=cig, cigar_off=cig_off,
                       n_cigar=torch.empty(n, dtype=torch.int32, device=dev),
                       read_start=torch.empty(n, dtype=torch.int32, device=dev),
                       ref_start=torch.empty(n, dtype=torch.int32, device=dev),
                       counts=torch.empty((n, 4), dtype=torch.int32, device=dev))


def sw_forward(inp: SwInputs, ma: int, mi: int, go: int, ge: int, half_band: int,
               traceback: bool = True, rowkeys: bool = False,
               order: Optional[torch.Tensor] = None, tile: Optional[SwTile] = None,
               workspace: Optional[dict] = None, cigar: bool = False, split_rows: int = 0,
               split_slices: int = 0, prio_rows: int = 0, max_waves: int = 0) -> SwForward:
    """Forward pass.  ``cigar=True`` also walks every pair inside the same
    launch (GMFA_SW_CIGAR) and returns the CIGARs in ``.walk``.
    ``split_rows > 0`` (with ``cigar``, ``order`` and no N masks): pairs of at
    least that many query rows run on several wavefronts each, a slice of the
    band per wavefront (``split_slices`` of them, 0 = the first compiled
    count); the outputs are those of the plain launch.  ``prio_rows > 0``:
    pairs of at least that many query rows run at raised wave priority.
    ``max_waves > 0`` caps the persistent launch's wavefronts."""
    dev = inp.device
    n = inp.n_pairs
    tile = tile or batch_tile(half_band, n, dev, rowkeys=rowkeys)
    flags = 0
    if traceback:
        flags |= _lib.SW_TRACEBACK
    if inp.has_n:
        flags |= _lib.SW_NMASK
    if rowkeys:
        flags |= _lib.SW_ROWKEYS
    if cigar:
        if not traceback or rowkeys:
            raise ValueError("cigar=True needs traceback=True and no row keys")
        flags |= _lib.SW_CIGAR
    ws = workspace if workspace is not None else {}
    score = torch.empty(n, dtype=torch.int32, device=dev)
    qend = torch.empty(n, dtype=torch.int32, device=dev)
    tend = torch.empty(n, dtype=torch.int32, device=dev)
    tb = tb_off = rk = rk_off = None
    rk_off_host = None
    if traceback:
        if "tb" in ws:
            tb, tb_off = ws["tb"], ws["tb_off"]
        else:
            nb = tb_bytes(inp.qlen_host, tile)
            offs = np.zeros(n + 1, dtype=np.int64)
            np.cumsum(nb // 4, out=offs[1:])
            tb = torch.empty(max(int(offs[-1]), 1), dtype=torch.int32, device=dev)
            tb_off = torch.from_numpy(offs[:-1]).to(dev)
    if rowkeys:
        per = inp.qlen_host.astype(np.int64)
        offs = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(per, out=offs[1:])
        rk = torch.empty(max(int(offs[-1]), 1), dtype=torch.int32, device=dev)
        rk_off = torch.from_numpy(offs[:-1]).to(dev)
        rk_off_host = offs
    counter = ws.get("counter")
    if counter is None:
        counter = torch.zeros(4, dtype=torch.int32, device=dev)
    if split_rows and counter.numel() < 4:
        raise ValueError("split launches need a work counter of 4 int32 words")
    out = SwOut()
    out.score, out.query_end, out.target_end = ptr(score), ptr(qend), ptr(tend)
    out.tb, out.tb_off = ptr(tb), ptr(tb_off)
    out.rowkey, out.rowkey_off = ptr(rk), ptr(rk_off)
    out.work_counter = ptr(counter)
    walk = None
    if cigar:
        walk = _cigar_buffers(inp, ws)
        out.cigar_off, out.cigar = ptr(walk.cigar_off), ptr(walk.cigar)
        out.n_cigar, out.read_start = ptr(walk.n_cigar), ptr(walk.read_start)
        out.ref_start, out.counts = ptr(walk.ref_start), ptr(walk.counts)
    params = make_params(ma, mi, go, ge, half_band, flags, tile, split_rows, split_slices,
                         prio_rows, max_waves)
    batch = inp.batch_struct(order)
    check(lib().gmfa_sw_forward(ctypes.byref(batch), ctypes.byref(params), ctypes.byref(out),
                                stream_of(dev)), "gmfa_sw_forward")
    params.split_rows = params.split_slices = params.prio_rows = params.max_waves = 0
    params.flags = int(flags) & ~_lib.SW_CIGAR  # a later standalone walk uses the same params
    return SwForward(score=score, query_end=qend, target_end=tend, tb=tb, tb_off=tb_off,
                     rowkey=rk, rowkey_off_host=rk_off_host, tile=tile, params=params, walk=walk)


def sw_forward_noalloc(inp: SwInputs, ma: int, mi: int, go: int, ge: int, half_band: int,
                       order: Optional[torch.Tensor] = None, tile: Optional[SwTile] = None,
                       outputs: Optional[dict] = None) -> SwForward:
    """Forward pass with the CIGAR walk fused, through gmfa_sw_forward_noalloc:
    the work counter and the traceback planes live in the library's static
    device scratch (one slab per wavefront), so the call allocates no device
    memory of its own.  ``outputs`` may hold preallocated output tensors
    (keys score, query_end, target_end and those of ``_cigar_buffers``'s
    record: cigar, cigar_off, n_cigar, read_start, ref_start, counts); the
    missing ones are allocated here by torch.  Calls that overlap on one GPU
    share the scratch and race."""
    dev = inp.device
    n = inp.n_pairs
    tile = tile or batch_tile(half_band, n, dev)
    flags = _lib.SW_TRACEBACK | _lib.SW_CIGAR
    if inp.has_n:
        flags |= _lib.SW_NMASK
    o = dict(outputs or {})
    for key in ("score", "query_end", "target_end", "n_cigar", "read_start", "ref_start"):
        if key not in o:
            o[key] = torch.empty(n, dtype=torch.int32, device=dev)
    if "counts" not in o:
        o["counts"] = torch.empty((n, 4), dtype=torch.int32, device=dev)
    if "cigar" not in o:
        cap = inp.qlen_host.astype(np.int64) + inp.tlen_host.astype(np.int64) + 1
        offs = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(cap, out=offs[1:])
        o["cigar"] = torch.empty(max(int(offs[-1]), 1), dtype=torch.int32, device=dev)
        o["cigar_off"] = torch.from_numpy(offs[:-1]).to(dev)
    out = SwOut()
    out.score, out.query_end, out.target_end = ptr(o["score"]), ptr(o["query_end"]), \
        ptr(o["target_end"])
    out.cigar_off, out.cigar = ptr(o["cigar_off"]), ptr(o["cigar"])
    out.n_cigar, out.read_start = ptr(o["n_cigar"]), ptr(o["read_start"])
    out.ref_start, out.counts = ptr(o["ref_start"]), ptr(o["counts"])
    params = make_params(ma, mi, go, ge, half_band, flags, tile)
    batch = inp.batch_struct(order)
    check(lib().gmfa_sw_forward_noalloc(ctypes.byref(batch), ctypes.byref(params),
                                        ctypes.byref(out), stream_of(dev)),
          "gmfa_sw_forward_noalloc")
    walk = SwTraceback(cigar=o["cigar"], cigar_off=o["cigar_off"], n_cigar=o["n_cigar"],
                       read_start=o["read_start"], ref_start=o["ref_start"], counts=o["counts"])
    return SwForward(score=o["score"], query_end=o["query_end"], target_end=o["target_end"],
                     tb=None, tb_off=None, rowkey=None, rowkey_off_host=None, tile=tile,
                     params=params, walk=walk)


class SwGraph:
    """One batch step on a resident batch -- the work-counter reset and the
    persistent forward kernel, with the CIGAR walk fused when ``cigar`` --
    captured in a HIP graph: ``replay()`` relaunches it with no host work per
    launch.  Inputs, workspace and outputs keep their device addresses, so new
    pairs with the same lengths and offsets can be written into ``inp``'s
    tensors between replays.  The library's own launches hold no memset
    nodes (the counter is reset by a kernel), and every buffer is allocated
    before or during capture, none by a replay."""

    def __init__(self, inp: SwInputs, scoring: tuple, half_band: int,
                 order: Optional[torch.Tensor] = None, tile: Optional[SwTile] = None,
                 workspace: Optional[dict] = None, cigar: bool = True, split_rows: int = 0,
                 split_slices: int = 0, prio_rows: int = 0, max_waves: int = 0):
        dev = inp.device
        self.tile = tile or batch_tile(half_band, inp.n_pairs, dev)
        self.ws = workspace if workspace is not None else make_workspace(inp, self.tile,
                                                                        cigar=cigar)

        def step():
            return sw_forward(inp, *scoring, half_band, traceback=True, order=order,
                              tile=self.tile, workspace=self.ws, cigar=cigar,
                              split_rows=split_rows, split_slices=split_slices,
                              prio_rows=prio_rows, max_waves=max_waves)

        # one eager pass on a side stream first (module load, allocator warm-up)
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            step()
        torch.cuda.current_stream(dev).wait_stream(side)
        torch.cuda.synchronize(dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.out = step()

    def replay(self) -> SwForward:
        self.graph.replay()
        return self.out


@dataclass
class SwTraceback:
    cigar: torch.Tensor      # flat BAM runs (len << 4 | op), int32 view
    cigar_off: torch.Tensor  # int64 [n]
    n_cigar: torch.Tensor
    read_start: torch.Tensor
    ref_start: torch.Tensor
    counts: torch.Tensor     # (n, 4): =, X, I, D


def sw_traceback(inp: SwInputs, fwd: SwForward, score=None, query_end=None, target_end=None,
                 workspace: Optional[dict] = None) -> SwTraceback:
    if fwd.tb is None:
        raise ValueError("traceback needs the forward pass with traceback=True")
    dev = inp.device
    ws = workspace if workspace is not None else {}
    bufs = _cigar_buffers(inp, ws)
    cig, cig_off = bufs.cigar, bufs.cigar_off
    ncig, rs, ts, counts = bufs.n_cigar, bufs.read_start, bufs.ref_start, bufs.counts
    score = fwd.score if score is None else score
    query_end = fwd.query_end if query_end is None else query_end
    target_end = fwd.target_end if target_end is None else target_end
    batch = inp.batch_struct()
    check(lib().gmfa_sw_traceback(ctypes.byref(batch), ctypes.byref(fwd.params), ptr(score),
                                  ptr(query_end), ptr(target_end), ptr(fwd.tb), ptr(fwd.tb_off),
                                  ptr(cig_off), ptr(cig), ptr(ncig), ptr(rs), ptr(ts),
                                  ptr(counts), stream_of(dev)), "gmfa_sw_traceback")
    return SwTraceback(cigar=cig, cigar_off=cig_off, n_cigar=ncig, read_start=rs, ref_start=ts,
                       counts=counts)


def decode_cigars(tbk: SwTraceback) -> list[list[tuple[str, int]]]:
    """Device CIGAR runs -> list of [(op, n)] per pair (the runs are gathered
    on the device, so only they cross to the host).  A pair without outputs
    (n_cigar < 0: a served request that found no traceback pages) decodes as
    an empty list."""
    off = tbk.cigar_off.to(torch.int64)
    nc = tbk.n_cigar.to(torch.int64).clamp_min(0)
    n = int(nc.numel())
    if n == 0:
        return []
    total = int(nc.sum().item())
    if total:
        starts = torch.cumsum(nc, 0) - nc
        owner = torch.repeat_interleave(torch.arange(n, device=nc.device), nc)
        within = torch.arange(total, device=nc.device) - starts[owner]
        runs = tbk.cigar[off[owner] + within].cpu().numpy().view(np.uint32)
    else:
        runs = np.zeros(0, np.uint32)
    nc_h = nc.cpu().numpy()
    ops = (runs & 15).astype(np.int64)
    lens = (runs >> 4).astype(np.int64)
    chars = np.array([_lib.OP_CHAR.get(c, "?") for c in range(16)])[ops]
    out = []
    x = 0
    for p in range(n):
        k = int(nc_h[p])
        out.append(list(zip(chars[x:x + k].tolist(), lens[x:x + k].tolist())))
        x += k
    return out


def sw_xdrop(fwd: SwForward, inp: SwInputs, group_off: np.ndarray, x_drop: float, hb: int):
    """Batch-wide x-drop on the device, one reference batch per group
    (pairs group_off[g] .. group_off[g+1]-1).  Returns device int32 (score,
    query_end, target_end)."""
    if fwd.rowkey is None:
        raise ValueError("x-drop needs the forward pass with rowkeys=True")
    go = np.asarray(group_off, dtype=np.int64)
    sizes = np.diff(go)
    if len(go) < 1 or go[0] != 0 or go[-1] != inp.n_pairs or (sizes < 0).any():
        raise ValueError("group_off must run from 0 to n_pairs, nondecreasing")
    ng = len(go) - 1
    dev = inp.device
    n = inp.n_pairs
    score = torch.zeros(n, dtype=torch.int32, device=dev)
    qe = torch.zeros(n, dtype=torch.int32, device=dev)
    te = torch.zeros(n, dtype=torch.int32, device=dev)
    if ng == 0 or n == 0:
        return score, qe, te
    mg = int(sizes.max())
    thr = int(np.floor(x_drop)) if x_drop > 0 else -1
    rk_off = torch.from_numpy(fwd.rowkey_off_host[:-1].astype(np.int64)).to(dev)
    god = torch.from_numpy(go.astype(np.int32)).to(dev)
    scratch = torch.empty(4 * n, dtype=torch.int32, device=dev) if mg > 64 else None
    check(lib().gmfa_sw_xdrop(ptr(fwd.rowkey), ptr(rk_off), ptr(inp.qlen), ptr(inp.band_offset),
                              ptr(god), ng, max(mg, 1), thr, int(hb), ptr(scratch), ptr(score),
                              ptr(qe), ptr(te), stream_of(dev)), "gmfa_sw_xdrop")
    return score, qe, te


# ----------------------------------------------------------------------------
# chaining
# ----------------------------------------------------------------------------


def chain_dp(read_end: torch.Tensor, ref_end: torch.Tensor, weight: torch.Tensor,
             seg_off: torch.Tensor, max_lookback: int, max_gap: int,
             pen_lut: Optional[torch.Tensor] = None, gap_open: int = 0, gap_extend: int = 0,
             bonus: Optional[torch.Tensor] = None):
    """Segmented chaining DP on the device.

    float64 weights -> float64 profile with ``pen_lut``; int32 weights -> the
    integer profile (penalty = open + extend*gap)."""
    dev = read_end.device
    n_seg = seg_off.numel() - 1
    n = read_end.numel()
    parent = torch.empty(n, dtype=torch.int64, device=dev)
    if weight.dtype == torch.float64:
        if pen_lut is None:
            raise ValueError("float64 chaining needs pen_lut")
        f = torch.empty(n, dtype=torch.float64, device=dev)
        check(lib().gmfa_chain_dp_f64(ptr(read_end), ptr(ref_end), ptr(weight), ptr(seg_off),
                                      n_seg, int(max_lookback), int(max_gap), ptr(pen_lut),
                                      ptr(bonus), ptr(f), ptr(parent), stream_of(dev)),
              "gmfa_chain_dp_f64")
    elif weight.dtype == torch.int32:
        f = torch.empty(n, dtype=torch.int32, device=dev)
        check(lib().gmfa_chain_dp_i32(ptr(read_end), ptr(ref_end), ptr(weight), ptr(seg_off),
                                      n_seg, int(max_lookback), int(max_gap), int(gap_open),
                                      int(gap_extend), ptr(f), ptr(parent), stream_of(dev)),
              "gmfa_chain_dp_i32")
    else:
        raise ValueError("weight must be float64 or int32")
    return f, parent


def chain_backtrack(f: np.ndarray, parent: np.ndarray, min_anchors: int, min_score: float):
    """Host AffineChainer._backtrack through the library (members in read order)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    parent = np.ascontiguousarray(parent, dtype=np.int64)
    n = len(f)
    members = np.zeros(max(n, 1), dtype=np.int64)
    starts = np.zeros(n + 1, dtype=np.int64)
    scores = np.zeros(max(n, 1), dtype=np.float64)
    nc = check(lib().gmfa_chain_backtrack(n, f.ctypes.data, parent.ctypes.data, int(min_anchors),
                                          float(min_score), members.ctypes.data,
                                          starts.ctypes.data, scores.ctypes.data),
               "gmfa_chain_backtrack")
    return [(members[starts[c]:starts[c + 1]], float(scores[c])) for c in range(nc)]


@dataclass
class ChainBatchResult:
    """Chains of a batch of reads from the device pipeline (host arrays)."""

    n_kept: np.ndarray     # [n_reads]
    kept: np.ndarray       # [n_reads, max_chains] chain record indices
    score: np.ndarray      # per chain record
    span: np.ndarray       # [records, 4] read_start, read_end, ref_start, ref_end
    mem: np.ndarray        # [records, 2] member offset, count
    members: np.ndarray    # anchor ids (read-local), read order per chain
    strand: np.ndarray     # per chain record: +1 / -1


def chain_upload(read_end, ref_end, read_pos, ref_pos, weight, anchor_id, seg_off, read_seg,
                 read_base, device) -> dict:
    """Copy a batch of segments (anchors back to back in their original order)
    to the device for chain_run."""
    dev = torch.device(device)

    def up(a, dt):
        return torch.from_numpy(np.ascontiguousarray(np.asarray(a).astype(dt))).to(dev)

    return {"re": up(read_end, np.int64), "fe": up(ref_end, np.int64),
            "w": up(weight, np.float64), "rp": up(read_pos, np.int64),
            "fp": up(ref_pos, np.int64), "id": up(anchor_id, np.int64),
            "seg_off": up(seg_off, np.int64), "read_seg": up(read_seg, np.int32),
            "read_base": up(read_base, np.int64), "n": int(len(read_end)),
            "n_seg": int(len(seg_off) - 1), "n_reads": int(len(read_seg))}


def chain_run(u: dict, max_lookback: int, max_gap: int, pen_lut: torch.Tensor, min_anchors: int,
              min_score: float, score_ratio: float, max_overlap: float, max_chains: int) -> dict:
    """Sort into DP order, DP, backtrack and select for an uploaded batch, all
    on the device (results stay there)."""
    n, ns, nr = u["n"], u["n_seg"], u["n_reads"]
    dev = u["re"].device
    src = [u["re"], u["fe"], u["w"], u["rp"], u["fp"], u["id"]]
    o_re, o_fe = torch.empty_like(src[0]), torch.empty_like(src[1])
    o_w, o_rp = torch.empty_like(src[2]), torch.empty_like(src[3])
    o_fp, o_id = torch.empty_like(src[4]), torch.empty_like(src[5])
    scratch = torch.empty(max(2 * n, 1), dtype=torch.int32, device=dev)
    s = stream_of(dev)
    check(lib().gmfa_chain_sort(ptr(u["seg_off"]), ns, n, *[ptr(t) for t in src], ptr(scratch),
                                ptr(o_re), ptr(o_fe), ptr(o_w), ptr(o_rp), ptr(o_fp), ptr(o_id),
                                s), "gmfa_chain_sort")
    f, parent = chain_dp(o_re, o_fe, o_w, u["seg_off"], max_lookback, max_gap, pen_lut=pen_lut)
    # every output is written where it is read (n_chains and n_kept for every
    # segment / read, records and members for the kept chains): no zero fill
    # except the parent-error flag
    r = {"claimed": torch.empty(max(n, 1), dtype=torch.uint8, device=dev),
         "n_chains": torch.empty(max(ns, 1), dtype=torch.int32, device=dev),
         "score": torch.empty(max(n, 1), dtype=torch.float64, device=dev),
         "span": torch.empty((max(n, 1), 4), dtype=torch.int64, device=dev),
         "mem": torch.empty((max(n, 1), 2), dtype=torch.int64, device=dev),
         "members": torch.empty(max(n, 1), dtype=torch.int64, device=dev),
         "bad": torch.zeros(1, dtype=torch.int32, device=dev),
         "n_kept": torch.empty(max(nr, 1), dtype=torch.int32, device=dev),
         "kept": torch.empty((max(nr, 1), int(max_chains)), dtype=torch.int64, device=dev)}
    check(lib().gmfa_chain_backtrack_device(ptr(u["seg_off"]), ns, n, ptr(f), ptr(parent),
                                            ptr(o_rp), ptr(o_re), ptr(o_fp), ptr(o_fe), ptr(o_id),
                                            int(min_anchors), float(min_score), ptr(scratch),
                                            ptr(r["claimed"]), ptr(r["n_chains"]),
                                            ptr(r["score"]), ptr(r["span"]), ptr(r["mem"]),
                                            ptr(r["members"]), ptr(r["bad"]), s),
          "gmfa_chain_backtrack_device")
    sel_scratch = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
    check(lib().gmfa_chain_select(nr, ptr(u["read_seg"]), ptr(u["seg_off"]), ptr(r["n_chains"]),
                                  ptr(r["score"]), ptr(r["span"]), float(score_ratio),
                                  float(max_overlap), int(max_chains), ptr(sel_scratch),
                                  ptr(u["read_base"]), ptr(r["n_kept"]), ptr(r["kept"]), s),
          "gmfa_chain_select")
    return r


def chain_pipeline(read_end: np.ndarray, ref_end: np.ndarray, read_pos: np.ndarray,
                   ref_pos: np.ndarray, weight: np.ndarray, anchor_id: np.ndarray,
                   seg_off: np.ndarray, seg_strand: np.ndarray, read_seg: np.ndarray,
                   read_base: np.ndarray, max_lookback: int, max_gap: int, pen_lut: torch.Tensor,
                   min_anchors: int, min_score: float, score_ratio: float, max_overlap: float,
                   max_chains: int, device) -> ChainBatchResult:
    """Upload, chain_run, one readback of the results."""
    u = chain_upload(read_end, ref_end, read_pos, ref_pos, weight, anchor_id, seg_off, read_seg,
                     read_base, device)
    r = chain_run(u, max_lookback, max_gap, pen_lut, min_anchors, min_score, score_ratio,
                  max_overlap, max_chains)
    if int(r["bad"].item()) != 0:
        raise RuntimeError("chaining DP produced a parent outside its segment")
    n, ns, nr = u["n"], u["n_seg"], u["n_reads"]
    seg_of_rec = np.repeat(np.arange(ns), np.diff(np.asarray(seg_off, np.int64)))
    return ChainBatchResult(n_kept=r["n_kept"].cpu().numpy()[:nr],
                            kept=r["kept"].cpu().numpy()[:nr], score=r["score"].cpu().numpy()[:n],
                            span=r["span"].cpu().numpy()[:n], mem=r["mem"].cpu().numpy()[:n],
                            members=r["members"].cpu().numpy()[:n],
                            strand=np.asarray(seg_strand)[seg_of_rec] if n else np.zeros(0))


@dataclass
class WfaResult:
    distance: np.ndarray  # host, -1 above the budget
    walk: "SwTraceback"   # device CIGAR runs / counts (read/ref starts unused)
    status: torch.Tensor  # device int32, nonzero if a traceback left the first wavefront


def wfa_align_device(inp: SwInputs, max_distance: int) -> WfaResult:
    """Batched unit-cost wavefront alignment (extension.py:308-463) of every
    pair of ``inp``, results left on the device.  Two launches: the
    distances (read back to size the histories), then the alignments."""
    dev = inp.device
    n = inp.n_pairs
    ql = inp.qlen_host.astype(np.int64)
    tl = inp.tlen_host.astype(np.int64)
    if n == 0:
        z = torch.zeros(0, dtype=torch.int32, device=dev)
        return WfaResult(np.zeros(0, np.int32), SwTraceback(z, z.long(), z, z, z, z.view(0, 4)),
                         torch.zeros(1, dtype=torch.int32, device=dev))
    dist = torch.empty(n, dtype=torch.int32, device=dev)
    cap = int(min(int(max_distance), int(np.maximum(ql, tl).max())))
    scratch = torch.empty(n * 2 * (2 * cap + 3), dtype=torch.int32, device=dev)
    b = inp.batch_struct()
    s = stream_of(dev)
    check(lib().gmfa_wfa_distance(ctypes.byref(b), int(max_distance), cap, ptr(scratch),
                                  ptr(dist), s), "gmfa_wfa_distance")
    d_h = dist.cpu().numpy()
    del scratch
    both = (ql > 0) & (tl > 0) & (d_h >= 0)
    hsz = np.where(both, (d_h.astype(np.int64) + 1) ** 2, 0)
    hoff = np.zeros(n + 1, np.int64)
    np.cumsum(hsz, out=hoff[1:])
    hist = torch.empty(max(int(hoff[-1]), 1), dtype=torch.int32, device=dev)
    coff = np.zeros(n + 1, np.int64)
    np.cumsum(ql + tl + 1, out=coff[1:])
    cig = torch.empty(max(int(coff[-1]), 1), dtype=torch.int32, device=dev)
    ncig = torch.zeros(n, dtype=torch.int32, device=dev)
    counts = torch.zeros((n, 4), dtype=torch.int32, device=dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    d_hoff = torch.from_numpy(hoff[:-1]).to(dev)
    d_coff = torch.from_numpy(coff[:-1]).to(dev)
    check(lib().gmfa_wfa_align(ctypes.byref(b), ptr(dist), ptr(d_hoff), ptr(hist), ptr(d_coff),
                               ptr(cig), ptr(ncig), ptr(counts), ptr(status), s), "gmfa_wfa_align")
    walk = SwTraceback(cigar=cig, cigar_off=d_coff, n_cigar=ncig, read_start=ncig,
                       ref_start=ncig, counts=counts)
    return WfaResult(distance=d_h, walk=walk, status=status)


def wfa_align(inp: SwInputs, max_distance: int):
    """wfa_align_device + host decoding: (distance, CIGARs (None above the
    budget), op counts)."""
    r = wfa_align_device(inp, max_distance)
    n = len(r.distance)
    if n == 0:
        return r.distance, [], np.zeros((0, 4), np.int32)
    if int(r.status.item()) != 0:
        raise RuntimeError("WFA traceback left the first wavefront")
    cigs = decode_cigars(r.walk)
    cnt = r.walk.counts.cpu().numpy()
    return r.distance, [None if r.distance[p] < 0 else cigs[p] for p in range(n)], cnt


def cigar_bam(cigar: torch.Tensor, cigar_off: torch.Tensor, n_cigar: torch.Tensor,
              read_start: torch.Tensor, read_end: torch.Tensor, read_len: torch.Tensor,
              ref_start: torch.Tensor, ref_end: torch.Tensor, strand: torch.Tensor,
              n_cigar_host: np.ndarray):
    """BAM CIGAR records on the device (gmfa_cigar_bam): returns the runs
    (len << 4 | op, int32 view), their offsets, counts and flags."""
    dev = cigar.device
    n = int(len(n_cigar_host))
    cap = np.asarray(n_cigar_host, np.int64) + 2
    off = np.zeros(n + 1, np.int64)
    np.cumsum(cap, out=off[1:])
    out = torch.zeros(max(int(off[-1]), 1), dtype=torch.int32, device=dev)
    d_off = torch.from_numpy(off[:-1]).to(dev)
    out_n = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
    flag = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)

    def i32(t):
        return t.to(device=dev, dtype=torch.int32).contiguous()

    def i64(t):
        return t.to(device=dev, dtype=torch.int64).contiguous()

    args = [i32(read_start), i32(read_end), i32(read_len), i64(ref_start), i64(ref_end),
            i32(strand)]
    check(lib().gmfa_cigar_bam(n, ptr(cigar), ptr(i64(cigar_off)), ptr(i32(n_cigar)),
                               *[ptr(t) for t in args], ptr(out), ptr(d_off), ptr(out_n),
                               ptr(flag), stream_of(dev)), "gmfa_cigar_bam")
    return out, d_off, out_n[:n], flag[:n]


def autotune_launch(inp: SwInputs, half_band: int, scoring: tuple, order: torch.Tensor,
                    tile: Optional[SwTile] = None, workspace: Optional[dict] = None,
                    tops=(128, 256, 384, 512), reps: int = 3, margin: float = 0.03):
    """Choose how a batch of uneven lengths is launched: the plain launch,
    the plain launch with the ``k`` longest pairs at raised wave priority
    (``prio_rows``), or a split launch (the ``k`` longest pairs over several
    wavefronts, every compiled split variant), for ``k`` in ``tops`` (a split
    launch only shortens the tail once it takes every pair whose cells are
    well above one SIMD's share of the batch: on the noisy 8,192-pair batch
    the 40 longest pairs are all ~100 kb, and splitting fewer than ~100 pairs
    measured no faster than the plain launch).  Each
    candidate is timed on the whole batch itself (the tail these options
    shorten is a property of the whole batch, not of a sample), the median of
    ``reps`` launches (single launches of one configuration spread by several
    percent on an uneven batch); the plain launch is kept unless another is
    faster by more than ``margin``.  Returns (keyword arguments for sw_forward:
    split_rows, split_slices, prio_rows; {label: ms})."""
    dev = inp.device
    tile = tile or batch_tile(half_band, inp.n_pairs, dev)
    ws = workspace if workspace is not None else make_workspace(inp, tile, cigar=True)
    s = torch.cuda.current_stream(dev)

    def best_ms(kw):
        sw_forward(inp, *scoring, half_band, traceback=True, order=order, tile=tile,
                   workspace=ws, cigar=True, **kw)
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(reps)]
        for e0, e1 in ev:
            e0.record(s)
            sw_forward(inp, *scoring, half_band, traceback=True, order=order, tile=tile,
                       workspace=ws, cigar=True, **kw)
            e1.record(s)
        torch.cuda.synchronize(dev)
        return round(float(np.median([e0.elapsed_time(e1) for e0, e1 in ev])), 4)

    plain = {"split_rows": 0, "split_slices": 0, "prio_rows": 0}
    timings = {"plain": best_ms(plain)}
    best, best_t = plain, timings["plain"] * (1.0 - margin)
    ql = np.sort(np.asarray(inp.qlen_host))[::-1]
    cands = []
    for k in tops:
        if k <= len(ql):
            cands.append((f"prio_top{k}", dict(plain, prio_rows=int(ql[k - 1]))))
    variants = split_variants(half_band, tile.cells_per_lane) if not inp.has_n else []
    for _, slices in variants:
        for k in tops:
            if k <= len(ql):
                cands.append((f"split{slices}_top{k}",
                              dict(plain, split_rows=int(ql[k - 1]), split_slices=slices)))
    for label, kw in cands:
        ms = best_ms(kw)
        timings[label] = ms
        if ms < best_t:
            best, best_t = kw, ms
    return best, timings


def _tile_ms(inp: SwInputs, half_band: int, scoring: tuple, tile: SwTile, reps: int,
             order: Optional[torch.Tensor]) -> float:
    """Best of ``reps`` timed launches of the forward pass with the fused
    walk on ``inp`` with ``tile`` (after one untimed launch)."""
    ws = make_workspace(inp, tile, cigar=True)
    sw_forward(inp, *scoring, half_band, traceback=True, order=order, tile=tile, workspace=ws,
               cigar=True)
    s = torch.cuda.current_stream(inp.device)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    for e0, e1 in ev:
        e0.record(s)
        sw_forward(inp, *scoring, half_band, traceback=True, order=order, tile=tile,
                   workspace=ws, cigar=True)
        e1.record(s)
    torch.cuda.synchronize(inp.device)
    del ws
    return min(e0.elapsed_time(e1) for e0, e1 in ev)


def autotune_tile(inp: SwInputs, half_band: int, scoring: tuple, reps: int = 3,
                  order: Optional[torch.Tensor] = None, margin: float = 0.03,
                  confirm: Optional[tuple] = None):
    """Time every compiled tile that holds the band (the exact/specialised
    variant where one exists, e.g. 8x64 plus the scalar cell for W = 513) on
    ``inp`` with the fused walk; returns the chosen tile and
    {(cells_per_lane, lanes_per_pair): ms}, each the best of ``reps`` timed
    launches (the least disturbed one).  The tile an untuned call runs
    (``batch_tile(half_band, n_pairs)``) is kept unless another tile is faster
    by more than ``margin``.

    ``confirm`` = (whole-batch SwInputs, its order) when ``inp`` is a sample:
    the sample's two fastest tiles and the whole batch's default are timed
    again on the whole batch and the choice is made there
    ({("full", C, G): ms} join the timings).  A sample ranks near-equal tiles
    unreliably even when it spans the length distribution: it holds fewer
    wavefronts per SIMD than the batch, so a tile that holds more pairs per
    wavefront or needs fewer registers is favoured differently, and the
    batch's tail (its longest pairs against the rest) is not the sample's.
    W = 513: 9x64 24.82 ms against 24.94 ms for 8x64 plus the scalar cell on
    every 4th pair of the long-read batch by cell count, then 74.8 ms against
    66.2 ms on the whole batch."""
    timings = {}
    default = batch_tile(half_band, inp.n_pairs, inp.device)
    same = lambda t, u: (t.cells_per_lane, t.lanes_per_pair, t.lanes_stored) == \
        (u.cells_per_lane, u.lanes_per_pair, u.lanes_stored)  # noqa: E731
    ranked = []
    for C, G in list_tiles():
        try:  # the specialised variant for this band, or the general tile if it holds it
            tile = resolve_tile(half_band, C, G)
        except (RuntimeError, ValueError):
            continue
        ms = _tile_ms(inp, half_band, scoring, tile, reps, order)
        timings[(C, G)] = round(ms, 4)
        ranked.append((ms, tile))
    ranked.sort(key=lambda x: x[0])
    if confirm is not None:
        full, full_order = confirm
        cands = [t for _, t in ranked[:2]]
        fdefault = batch_tile(half_band, full.n_pairs, full.device)
        if not any(same(t, fdefault) for t in cands):
            cands.append(fdefault)
        fms = []
        for t in cands:
            ms = _tile_ms(full, half_band, scoring, t, reps, full_order)
            timings[("full", t.cells_per_lane, t.lanes_per_pair)] = round(ms, 4)
            fms.append((ms, t))
        ranked, default = sorted(fms, key=lambda x: x[0]), fdefault
    best_ms, best = ranked[0]
    for ms, t in ranked:
        if same(t, default) and ms <= best_ms * (1.0 + margin):
            return t, timings
    return best, timings

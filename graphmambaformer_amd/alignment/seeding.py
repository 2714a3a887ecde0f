"""Base alphabet of the path (graphmambaformer/alignment/seeding.py:46-80).

A, C, G, T -> 1..4; anything else -> N (5); 0 is reserved (sentinel); padding
is negative.  The device kernels store bases 2-bit packed (code - 1) with a
separate N mask; ``pack`` lives in ``graphmambaformer_amd.kernels``.
"""

from __future__ import annotations

import numpy as np

SENTINEL = 0
BASE_CODES = {"A": 1, "C": 2, "G": 3, "T": 4}
N_CODE = 5
ALPHABET_SIZE = 6

_COMPLEMENT = np.array([0, 4, 3, 2, 1, 5], dtype=np.int8)
_ENCODE = np.full(256, N_CODE, dtype=np.int8)
for _b, _c in BASE_CODES.items():
    _ENCODE[ord(_b)] = _c
    _ENCODE[ord(_b.lower())] = _c
_DECODE = np.array(list("$ACGTN"), dtype="<U1")


def encode_bases(seq: str) -> np.ndarray:
    if not seq:
        return np.zeros(0, dtype=np.int8)
    return _ENCODE[np.frombuffer(seq.encode("ascii", errors="replace"), dtype=np.uint8)]


def decode_bases(codes: np.ndarray) -> str:
    return "".join(_DECODE[np.clip(codes, 0, N_CODE)])


def reverse_complement_codes(codes: np.ndarray) -> np.ndarray:
    return _COMPLEMENT[codes[::-1]]


# ----------------------------------------------------------------------------
# k-mer table + GPU lookup (seeding.py:83-267, 720-798)
# ----------------------------------------------------------------------------

MAX_K = 31
_U = np.uint64


def hash64(values: np.ndarray) -> np.ndarray:
    """minimap2's invertible 64-bit integer mix (seeding.py:83-98), uint64."""
    x = np.asarray(values).astype(np.uint64)
    with np.errstate(over="ignore"):
        x = (x << _U(21)) - x - _U(1)  # ~x + (x << 21)
        x = x ^ (x >> _U(24))
        x = x * _U(265)  # x + (x << 3) + (x << 8)
        x = x ^ (x >> _U(14))
        x = x * _U(21)  # x + (x << 2) + (x << 4)
        x = x ^ (x >> _U(28))
        x = x * _U(1 + (1 << 31))  # x + (x << 31)
    return x


def pack_kmers(codes: np.ndarray, k: int):
    """(packed, valid) of every k-mer (seeding.py:101-126): first base most
    significant, A..T = 0..3; windows with a non-ACGT base are invalid and
    pack to 0."""
    if not 1 <= k <= MAX_K:
        raise ValueError(f"k must be in [1, {MAX_K}], got {k}")
    codes = np.asarray(codes)
    m = len(codes) - k + 1
    if m <= 0:
        return np.zeros(0, np.int64), np.zeros(0, bool)
    d = codes.astype(np.int64) - 1
    bad = d > 3  # N (and any code above it); A..T are 0..3
    # count of bad bases per window by prefix sums; value by Horner over k
    cb = np.concatenate([[0], np.cumsum(bad, dtype=np.int64)])
    valid = (cb[k:] - cb[:m]) == 0
    packed = np.zeros(m, np.int64)
    for o in range(k):
        packed = packed * 4 + d[o:o + m]
    packed[~valid] = 0
    return packed, valid


def _check_codes(codes: np.ndarray, what: str) -> np.ndarray:
    codes = np.asarray(codes)
    if codes.size and (codes.min() < 1 or codes.max() > 5):
        raise ValueError(f"{what}: base codes must be 1..4 (ACGT) or 5 (N)")
    return codes


def minimizer_mask(hashes: np.ndarray, window: int) -> np.ndarray:
    """Positions attaining the minimum of some window of `window` consecutive
    hashes (seeding.py:172-199; ties all kept)."""
    h = np.asarray(hashes)
    m = len(h)
    if m == 0:
        return np.zeros(0, bool)
    w = max(1, min(int(window), m))
    nw = m - w + 1
    sw = np.lib.stride_tricks.sliding_window_view(h, w)
    wmin = sw.min(axis=1)
    sel = np.zeros(m, bool)
    for o in range(w):
        sel[o:o + nw] |= h[o:o + nw] == wmin
    return sel


class KmerTable:
    """Sorted unique keys + CSR positions (seeding.py:202-267)."""

    def __init__(self, keys, offsets, positions, k, span=None):
        self.keys, self.offsets, self.positions = keys, offsets, positions
        self.k, self.span = k, span or k

    @classmethod
    def build(cls, packed, valid, k, span=None, max_occ=0) -> "KmerTable":
        pos = np.flatnonzero(valid).astype(np.int64)
        codes = packed[pos]
        order = np.argsort(codes, kind="stable")
        codes, pos = codes[order], pos[order]
        if len(codes):
            first = np.concatenate([[True], codes[1:] != codes[:-1]])
            starts = np.flatnonzero(first)
            keys = codes[starts]
            counts = np.diff(np.concatenate([starts, [len(codes)]]))
        else:
            keys, counts = np.zeros(0, np.int64), np.zeros(0, np.int64)
        if max_occ > 0 and len(keys) and counts.max() > max_occ:
            keep = counts <= max_occ
            pos = pos[np.repeat(keep, counts)]
            keys, counts = keys[keep], counts[keep]
        offsets = np.zeros(len(keys) + 1, np.int64)
        np.cumsum(counts, out=offsets[1:])
        return cls(keys.astype(np.int64), offsets, pos, k, span)

    def __len__(self) -> int:
        return len(self.keys)

    def lookup(self, queries: np.ndarray):
        queries = np.asarray(queries, np.int64)
        if len(self.keys) == 0 or len(queries) == 0:
            z = np.zeros(len(queries), np.int64)
            return z, z.copy()
        idx = np.searchsorted(self.keys, queries)
        cl = np.minimum(idx, len(self.keys) - 1)
        hit = self.keys[cl] == queries
        return (np.where(hit, self.offsets[cl], 0),
                np.where(hit, self.offsets[cl + 1] - self.offsets[cl], 0))


class GPUKmerIndex:
    """GPU-resident k-mer table with batched lookup (seeding.py:720-798).

    The table is built on the host exactly as the reference builds it
    (minimizers of the reference, sorted unique keys, CSR positions,
    ``max_occ``) and kept on the device with a directory over the top bits of
    the key, so a lookup searches one bucket.  ``query`` seeds one read,
    ``query_batch`` many reads in one pass of four kernels."""

    name = "gpu_kmer"

    def __init__(self, ref_codes: np.ndarray, k: int = 15, window: int = 10, max_occ: int = 0,
                 device=None, dir_bits: int = 22):
        import torch

        self.k, self.window = int(k), int(window)
        if not 1 <= self.window <= 128:
            raise ValueError("window must be in [1, 128]")
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        if self.device.type != "cuda":
            raise RuntimeError("GPUKmerIndex of graphmambaformer_amd needs a HIP device")
        packed, valid = pack_kmers(_check_codes(ref_codes, "reference"), self.k)
        selected = valid & minimizer_mask(hash64(packed), self.window)
        self._table = KmerTable.build(packed, selected, self.k, max_occ=max_occ)
        t = self._table
        self.dir_bits = int(min(dir_bits, 2 * self.k))
        shift = 2 * self.k - self.dir_bits
        bounds = np.arange((1 << self.dir_bits) + 1, dtype=np.int64) << shift
        dir_ = np.searchsorted(t.keys, bounds, side="left").astype(np.int32)
        dev = self.device
        self.keys = torch.from_numpy(t.keys).to(dev)
        self.offsets = torch.from_numpy(t.offsets).to(dev)
        self.positions = torch.from_numpy(t.positions.astype(np.int64)).to(dev)
        self.dir = torch.from_numpy(dir_).to(dev)

    def _index(self):
        from .._lib import KmerIndex, ptr

        ix = KmerIndex()
        ix.keys, ix.offsets, ix.positions = ptr(self.keys), ptr(self.offsets), ptr(self.positions)
        ix.dir = ptr(self.dir)
        ix.n_keys = int(self.keys.numel())
        ix.k, ix.dir_bits = self.k, self.dir_bits
        return ix

    def lookup(self, queries):
        """Batched (start, count) for a flat tensor of packed k-mers (int32
        each, as the reference returns them)."""
        import ctypes

        import torch

        from .._lib import check, lib, ptr, stream_of

        q = torch.as_tensor(queries).to(device=self.device, dtype=torch.int64).contiguous()
        n = q.numel()
        start = torch.zeros(n, dtype=torch.int64, device=self.device)
        count = torch.zeros(n, dtype=torch.int32, device=self.device)
        ix = self._index()
        check(lib().gmfa_kmer_lookup(ctypes.byref(ix), ptr(q), n, ptr(start), ptr(count),
                                     stream_of(self.device)), "gmfa_kmer_lookup")
        return start.to(torch.int32), count

    def query(self, read_codes: np.ndarray):
        """(read_pos, ref_pos, length) int64 of one read (seeding.py:777-798)."""
        return self.query_batch([read_codes])[0]

    def prepare(self, reads):
        """Pack and upload a batch of reads (codes 1..5) with its tiling; the
        returned batch can be seeded repeatedly by ``seed``."""
        import torch

        from .. import kernels as K
        from .._lib import KmerReads, ptr

        dev = self.device
        k = self.k
        reads = [_check_codes(r, "read") for r in reads]
        n = len(reads)
        lens = np.array([len(r) for r in reads], np.int64)
        m = np.maximum(lens - k + 1, 0)
        pos_off = np.zeros(n + 1, np.int64)
        np.cumsum(m, out=pos_off[1:])
        ntile = (m + 255) // 256
        tile_read = np.repeat(np.arange(n, dtype=np.int32), ntile)
        tile_start = (np.arange(int(ntile.sum()), dtype=np.int64) -
                      np.repeat(np.cumsum(ntile) - ntile, ntile)).astype(np.int32) * 256
        if n:
            w, msk, roff, has_n = K.pack_host(reads, alphabet="ref")
        else:
            w, msk, roff, has_n = (np.zeros(1, np.uint32), np.zeros(1, np.uint32),
                                   np.zeros(0, np.int64), False)

        def up(a):
            return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

        bufs = {"seq": up(w.view(np.int32)), "mask": up(msk.view(np.int32)) if has_n else None,
                "roff": up(roff), "len": up(lens.astype(np.int32)), "poff": up(pos_off),
                "tr": up(tile_read), "ts": up(tile_start)}
        rd = KmerReads()
        rd.seq, rd.mask, rd.n_words = ptr(bufs["seq"]), ptr(bufs["mask"]), int(bufs["seq"].numel())
        rd.read_off, rd.read_len, rd.pos_off = ptr(bufs["roff"]), ptr(bufs["len"]), ptr(bufs["poff"])
        rd.tile_read, rd.tile_start = ptr(bufs["tr"]), ptr(bufs["ts"])
        rd.n_reads, rd.n_tiles, rd.window = n, int(len(tile_read)), self.window
        npos = int(pos_off[-1])
        bufs["pstart"] = torch.empty(max(npos, 1), dtype=torch.int64, device=dev)
        bufs["pcount"] = torch.empty(max(npos, 1), dtype=torch.int32, device=dev)
        bufs["paoff"] = torch.empty(max(npos, 1), dtype=torch.int64, device=dev)
        bufs["rtot"] = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
        bufs["raoff"] = torch.empty(n + 1, dtype=torch.int64, device=dev)
        return rd, bufs

    def seed(self, prepared):
        """Seed a prepared batch: device (read_pos, ref_pos) of all anchors,
        read by read, and the host offsets [n_reads + 1] (one readback of the
        anchor counts sizes the output)."""
        import ctypes

        import torch

        from .._lib import check, lib, ptr, stream_of

        rd, b = prepared
        dev = self.device
        ix = self._index()
        s = stream_of(dev)
        check(lib().gmfa_kmer_seed(ctypes.byref(ix), ctypes.byref(rd), ptr(b["pstart"]),
                                   ptr(b["pcount"]), ptr(b["paoff"]), ptr(b["rtot"]),
                                   ptr(b["raoff"]), s), "gmfa_kmer_seed")
        offs = b["raoff"].cpu().numpy()
        total = int(offs[-1])
        read_pos = torch.empty(max(total, 1), dtype=torch.int64, device=dev)
        ref_pos = torch.empty(max(total, 1), dtype=torch.int64, device=dev)
        check(lib().gmfa_kmer_emit(ctypes.byref(ix), ctypes.byref(rd), ptr(b["pstart"]),
                                   ptr(b["pcount"]), ptr(b["paoff"]), ptr(b["raoff"]),
                                   ptr(read_pos), ptr(ref_pos), s), "gmfa_kmer_emit")
        return read_pos[:total], ref_pos[:total], offs

    def query_batch_device(self, reads):
        """Seed many reads at once (pack, upload, seed)."""
        return self.seed(self.prepare(reads))

    def query_batch(self, reads):
        """Per read (read_pos, ref_pos, length) int64 host arrays."""
        rp, fp, offs = self.query_batch_device(reads)
        rp, fp = rp.cpu().numpy(), fp.cpu().numpy()
        out = []
        for r in range(len(offs) - 1):
            a, b = int(offs[r]), int(offs[r + 1])
            out.append((rp[a:b].copy(), fp[a:b].copy(), np.full(b - a, self.k, np.int64)))
        return out

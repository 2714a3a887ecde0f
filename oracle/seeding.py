"""TEST INFRASTRUCTURE ONLY -- CPU restatement of the reference's k-mer
seeding (graphmambaformer/alignment/seeding.py of sarakh1999/GraphMambaFormer),
the checker for the GPU k-mer lookup.  Only tests/, __graft_entry__.smoke() and
bench.py's CPU baseline may use it.

Functions and the reference lines they restate:
  encode          seeding.py:64-69   A,C,G,T -> 1..4, anything else -> 5 (N)
  hash64          seeding.py:83-98   minimap2's invertible 64-bit integer hash
                                     (Thomas Wang's 64-bit mix, mod 2^64)
  pack_kmers      seeding.py:101-126 2-bit Horner packing, N windows -> 0 + invalid
  minimizer_mask  seeding.py:172-199 window minima of the hashes, all ties kept
  KmerTable       seeding.py:202-267 sorted unique keys + CSR positions, max_occ
  kmer_query      seeding.py:777-798 GPUKmerIndex.query: selected = valid &
                                     minimizer; lookup; expand hits in position
                                     order, table order within a position

Pins (tests/test_seeding.py): the reference's golden vectors
(tests/golden/seeding.json: hash64 values, packed/valid/selected masks, whole
GPUKmerIndex.query outputs), the reference's own tests restated
(test_alignment_stages.py:128-160 explicit-window minima and packing round
trip, 350-362 GPU index == minimizer index), brute force (dict index,
per-window minima), and hash64's invertibility (the inverse map restores
every input).
"""

from __future__ import annotations

import numpy as np

_M64 = (1 << 64) - 1
MAX_K = 31

_ENC = np.full(256, 5, dtype=np.int8)
for _c, _v in zip("ACGTacgt", (1, 2, 3, 4, 1, 2, 3, 4)):
    _ENC[ord(_c)] = _v


def encode(seq: str) -> np.ndarray:
    if not seq:
        return np.zeros(0, np.int8)
    return _ENC[np.frombuffer(seq.encode("ascii", errors="replace"), np.uint8)]


def hash64(values: np.ndarray) -> np.ndarray:
    """Wang/minimap2 64-bit mix on uint64 (wrapping arithmetic)."""
    x = np.asarray(values).astype(np.uint64, copy=True)
    u = np.uint64
    with np.errstate(over="ignore"):
        x = ~x + (x << u(21))
        x ^= x >> u(24)
        x = x + (x << u(3)) + (x << u(8))
        x ^= x >> u(14)
        x = x + (x << u(2)) + (x << u(4))
        x ^= x >> u(28)
        x = x + (x << u(31))
    return x


def hash64_inverse(h: int) -> int:
    """Inverse of hash64 on one value (each step is a bijection of Z/2^64)."""
    def inv_xorshift(y, s):
        x = y
        for _ in range(64 // s + 1):
            x = y ^ (x >> s)
        return x & _M64

    def inv_mul(y, m):  # y = x * m mod 2^64, m odd
        return (y * pow(m, -1, 1 << 64)) & _M64

    x = h & _M64
    x = inv_mul(x, 1 + (1 << 31))
    x = inv_xorshift(x, 28)
    x = inv_mul(x, 1 + 4 + 16)
    x = inv_xorshift(x, 14)
    x = inv_mul(x, 1 + 8 + 256)
    x = inv_xorshift(x, 24)
    # x = ~v + (v << 21) = (v << 21) - v - 1 = v * (2^21 - 1) - 1
    x = inv_mul((x + 1) & _M64, (1 << 21) - 1)
    return x


def pack_kmers(codes: np.ndarray, k: int):
    if not 1 <= k <= MAX_K:
        raise ValueError(f"k must be in [1, {MAX_K}]")
    codes = np.asarray(codes)
    m = len(codes) - k + 1
    if m <= 0:
        return np.zeros(0, np.int64), np.zeros(0, bool)
    digits = codes.astype(np.int64) - 1
    bad = digits > 3
    packed = np.zeros(m, np.int64)
    invalid = np.zeros(m, bool)
    for o in range(k):
        packed = packed * 4 + digits[o:o + m]
        invalid |= bad[o:o + m]
    packed[invalid] = 0
    return packed, ~invalid


def minimizer_mask(hashes: np.ndarray, window: int) -> np.ndarray:
    m = len(hashes)
    if m == 0:
        return np.zeros(0, bool)
    w = max(1, min(window, m))
    nw = m - w + 1
    wmin = hashes[:nw].copy()
    for o in range(1, w):
        np.minimum(wmin, hashes[o:o + nw], out=wmin)
    sel = np.zeros(m, bool)
    for o in range(w):
        sel[o:o + nw] |= hashes[o:o + nw] == wmin
    return sel


class KmerTable:
    def __init__(self, keys, offsets, positions, k):
        self.keys, self.offsets, self.positions, self.k = keys, offsets, positions, k

    @classmethod
    def build(cls, packed, valid, k, max_occ=0):
        pos = np.flatnonzero(valid).astype(np.int64)
        codes = packed[pos]
        order = np.argsort(codes, kind="stable")
        codes, pos = codes[order], pos[order]
        keys, counts = np.unique(codes, return_counts=True)
        offsets = np.zeros(len(keys) + 1, np.int64)
        np.cumsum(counts, out=offsets[1:])
        if max_occ > 0 and len(keys) and counts.max() > max_occ:
            keep = counts <= max_occ
            starts, ends = offsets[:-1][keep], offsets[1:][keep]
            pos = np.concatenate([pos[s:e] for s, e in zip(starts, ends)] or
                                 [np.zeros(0, np.int64)])
            keys, counts = keys[keep], counts[keep]
            offsets = np.zeros(len(keys) + 1, np.int64)
            np.cumsum(counts, out=offsets[1:])
        return cls(keys.astype(np.int64), offsets, pos, k)

    def lookup(self, queries):
        queries = np.asarray(queries, np.int64)
        if len(self.keys) == 0 or len(queries) == 0:
            z = np.zeros(len(queries), np.int64)
            return z, z.copy()
        idx = np.searchsorted(self.keys, queries)
        cl = np.minimum(idx, len(self.keys) - 1)
        hit = self.keys[cl] == queries
        start = np.where(hit, self.offsets[cl], 0)
        count = np.where(hit, self.offsets[cl + 1] - self.offsets[cl], 0)
        return start, count


def build_index(ref_codes, k=15, window=10, max_occ=0) -> KmerTable:
    """GPUKmerIndex.__init__ (seeding.py:731-749)."""
    packed, valid = pack_kmers(ref_codes, k)
    sel = valid & minimizer_mask(hash64(packed), window)
    return KmerTable.build(packed, sel, k, max_occ=max_occ)


def kmer_query(table: KmerTable, read_codes, window=10):
    """GPUKmerIndex.query (seeding.py:777-798): (read_pos, ref_pos, length)."""
    k = table.k
    packed, valid = pack_kmers(read_codes, k)
    sel = valid & minimizer_mask(hash64(packed), window)
    pos = np.flatnonzero(sel)
    z = np.zeros(0, np.int64)
    if pos.size == 0:
        return z, z.copy(), z.copy()
    start, count = table.lookup(packed[pos])
    total = int(count.sum())
    if total == 0:
        return z, z.copy(), z.copy()
    read_pos = np.repeat(pos, count)
    base = np.repeat(np.concatenate([[0], np.cumsum(count)[:-1]]), count)
    within = np.arange(total, dtype=np.int64) - base
    ref_pos = table.positions[np.repeat(start, count) + within]
    return read_pos.astype(np.int64), ref_pos.astype(np.int64), np.full(total, k, np.int64)

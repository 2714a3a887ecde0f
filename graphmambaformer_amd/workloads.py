"""Seeded synthetic workloads for the SW and chaining path.

All generators use numpy ``default_rng(seed)`` (PCG64) and emit base codes
0..3 (A, C, G, T).  Shapes follow the benchmark configurations:

  sw_short   1,024 pairs, 150 bp reads, target = read with 1% substitutions and
             0.2% 1-3 bp indel events plus 32 random bp on each side; band
             half-width 32 centred at offset 32; scoring 2/-4/-4/-2.
  sw_long    16,384 pairs, length lognormal(median 12 kb, sigma 0.4) clipped to
             5-25 kb, 0.5% error events (70% 1-2 bp indels, 30% substitutions),
             band half-width 256.
  sw_noisy   8,192 pairs, length lognormal(median 8 kb, sigma 1.0) clipped to
             500 bp-100 kb, 6% errors (60% indels weighted x2 inside homopolymer
             runs >= 3 bases, 40% substitutions), band half-width 751.
  chains     4,096 reads, anchors per read ~ Poisson(400) on a 10 kb read, 90%
             on one collinear diagonal with Gaussian jitter sigma 3, 10% uniform
             decoys over a 46 Mbp reference; span 15, weight 15.

The homopolymer-biased indel model mirrors data/synthetic.py:521-528 of the
reference (indels lengthen or shorten runs).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

MATCH, MISMATCH, GAP_OPEN, GAP_EXTEND = 2, 4, 4, 2


def homopolymer_run_mask(seq: np.ndarray, min_run: int = 3) -> np.ndarray:
    """True at positions inside runs of >= min_run identical bases."""
    n = len(seq)
    if n == 0:
        return np.zeros(0, bool)
    change = np.flatnonzero(np.diff(seq) != 0) + 1
    starts = np.concatenate([[0], change])
    ends = np.concatenate([change, [n]])
    lens = ends - starts
    return np.repeat(lens >= min_run, lens)


def mutate(rng, seq: np.ndarray, rate: float, indel_frac: float, indel_max: int,
           homopolymer_weight: float = 1.0, sub_rate: float | None = None,
           indel_rate: float | None = None) -> np.ndarray:
    """Apply substitution / insertion / deletion events to a 0..3 code array.

    Either ``rate`` (events per base) split by ``indel_frac``, or explicit
    ``sub_rate`` / ``indel_rate``.  Indel lengths are uniform 1..indel_max,
    insertion/deletion 50/50; inside homopolymer runs (weight > 1) indel
    probability is scaled and inserted bases repeat the run base."""
    n = len(seq)
    if sub_rate is None:
        sub_rate = rate * (1.0 - indel_frac)
        indel_rate = rate * indel_frac
    w = np.ones(n)
    hp = homopolymer_run_mask(seq) if homopolymer_weight != 1.0 else np.zeros(n, bool)
    if homopolymer_weight != 1.0:
        w = np.where(hp, homopolymer_weight, 1.0)
        w = w / w.mean() if n else w
    s = seq.copy()
    sub = rng.random(n) < sub_rate
    s[sub] = (s[sub] + rng.integers(1, 4, size=int(sub.sum()))) % 4
    ev = np.flatnonzero(rng.random(n) < indel_rate * w)
    if len(ev) == 0:
        return s.astype(np.uint8)
    lens = rng.integers(1, indel_max + 1, size=len(ev))
    is_ins = rng.random(len(ev)) < 0.5
    counts = np.ones(n, dtype=np.int64)
    for p, L in zip(ev[~is_ins], lens[~is_ins]):
        counts[p:p + L] = 0
    ins_p, ins_l = ev[is_ins], lens[is_ins]
    np.add.at(counts, ins_p, ins_l)
    keep = counts[ins_p] > ins_l  # insertion anchored on a deleted base is dropped
    out = np.repeat(s, counts)
    ends = np.cumsum(counts)
    rand_ins = ~hp[ins_p] & keep
    for p, L in zip(ins_p[rand_ins], ins_l[rand_ins]):
        e = ends[p]
        out[e - L:e] = rng.integers(0, 4, size=L)
    return out.astype(np.uint8)


@dataclass
class SwWorkload:
    name: str
    queries: list
    targets: list
    band_offset: np.ndarray
    half_band: int
    match: int = MATCH
    mismatch: int = MISMATCH
    gap_open: int = GAP_OPEN
    gap_extend: int = GAP_EXTEND
    meta: dict = field(default_factory=dict)

    @property
    def n_pairs(self) -> int:
        return len(self.queries)

    def cells(self) -> np.ndarray:
        """In-band cells per pair: sum_i #{d : 1 <= i + off - hb + d <= tlen}."""
        return band_cells(np.array([len(q) for q in self.queries]),
                          np.array([len(t) for t in self.targets]), self.band_offset,
                          self.half_band)

    def as_ref_codes(self, k: int):
        """Pair k in the reference alphabet (A,C,G,T = 1..4)."""
        return (self.queries[k].astype(np.int16) + 1, self.targets[k].astype(np.int16) + 1)


def band_cells(qlen, tlen, off, hb) -> np.ndarray:
    qlen = np.asarray(qlen, np.int64)
    tlen = np.asarray(tlen, np.int64)
    off = np.asarray(off, np.int64)
    W = 2 * hb + 1
    out = np.zeros(len(qlen), np.int64)
    for p in range(len(qlen)):
        i = np.arange(1, qlen[p] + 1, dtype=np.int64)
        lo = np.maximum(0, hb + 1 - off[p] - i)
        hi = np.minimum(W - 1, tlen[p] + hb - off[p] - i)
        out[p] = np.maximum(hi - lo + 1, 0).sum()
    return out


def sw_short(n_pairs: int = 1024, seed: int = 0) -> SwWorkload:
    rng = np.random.default_rng(seed)
    qs, ts = [], []
    for _ in range(n_pairs):
        read = rng.integers(0, 4, size=150).astype(np.uint8)
        t = mutate(rng, read, 0.0, 0.0, 3, sub_rate=0.01, indel_rate=0.002)
        t = np.concatenate([rng.integers(0, 4, size=32), t, rng.integers(0, 4, size=32)])
        qs.append(read)
        ts.append(t.astype(np.uint8))
    return SwWorkload("sw_short", qs, ts, np.full(n_pairs, 32, np.int64), 32)


def _lognormal_lengths(rng, n, median, sigma, lo, hi):
    return np.clip(np.round(rng.lognormal(np.log(median), sigma, size=n)), lo, hi).astype(np.int64)


def sw_long(n_pairs: int = 16384, seed: int = 2) -> SwWorkload:
    rng = np.random.default_rng(seed)
    lens = _lognormal_lengths(rng, n_pairs, 12000, 0.4, 5000, 25000)
    qs, ts = [], []
    for L in lens:
        tmpl = rng.integers(0, 4, size=int(L)).astype(np.uint8)
        qs.append(mutate(rng, tmpl, 0.005, 0.7, 2))
        ts.append(tmpl)
    return SwWorkload("sw_long", qs, ts, np.zeros(n_pairs, np.int64), 256)


def sw_noisy(n_pairs: int = 8192, seed: int = 3) -> SwWorkload:
    rng = np.random.default_rng(seed)
    lens = _lognormal_lengths(rng, n_pairs, 8000, 1.0, 500, 100000)
    qs, ts = [], []
    for L in lens:
        tmpl = rng.integers(0, 4, size=int(L)).astype(np.uint8)
        qs.append(mutate(rng, tmpl, 0.06, 0.6, 3, homopolymer_weight=2.0))
        ts.append(tmpl)
    return SwWorkload("sw_noisy", qs, ts, np.zeros(n_pairs, np.int64), 751)


def subsample(w: SwWorkload, idx) -> SwWorkload:
    idx = np.asarray(idx, np.int64)
    return SwWorkload(w.name + "_sub", [w.queries[i] for i in idx], [w.targets[i] for i in idx],
                      w.band_offset[idx], w.half_band, w.match, w.mismatch, w.gap_open,
                      w.gap_extend, dict(w.meta, subsample=idx.tolist()))


@dataclass
class ChainWorkload:
    read_end: np.ndarray  # int64, all reads back to back, DP order per read
    ref_end: np.ndarray
    weight: np.ndarray    # int32
    seg_off: np.ndarray   # int64 [n_reads + 1]
    max_lookback: int = 50
    max_gap: int = 5000
    gap_open: int = 6
    gap_extend: int = 1

    @property
    def n_reads(self) -> int:
        return len(self.seg_off) - 1

    def pair_evaluations(self) -> int:
        n = np.diff(self.seg_off)
        L = self.max_lookback
        tot = 0
        for k in n:
            i = np.arange(k)
            tot += int(np.minimum(i, L).sum())
        return tot


def chains(n_reads: int = 4096, seed: int = 1, read_len: int = 10000,
           ref_len: int = 46_000_000, span: int = 15) -> ChainWorkload:
    rng = np.random.default_rng(seed)
    res, fes, offs = [], [], [0]
    for _ in range(n_reads):
        n = int(rng.poisson(400))
        n_col = int(round(0.9 * n))
        rp_c = np.sort(rng.integers(0, read_len - span + 1, size=n_col))
        start = int(rng.integers(0, ref_len - 2 * read_len))
        fp_c = start + rp_c + np.round(rng.normal(0.0, 3.0, size=n_col)).astype(np.int64)
        rp_d = rng.integers(0, read_len - span + 1, size=n - n_col)
        fp_d = rng.integers(0, ref_len, size=n - n_col)
        rp = np.concatenate([rp_c, rp_d]).astype(np.int64)
        fp = np.concatenate([fp_c, fp_d]).astype(np.int64)
        re_, fe = rp + span, fp + span
        order = np.lexsort((re_, fe))  # AnchorSet.sorted_by_ref (types.py:149-152)
        res.append(re_[order])
        fes.append(fe[order])
        offs.append(offs[-1] + n)
    read_end = np.concatenate(res) if res else np.zeros(0, np.int64)
    ref_end = np.concatenate(fes) if fes else np.zeros(0, np.int64)
    return ChainWorkload(read_end=read_end, ref_end=ref_end,
                         weight=np.full(len(read_end), span, np.int32),
                         seg_off=np.array(offs, np.int64))


# ----------------------------------------------------------------------------
# serving: requests against a resident reference
# ----------------------------------------------------------------------------


@dataclass
class ServePool:
    """Distinct requests of one band class.  Request k aligns query k against
    the reference window [t_start[k], t_start[k] + t_len[k])."""

    name: str
    half_band: int
    queries: list
    t_start: np.ndarray  # int64, base offset in the reference
    t_len: np.ndarray    # int32
    band_offset: np.ndarray  # int32

    @property
    def n(self) -> int:
        return len(self.queries)

    def cells(self) -> np.ndarray:
        return band_cells([len(q) for q in self.queries], self.t_len, self.band_offset,
                          self.half_band)

    def pair(self, k: int, reference: np.ndarray):
        """(query, target) of request k as 0..3 codes."""
        s = int(self.t_start[k])
        return self.queries[k], reference[s:s + int(self.t_len[k])]


@dataclass
class ServingMix:
    reference: np.ndarray  # uint8 codes 0..3
    pools: list            # ServePool per class
    mix: tuple             # request share of each class
    match: int = MATCH
    mismatch: int = MISMATCH
    gap_open: int = GAP_OPEN
    gap_extend: int = GAP_EXTEND


def serving_mix(seed: int = 4, ref_len: int = 46_000_000,
                pool_sizes: tuple = (2048, 1024, 512)) -> ServingMix:
    """A random resident reference and request pools of three band classes:
    short (150 bp reads against 214 bp windows, half_band 32 at offset 32, 1%
    substitutions + 0.2% indels), long (lognormal(12 kb, 0.4) clipped 5-25 kb,
    0.5% errors, half_band 256) and noisy (lognormal(8 kb, 1.0) clipped
    0.5-100 kb, 6% homopolymer-weighted errors, half_band 751).  Windows start
    at uniform positions; the query is the mutated window."""
    rng = np.random.default_rng(seed)
    ref = rng.integers(0, 4, size=ref_len, dtype=np.uint8)
    pools = []
    # short reads
    n = pool_sizes[0]
    pos = rng.integers(0, ref_len - 214, size=n)
    qs = [mutate(rng, ref[p + 32:p + 182], 0.0, 0.0, 3, sub_rate=0.01, indel_rate=0.002)
          for p in pos]
    pools.append(ServePool("short", 32, qs, pos.astype(np.int64), np.full(n, 214, np.int32),
                           np.full(n, 32, np.int32)))
    # long reads
    n = pool_sizes[1]
    lens = _lognormal_lengths(rng, n, 12000, 0.4, 5000, 25000)
    pos = np.array([rng.integers(0, ref_len - L) for L in lens], np.int64)
    qs = [mutate(rng, ref[p:p + L], 0.005, 0.7, 2) for p, L in zip(pos, lens)]
    pools.append(ServePool("long", 256, qs, pos, lens.astype(np.int32), np.zeros(n, np.int32)))
    # noisy reads
    n = pool_sizes[2]
    lens = _lognormal_lengths(rng, n, 8000, 1.0, 500, 100000)
    pos = np.array([rng.integers(0, ref_len - L) for L in lens], np.int64)
    qs = [mutate(rng, ref[p:p + L], 0.06, 0.6, 3, homopolymer_weight=2.0)
          for p, L in zip(pos, lens)]
    pools.append(ServePool("noisy", 751, qs, pos, lens.astype(np.int32), np.zeros(n, np.int32)))
    return ServingMix(ref, pools, (0.6, 0.3, 0.1))


def serving_requests(mix: ServingMix, n_requests: int, seed: int = 5):
    """Class and pool index of each request of a stream (class drawn with the
    mix shares, pool entry uniformly)."""
    rng = np.random.default_rng(seed)
    cls = rng.choice(len(mix.pools), size=n_requests, p=np.asarray(mix.mix) / sum(mix.mix))
    idx = np.array([rng.integers(0, mix.pools[c].n) for c in cls], np.int64)
    return cls.astype(np.int32), idx


def poisson_arrivals(n: int, rate: float, seed: int = 6) -> np.ndarray:
    """Arrival times (s) of a Poisson stream of n requests at `rate` per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def seeding_reads(seed: int = 7, ref_len: int = 46_000_000, n_reads: int = 4096,
                  read_len: int = 10_000):
    """A random reference (codes 1..4) and reads sampled from it with 1% errors
    (half substitutions, half 1-3 bp indels); for k-mer seeding."""
    rng = np.random.default_rng(seed)
    ref = rng.integers(0, 4, size=ref_len, dtype=np.uint8)
    reads = []
    for _ in range(n_reads):
        p = int(rng.integers(0, ref_len - read_len))
        reads.append(mutate(rng, ref[p:p + read_len], 0.01, 0.5, 3).astype(np.int8) + 1)
    return (ref + 1).astype(np.int8), reads

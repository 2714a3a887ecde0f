"""ctypes wrapper over ``liboracle.so`` (built from ``gmfa_oracle.c``).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Pins (tests/test_oracle.py):
  sw_banded / sw_traceback: KAT-SW1, identical sequences, unbanded Gotoh,
      float32 transcription (ref_semantics), CIGAR rescoring, golden vectors
      written from the reference (tests/golden/).
  chain_dp_f64 / chain_dp_i32: KAT-CH1, KAT-CH2, O(n^2) brute force, golden
      vectors.
  chain_backtrack / select: golden vectors, hand cases from
      tests/test_alignment_stages.py:586-617 of the reference.
"""

from __future__ import annotations

import ctypes
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

OPS = "=XID"


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    import subprocess

    subprocess.run(["make", "-C", _HERE, "-s"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        i32 = ctypes.c_int32
        L.oracle_sw_banded.argtypes = [i64, i64, i64, P, P, P, P, P, i32, i32, i32, i32,
                                       i32, i64, P, P, P, P, P, P, P]
        L.oracle_sw_banded.restype = ctypes.c_int
        L.oracle_sw_traceback.argtypes = [P, P, P, i64, i32, i64, P, P, i64, i64, i32, i32,
                                          i32, i32, P, P, P]
        L.oracle_sw_traceback.restype = i64
        L.oracle_chain_dp_f64.argtypes = [i64, P, P, P, i32, i64, P, P, P, P]
        L.oracle_chain_dp_f64.restype = None
        L.oracle_chain_dp_i32.argtypes = [i64, P, P, P, i32, i64, i32, i32, P, P]
        L.oracle_chain_dp_i32.restype = None
        L.oracle_chain_backtrack.argtypes = [i64, P, P, i64, ctypes.c_double, P, P, P]
        L.oracle_chain_backtrack.restype = i64
        L.oracle_sw_align_batch.argtypes = [i64, P, P, P, P, P, P, P, i32, i32, i32, i32, i32,
                                            i32, P, P, P, P, P, P]
        L.oracle_sw_align_batch.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def xdrop_threshold(x_drop) -> int:
    """Integer form of the reference's `> x_drop` test (-1 = disabled)."""
    return int(np.floor(x_drop)) if x_drop > 0 else -1


def sw_banded(query, target, qlen, tlen, off, hb, ma, mi, go, ge, x_drop=0.0, store=False):
    """Batched banded SW (extension.py:78-204).  Returns a dict of numpy arrays."""
    q = np.ascontiguousarray(query, dtype=np.int16)
    t = np.ascontiguousarray(target, dtype=np.int16)
    if q.ndim == 1:
        q = q[None]
    if t.ndim == 1:
        t = t[None]
    B, M = q.shape
    N = t.shape[1]
    qlen = np.ascontiguousarray(qlen, dtype=np.int64).reshape(B)
    tlen = np.ascontiguousarray(tlen, dtype=np.int64).reshape(B)
    off = np.ascontiguousarray(off, dtype=np.int64).reshape(B)
    W = 2 * hb + 1
    score = np.zeros(B, np.int32)
    qend = np.zeros(B, np.int64)
    tend = np.zeros(B, np.int64)
    rows = np.zeros(1, np.int64)
    if store:
        H = np.zeros((B, M + 1, W), np.int32)
        E = np.zeros((B, M + 1, W), np.int32)
        F = np.zeros((B, M + 1, W), np.int32)
        hp, ep, fp = _p(H), _p(E), _p(F)
    else:
        H = E = F = None
        hp = ep = fp = None
    rc = lib().oracle_sw_banded(B, M, N, _p(q), _p(t), _p(qlen), _p(tlen), _p(off), hb, ma,
                                mi, go, ge, xdrop_threshold(x_drop), _p(score), _p(qend), _p(tend), hp, ep,
                                fp, _p(rows))
    if rc != 0:
        raise MemoryError("oracle_sw_banded allocation failed")
    out = dict(score=score, query_end=qend, target_end=tend, rows_done=int(rows[0]),
               half_band=hb, band_offset=off)
    if store:
        out.update(H=H, E=E, F=F)
    return out


def rle(ops):
    out = []
    for op in ops:
        if out and out[-1][0] == op:
            out[-1] = (op, out[-1][1] + 1)
        else:
            out.append((op, 1))
    return out


def sw_traceback(res, index, q, t, ma, mi, go, ge):
    """traceback_banded (extension.py:207-302) for pair ``index`` of ``res``."""
    H = np.ascontiguousarray(res["H"][index])
    E = np.ascontiguousarray(res["E"][index])
    F = np.ascontiguousarray(res["F"][index])
    M = H.shape[0] - 1
    q = np.ascontiguousarray(q, dtype=np.int16)
    t = np.ascontiguousarray(t, dtype=np.int16)
    qend = int(res["query_end"][index])
    tend = int(res["target_end"][index])
    ops = np.zeros(qend + tend + 1, np.uint8)
    rs = np.zeros(1, np.int64)
    ts = np.zeros(1, np.int64)
    n = lib().oracle_sw_traceback(_p(H), _p(E), _p(F), M, res["half_band"],
                                  int(res["band_offset"][index]), _p(q), _p(t), qend, tend,
                                  ma, mi, go, ge, _p(ops), _p(rs), _p(ts))
    oplist = [OPS[o] for o in ops[:n]]
    cigar = rle(oplist)
    counts = {k: 0 for k in OPS}
    for op, c in cigar:
        counts[op] += c
    return dict(score=int(res["score"][index]), cigar=cigar, read_start=int(rs[0]),
                read_end=qend, ref_start=int(ts[0]), ref_end=tend, n_match=counts["="],
                n_mismatch=counts["X"], n_insertion=counts["I"], n_deletion=counts["D"])


def sw_align_pairs(queries, targets, offs, hb, ma, mi, go, ge, threads=1):
    """Independent pairs (B = 1 groups), forward + traceback each.

    This is the reference CPU execution model: one pair per task fanned out
    over host threads (pipeline.py:355-363, accel/parallel.py:60-103).
    """

    def one(k):
        q = np.asarray(queries[k], np.int16)
        t = np.asarray(targets[k], np.int16)
        r = sw_banded(q[None], t[None], [len(q)], [len(t)], [offs[k]], hb, ma, mi, go, ge,
                      store=True)
        return sw_traceback(r, 0, q, t, ma, mi, go, ge)

    if threads <= 1:
        return [one(k) for k in range(len(queries))]
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(one, range(len(queries))))


def sw_align_batch(queries, targets, offs, hb, ma, mi, go, ge, threads=1):
    """sw_align_pairs in one C call (oracle_sw_align_batch): independent pairs,
    forward + traceback on ``threads`` POSIX threads, no per-pair Python work.
    Returns a dict of arrays: score, read_end, ref_end, read_start, ref_start,
    n_cols (CIGAR columns)."""
    n = len(queries)
    ql = np.array([len(x) for x in queries], np.int64)
    tl = np.array([len(x) for x in targets], np.int64)
    qo = np.zeros(n, np.int64)
    to = np.zeros(n, np.int64)
    if n:
        qo[1:] = np.cumsum(ql)[:-1]
        to[1:] = np.cumsum(tl)[:-1]
    qc = np.ascontiguousarray(np.concatenate([np.asarray(x, np.int16) for x in queries])
                              if n else np.zeros(0, np.int16))
    tc = np.ascontiguousarray(np.concatenate([np.asarray(x, np.int16) for x in targets])
                              if n else np.zeros(0, np.int16))
    bo = np.ascontiguousarray(offs, dtype=np.int64).reshape(n)
    out = {k: np.zeros(n, np.int64) for k in ("read_end", "ref_end", "read_start", "ref_start",
                                               "n_cols")}
    score = np.zeros(n, np.int32)
    rc = lib().oracle_sw_align_batch(n, _p(qc), _p(qo), _p(ql), _p(tc), _p(to), _p(tl), _p(bo),
                                     hb, ma, mi, go, ge, int(threads), _p(score),
                                     _p(out["read_end"]), _p(out["ref_end"]),
                                     _p(out["read_start"]), _p(out["ref_start"]),
                                     _p(out["n_cols"]))
    if rc != 0:
        raise MemoryError("oracle_sw_align_batch allocation failed")
    out["score"] = score
    return out


def sw_score_pairs(queries, targets, offs, hb, ma, mi, go, ge, threads=1):
    """Forward only, one pair per task (score / ends)."""

    def one(k):
        q = np.asarray(queries[k], np.int16)
        t = np.asarray(targets[k], np.int16)
        r = sw_banded(q[None], t[None], [len(q)], [len(t)], [offs[k]], hb, ma, mi, go, ge)
        return int(r["score"][0]), int(r["query_end"][0]), int(r["target_end"][0])

    if threads <= 1:
        return [one(k) for k in range(len(queries))]
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(one, range(len(queries))))


def penalty_lut(gap_open, gap_extend, log_coeff, max_gap):
    """_penalty(g) for g in [0, max_gap), with the reference's numpy expression
    (chaining.py:120-126) evaluated on a contiguous int64 array."""
    g = np.arange(max(int(max_gap), 1), dtype=np.int64)
    return np.where(g > 0, gap_open + gap_extend * g + log_coeff * np.log2(g + 1.0), 0.0)


def chain_dp_f64(read_end, ref_end, weight, max_lookback, max_gap, pen_lut, bonus=None):
    re = np.ascontiguousarray(read_end, dtype=np.int64)
    fe = np.ascontiguousarray(ref_end, dtype=np.int64)
    w = np.ascontiguousarray(weight, dtype=np.float64)
    lut = np.ascontiguousarray(pen_lut, dtype=np.float64)
    n = len(re)
    f = np.zeros(n, np.float64)
    par = np.full(n, -1, np.int64)
    bp = None
    if bonus is not None:
        bonus = np.ascontiguousarray(bonus, dtype=np.float64)
        bp = _p(bonus)
    lib().oracle_chain_dp_f64(n, _p(re), _p(fe), _p(w), int(max_lookback), int(max_gap),
                              _p(lut), bp, _p(f), _p(par))
    return f, par


def chain_dp_i32(read_end, ref_end, weight, max_lookback, max_gap, gap_open, gap_extend):
    re = np.ascontiguousarray(read_end, dtype=np.int64)
    fe = np.ascontiguousarray(ref_end, dtype=np.int64)
    w = np.ascontiguousarray(weight, dtype=np.int32)
    n = len(re)
    f = np.zeros(n, np.int32)
    par = np.full(n, -1, np.int64)
    lib().oracle_chain_dp_i32(n, _p(re), _p(fe), _p(w), int(max_lookback), int(max_gap),
                              int(gap_open), int(gap_extend), _p(f), _p(par))
    return f, par


def chain_backtrack(f, parent, min_anchors, min_score):
    """AffineChainer._backtrack (chaining.py:530-576) over local DP indices.

    Returns a list of (members in read order, score)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    par = np.ascontiguousarray(parent, dtype=np.int64)
    n = len(f)
    members = np.zeros(max(n, 1), np.int64)
    starts = np.zeros(n + 1, np.int64)
    scores = np.zeros(max(n, 1), np.float64)
    nc = lib().oracle_chain_backtrack(n, _p(f), _p(par), int(min_anchors), float(min_score),
                                      _p(members), _p(starts), _p(scores))
    return [(members[starts[c]:starts[c + 1]].copy(), float(scores[c])) for c in range(nc)]


def select(chains, secondary_score_ratio, secondary_overlap, max_chains):
    """AffineChainer.select (chaining.py:579-617) over dicts with keys
    score, read_start, read_end.  Returns (kept indices, is_primary flags)."""
    if not chains:
        return [], []
    order = sorted(range(len(chains)), key=lambda k: -chains[k]["score"])
    best = chains[order[0]]["score"]
    cutoff = best * secondary_score_ratio
    kept, prim = [], []
    for k in order:
        c = chains[k]
        if c["score"] < cutoff and kept:
            break
        ov = 0.0
        for kk, p in zip(kept, prim):
            if not p:
                continue
            o = chains[kk]
            lo = max(c["read_start"], o["read_start"])
            hi = min(c["read_end"], o["read_end"])
            if hi > lo:
                span = max(min(c["read_end"] - c["read_start"],
                               o["read_end"] - o["read_start"]), 1)
                ov = max(ov, (hi - lo) / span)
        if ov > secondary_overlap:
            continue
        prim.append(not kept)
        kept.append(k)
        if len(kept) >= max_chains:
            break
    return kept, prim


# ----------------------------------------------------------------------------
# Stage drivers restated on top of the kernels above
# ----------------------------------------------------------------------------
_ENC = np.full(256, 5, np.int16)
for _c, _v in zip("ACGT", (1, 2, 3, 4)):
    _ENC[ord(_c)] = _v
    _ENC[ord(_c.lower())] = _v
_COMP = np.array([0, 4, 3, 2, 1, 5], np.int16)


def encode(seq: str) -> np.ndarray:
    """seeding.py:64-70: A,C,G,T -> 1..4, anything else -> 5."""
    return _ENC[np.frombuffer(seq.encode("ascii", errors="replace"), np.uint8)]


def revcomp(codes: np.ndarray) -> np.ndarray:
    """seeding.py:79-80."""
    return _COMP[codes[::-1]]


def chain_read(read_pos, ref_pos, length, strand, max_lookback=64, max_gap=5000,
               gap_open=6.0, gap_extend=0.05, log_coeff=0.5, min_chain_score=20.0,
               min_chain_anchors=1, secondary_score_ratio=0.6, secondary_overlap=0.5,
               max_chains=8):
    """AffineChainer.chain (chaining.py:620-642) without neural/graph context:
    per strand (+1 then -1) stable lexsort by (ref_end, read_end), weight =
    anchor length, DP, backtrack; then one select over both strands."""
    read_pos = np.asarray(read_pos, np.int64)
    ref_pos = np.asarray(ref_pos, np.int64)
    length = np.asarray(length, np.int64)
    strand = np.asarray(strand, np.int64)
    if len(read_pos) == 0:
        return []
    lut = penalty_lut(gap_open, gap_extend, log_coeff, max_gap)
    chains = []
    for s in (1, -1):
        idx = np.flatnonzero(strand == s)
        if len(idx) < min_chain_anchors:
            continue
        rp, fp, ln = read_pos[idx], ref_pos[idx], length[idx]
        order = np.lexsort((rp + ln, fp + ln))
        rp, fp, ln = rp[order], fp[order], ln[order]
        gidx = idx[order]
        f, par = chain_dp_f64(rp + ln, fp + ln, ln.astype(np.float64), max_lookback, max_gap,
                              lut)
        for members, score in chain_backtrack(f, par, min_chain_anchors, min_chain_score):
            chains.append(dict(anchor_idx=gidx[members], score=score, strand=s,
                               read_start=int(rp[members].min()),
                               read_end=int((rp + ln)[members].max()),
                               ref_start=int(fp[members].min()),
                               ref_end=int((fp + ln)[members].max())))
    kept, prim = select(chains, secondary_score_ratio, secondary_overlap, max_chains)
    out = []
    for k, p in zip(kept, prim):
        c = dict(chains[k])
        c["is_primary"] = p
        out.append(c)
    return out


def extend_chains(read_seq, chains, anchor_diagonals, ref_seq, half_band=64,
                  max_half_band=512, flank=100, max_window=32768, ma=2, mi=4, go=6, ge=2,
                  x_drop=600.0):
    """ExtensionEngine.extend_chains (extension.py:489-641), banded_sw path.

    ``chains``: dicts with strand, read_start, ref_start, anchor_idx;
    ``anchor_diagonals``: ref_pos - read_pos per anchor of the read."""
    if not chains:
        return []
    fwd = encode(read_seq)
    rev = revcomp(fwd)
    ref = encode(ref_seq)
    read_len, ref_len = len(fwd), len(ref)
    wins = []
    for c in chains:
        diag = c["ref_start"] - c["read_start"]
        start = max(0, diag - flank)
        end = min(ref_len, diag + read_len + flank)
        if end - start > max_window:
            end = start + max_window
        idx = np.asarray(c["anchor_idx"], np.int64)
        spread = 0
        if len(idx):
            dg = np.asarray(anchor_diagonals, np.int64)[idx]
            spread = int(dg.max() - dg.min())
        wins.append((start, end, min(half_band + spread, max_half_band)))
    hb = max(w[2] for w in wins)
    width = max(w[1] - w[0] for w in wins)
    B = len(chains)
    q = np.full((B, read_len), -1, np.int16)
    t = np.full((B, width), -1, np.int16)
    tlen = np.zeros(B, np.int64)
    off = np.zeros(B, np.int64)
    for b, (c, (start, end, _)) in enumerate(zip(chains, wins)):
        codes = fwd if c["strand"] > 0 else rev
        q[b, :len(codes)] = codes
        w = ref[start:end]
        t[b, :len(w)] = w
        tlen[b] = len(w)
        off[b] = (c["ref_start"] - c["read_start"]) - start
    res = sw_banded(q, t, np.full(B, read_len), tlen, off, hb, ma, mi, go, ge,
                    x_drop=x_drop, store=True)
    out = []
    for b, (c, (start, _e, _)) in enumerate(zip(chains, wins)):
        codes = fwd if c["strand"] > 0 else rev
        e = sw_traceback(res, b, codes, ref[start:start + int(tlen[b])], ma, mi, go, ge)
        e["ref_start"] += start
        e["ref_end"] += start
        cig = [("S", e["read_start"])] + e["cigar"] + [("S", read_len - e["read_end"])]
        merged = []
        for op, n in cig:
            if n <= 0:
                continue
            if merged and merged[-1][0] == op:
                merged[-1] = (op, merged[-1][1] + n)
            else:
                merged.append((op, n))
        e["cigar"] = merged
        out.append(e)
    return out


def xdrop_from_rowkeys(keys: np.ndarray, offs: np.ndarray, qlen_host: np.ndarray, x_drop: float,
                       hb: int, boff_host: np.ndarray, group_off=None):
    """Batch-wide x-drop (extension.py:186-193) from the forward kernel's
    per-row keys (key = H << 11 | 2047 - d, qlen keys per pair at offs[p]).

    The forward pass computed every row of every pair; the reference stops a
    batch after the first row i where every pair has
    row_max_seen - row_best > x_drop, so the best cell is the first strict
    maximum over rows <= i.  ``group_off``: batches as pair ranges (default:
    all pairs are one batch).  Returns host arrays (score, qend, tend)."""
    n = len(qlen_host)
    go = np.array([0, n], np.int64) if group_off is None else np.asarray(group_off, np.int64)
    score = np.zeros(n, dtype=np.int32)
    qe = np.zeros(n, dtype=np.int32)
    te = np.zeros(n, dtype=np.int32)
    for g in range(len(go) - 1):
        ps = range(int(go[g]), int(go[g + 1]))
        maxq = max((int(qlen_host[p]) for p in ps), default=0)
        if maxq == 0:
            continue
        rb = np.zeros((len(ps), maxq), dtype=np.int64)
        rd = np.zeros((len(ps), maxq), dtype=np.int64)
        for x, p in enumerate(ps):
            q = int(qlen_host[p])
            k = keys[offs[p]:offs[p] + q].astype(np.int64)
            rb[x, :q] = k >> 11
            rd[x, :q] = 2047 - (k & 2047)
        stop = maxq
        if x_drop > 0:
            thr = int(np.floor(x_drop))
            seen = np.maximum.accumulate(rb, axis=1)
            hit = np.flatnonzero(((seen - rb) > thr).all(axis=0))
            if len(hit):
                stop = int(hit[0]) + 1
        for x, p in enumerate(ps):
            r = rb[x, :stop]
            if r.max() <= 0:
                continue
            i = int(np.argmax(r))  # first row reaching the max = first strict improvement
            score[p] = int(r[i])
            qe[p] = i + 1
            te[p] = i + 1 + int(boff_host[p]) - hb + int(rd[x, i])
    return score, qe, te

"""TEST INFRASTRUCTURE ONLY -- CPU restatement of the reference's wavefront
aligner (graphmambaformer/alignment/extension.py:308-463 WavefrontAligner of
sarakh1999/GraphMambaFormer) and of the WFA branch of ExtensionEngine
(extension.py:643-730), the checker for the GPU WFA.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU baseline may use it.

  wfa_align   extension.py:352-423  global unit-cost alignment; diagonal
              k = i - j, offset = query bases consumed; wavefront s spans
              k in [-s, s]; next = max(clip(sub), clip(ins), clip(del)) with
              sub = prev[k] + 1, ins = prev[k-1] + 1, del = prev[k+1]; then the
              match extension (a query N never matches); done when diagonal
              n - m reaches n; distance above max_distance -> None
  traceback   extension.py:425-463  from the last wavefront back: the
              candidate with the largest (unclipped) offset, ties X, I, D;
              the gap to the reach is a run of '='
  wfa_score   extension.py:703-718  score of a CIGAR: matches * match -
              mismatches * mismatch - sum(open + (len-1) * extend) over gaps

Pins (tests/test_wfa.py): the reference's golden vectors
(tests/golden/wfa.json), Levenshtein distance by full DP (the reference's own
test, test_alignment_stages.py:1063-1091), CIGAR consumption (= X I sum to
the query, = X D to the target, X I D sum to the distance), the N case of
test_alignment_stages.py:1106-1110 and the over-budget signal.
"""

from __future__ import annotations

import numpy as np

N_CODE = 5


def _merge(ops):
    out = []
    for op, n in ops:
        if n <= 0:
            continue
        if out and out[-1][0] == op:
            out[-1] = (op, out[-1][1] + n)
        else:
            out.append((op, n))
    return out


def _extend(off, k, q, t):
    n, m = len(q), len(t)
    i = off
    if i < 0:
        return i
    j = i - k
    while i < n and 0 <= j < m and q[i] == t[j] and q[i] >= 0 and q[i] != N_CODE:
        i += 1
        j += 1
    return i


def wfa_align(q, t, max_distance=4096):
    """(distance, cigar) or None when the distance exceeds max_distance."""
    q = [int(x) for x in q]
    t = [int(x) for x in t]
    n, m = len(q), len(t)
    if n == 0 or m == 0:
        return max(n, m), _merge([("I", n), ("D", m)])
    final_k = n - m
    hist = [[_extend(0, 0, q, t)]]  # wavefront s: offsets of k = -s .. s
    s = 0
    while True:
        cur = hist[-1]
        if -s <= final_k <= s and cur[final_k + s] >= n:
            return s, _traceback(hist, final_k)
        if s >= max_distance:
            return None
        s += 1
        prev = cur

        def pred(k):  # offset of diagonal k in the previous wavefront (-1 if absent)
            x = k + (s - 1)
            return prev[x] if 0 <= x < len(prev) else -1

        def clip(v, k):
            c = v - k
            return -1 if (v < 0 or v > n or c < 0 or c > m) else v

        new = []
        for k in range(-s, s + 1):
            a = pred(k)
            sub = clip(a + 1 if a >= 0 else -1, k)
            b = pred(k - 1)
            ins = clip(b + 1 if b >= 0 else -1, k)
            dele = clip(pred(k + 1), k)
            new.append(_extend(max(sub, ins, dele), k, q, t))
        hist.append(new)


def _traceback(hist, final_k):
    ops = []
    k = final_k
    for s in range(len(hist) - 1, 0, -1):
        reach = hist[s][k + s]
        prev = hist[s - 1]

        def pred(kk):
            x = kk + (s - 1)
            if 0 <= x < len(prev) and prev[x] >= 0:
                return prev[x]
            return -1

        sub, ins, dele = pred(k), pred(k - 1), pred(k + 1)
        cands = [("X", sub + 1 if sub >= 0 else -1, k), ("I", ins + 1 if ins >= 0 else -1, k - 1),
                 ("D", dele, k + 1)]
        best = cands[0]
        for c in cands[1:]:
            if c[1] > best[1]:
                best = c
        op, start, k = best
        ops.extend(["="] * max(reach - start, 0))
        ops.append(op)
    # wavefront 0 holds diagonal 0 only (numpy indexing: -1 wraps onto it)
    if k not in (0, -1):
        raise IndexError("traceback left the first wavefront")
    ops.extend(["="] * hist[0][0])
    ops.reverse()
    rle = []
    for op in ops:
        if rle and rle[-1][0] == op:
            rle[-1] = (op, rle[-1][1] + 1)
        else:
            rle.append((op, 1))
    return _merge(rle)


def levenshtein(a, b):
    prev = list(range(len(b) + 1))
    for i in range(1, len(a) + 1):
        cur = [i] + [0] * len(b)
        for j in range(1, len(b) + 1):
            cur[j] = min(prev[j] + 1, cur[j - 1] + 1, prev[j - 1] + (a[i - 1] != b[j - 1]))
        prev = cur
    return prev[-1]


def wfa_score(cigar, match, mismatch, gap_open, gap_extend):
    cnt = {"=": 0, "X": 0, "I": 0, "D": 0}
    for op, n in cigar:
        cnt[op] = cnt.get(op, 0) + n
    gap = sum(gap_open + max(n - 1, 0) * gap_extend for op, n in cigar if op in ("I", "D"))
    return cnt["="] * match - cnt["X"] * mismatch - gap, cnt


def wfa_window(chain, read_len, ref_len):
    """Target window of the WFA branch (extension.py:654-662): start at the
    chain's read-zero diagonal, length read_len + the chain's net indel."""
    start = int(np.clip(chain["ref_start"] - chain["read_start"], 0, max(ref_len - 1, 0)))
    net = (chain["ref_end"] - chain["ref_start"]) - (chain["read_end"] - chain["read_start"])
    tl = max(1, read_len + net)
    end = min(ref_len, start + tl)
    return start, end

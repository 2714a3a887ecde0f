"""Pure-Python brute-force checks (small inputs only).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

These are independent formulations, not copies of the oracle's recurrences:

* ``full_gotoh_score``  unbanded full-matrix Gotoh with the textbook
  E/F recurrences over H (the reference's own test oracle,
  tests/test_alignment_stages.py:977-996).  Equals the banded result when the
  band covers the whole matrix.
* ``rescore_cigar``  scores a CIGAR with the affine scheme (gap of length L
  costs open + (L-1)*extend, extension.py:154-156); must reproduce the
  reported score exactly.
* ``chain_dp_bruteforce``  textbook O(n^2) chaining loop
  (tests/test_alignment_stages.py:372-397) with exact integer arithmetic for
  the integer profile, float64 otherwise.
* ``sw_bruteforce_all_optimal``  every cell reaching the best score (for tie
  tests: the reported end must be the lexicographically first (i, j)).
"""

from __future__ import annotations

import math

NEG = float("-inf")


def _sub(a, b, ma, mi):
    return ma if (a == b and a >= 0 and a != 5) else -mi


def full_gotoh_matrices(q, t, ma, mi, go, ge):
    n, m = len(q), len(t)
    H = [[0.0] * (m + 1) for _ in range(n + 1)]
    E = [[NEG] * (m + 1) for _ in range(n + 1)]
    F = [[NEG] * (m + 1) for _ in range(n + 1)]
    for i in range(1, n + 1):
        for j in range(1, m + 1):
            E[i][j] = max(H[i][j - 1] - go, E[i][j - 1] - ge)
            F[i][j] = max(H[i - 1][j] - go, F[i - 1][j] - ge)
            H[i][j] = max(0.0, H[i - 1][j - 1] + _sub(int(q[i - 1]), int(t[j - 1]), ma, mi),
                          E[i][j], F[i][j])
    return H, E, F


def full_gotoh_score(q, t, ma, mi, go, ge):
    H, _, _ = full_gotoh_matrices(q, t, ma, mi, go, ge)
    return max(max(row) for row in H)


def first_best_cell(q, t, ma, mi, go, ge):
    """(score, i, j) with the first row of strict improvement and, inside it, the
    first column -- the lexicographically smallest (i, j) among maxima."""
    H, _, _ = full_gotoh_matrices(q, t, ma, mi, go, ge)
    best, bi, bj = 0.0, 0, 0
    for i in range(1, len(q) + 1):
        for j in range(1, len(t) + 1):
            if H[i][j] > best:
                best, bi, bj = H[i][j], i, j
    return best, bi, bj


def rescore_cigar(cigar, q, t, read_start, ref_start, ma, mi, go, ge):
    """Affine score of an alignment given as a CIGAR over (=, X, I, D)."""
    i, j, s = read_start, ref_start, 0
    for op, n in cigar:
        if op in "=X":
            for _ in range(n):
                s += _sub(int(q[i]), int(t[j]), ma, mi)
                if (op == "=") != (_sub(int(q[i]), int(t[j]), ma, mi) == ma):
                    raise AssertionError("=/X op disagrees with the bases")
                i += 1
                j += 1
        elif op == "I":
            s -= go + (n - 1) * ge
            i += n
        elif op == "D":
            s -= go + (n - 1) * ge
            j += n
        else:
            raise AssertionError(f"unexpected op {op}")
    return s, i, j


def chain_dp_bruteforce(read_end, ref_end, weight, max_lookback, max_gap, penalty):
    """O(n^2) chaining, ``penalty(gap)`` a callable; ties keep the first j."""
    n = len(read_end)
    f = [0] * n
    parent = [-1] * n
    L = max(1, max_lookback)
    for i in range(n):
        best, bj = weight[i], -1
        for j in range(max(0, i - L), i):
            dq = int(read_end[i]) - int(read_end[j])
            dr = int(ref_end[i]) - int(ref_end[j])
            if dq <= 0 or dr <= 0 or dq > max_gap or dr > max_gap:
                continue
            s = (f[j] + min(dq, dr, weight[i])) - penalty(abs(dr - dq))
            if s > best:
                best, bj = s, j
        f[i], parent[i] = best, bj
    return f, parent


def int_penalty(go, ge):
    return lambda g: (go + ge * g) if g > 0 else 0


def float_penalty(go, ge, lc):
    return lambda g: (go + ge * g + lc * math.log2(g + 1.0)) if g > 0 else 0.0

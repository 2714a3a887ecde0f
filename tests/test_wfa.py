"""CPU pins of the WFA oracle (oracle/wfa.py): the reference's golden vectors,
Levenshtein by full DP, CIGAR consumption, the N case and the budget."""

import numpy as np

from oracle import oracle as O
from oracle import wfa as W
from tests.golden.load import load


def test_wfa_golden():
    for n, case in enumerate(load("wfa.json")["align"]):
        got = W.wfa_align(O.encode(case["q"]), O.encode(case["t"]), case["max_distance"])
        if case["distance"] is None:
            assert got is None, n
            continue
        d, cig = got
        assert d == case["distance"], n
        assert [list(x) for x in cig] == case["cigar"], n


def test_wfa_levenshtein_and_consumption():
    """test_alignment_stages.py:1063-1091 restated on random indel/sub cases."""
    rng = np.random.default_rng(5)
    for _ in range(150):
        L = int(rng.integers(1, 40))
        a = rng.integers(1, 5, size=L).tolist()
        b = list(a)
        for _ in range(int(rng.integers(0, 6))):
            p = int(rng.integers(0, max(len(b), 1)))
            act = int(rng.integers(0, 3))
            if act == 0 and b:
                b[p] = int(rng.integers(1, 5))
            elif act == 1 and b:
                b.pop(p)
            else:
                b.insert(p, int(rng.integers(1, 5)))
        d, cig = W.wfa_align(a, b)
        assert d == W.levenshtein(a, b)
        assert sum(k for op, k in cig if op in "=XI") == len(a)
        assert sum(k for op, k in cig if op in "=XD") == len(b)
        assert sum(k for op, k in cig if op in "XID") == d


def test_wfa_n_and_budget():
    """test_alignment_stages.py:1094-1110 restated."""
    d, cig = W.wfa_align(O.encode("AN"), O.encode("AN"))
    assert d == 1 and cig == [("=", 1), ("X", 1)]
    rng = np.random.default_rng(6)
    a, b = rng.integers(1, 5, size=60), rng.integers(1, 5, size=60)
    assert W.wfa_align(a, b, max_distance=2) is None

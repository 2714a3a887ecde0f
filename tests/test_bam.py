"""CPU pins of the BAM CIGAR record oracle (oracle/bam.py)."""

from oracle import bam as B


def test_record_cigar_hand_cases():
    assert B.record_cigar([("=", 5), ("X", 1), ("=", 3)], 2, 11, 15) == \
        [("S", 2), ("=", 5), ("X", 1), ("=", 3), ("S", 4)]
    assert B.record_cigar([("=", 4)], 0, 4, 4) == [("=", 4)]
    assert B.record_cigar([("S", 1), ("=", 4)], 2, 6, 6) == [("S", 3), ("=", 4)]
    assert B.record_cigar([], 0, 0, 7) == [("S", 7)]


def test_bam_record_hand_cases():
    c = [("S", 2), ("=", 5), ("I", 1), ("D", 2), ("X", 1), ("S", 3)]
    runs, flag = B.bam_record(c, 12, 100, 108, 1)
    assert flag == 0 and runs == [(2 << 4) | 4, (5 << 4) | 7, (1 << 4) | 1, (2 << 4) | 2,
                                  (1 << 4) | 8, (3 << 4) | 4]
    runs_r, flag_r = B.bam_record(c, 12, 100, 108, -1)
    assert flag_r == 0x10 and runs_r == list(reversed(runs))
    assert B.bam_record(c, 13, 100, 108, 1) == ([], 0x4)        # length rule
    assert B.bam_record(c, 13, 100, 108, -1) == ([], 0x14)      # reverse flag kept
    assert B.bam_record([], 12, 100, 108, 1) == ([], 0x4)       # no CIGAR
    assert B.bam_record(c, 12, 100, 100, 1) == ([], 0x4)        # empty reference span
    # invariants: reversal is an involution; mapped query length = read length
    back, _ = B.bam_record(list(reversed(c)), 12, 0, 1, -1)
    assert back == runs
    q = sum(r >> 4 for r in runs if (r & 15) in (0, 1, 4, 7, 8))
    assert q == 12

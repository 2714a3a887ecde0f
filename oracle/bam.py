"""TEST INFRASTRUCTURE ONLY -- CPU restatement of how the reference turns an
extension into the CIGAR and flag of a BAM record, the checker for
gmfa_cigar_bam.  Only tests/ may use it.

  record_cigar  extension.py:628-640  merge_cigar([S(read_start)] + cigar +
                [S(read_len - read_end)]); types.py:307-322 merge_cigar drops
                empty runs and merges equal neighbours
  bam_record    formats.py:52, 287-289, 343-382  mapped = CIGAR, sequence and
                ref_end > ref_start; strand -1 -> flag 0x10 and the CIGAR
                reversed; then the SAM rule: query-consuming length (M I S = X)
                must equal the read length, else unmapped (flag |= 0x4, the
                reverse flag kept, no CIGAR); ops coded M0 I1 D2 N3 S4 H5 P6 =7 X8

Pins (tests/test_bam.py): hand cases derived from those lines (clips on both
ends, zero clips, merges into a clip, reverse strand, the length rule
unmapping a forward and a reverse record, an empty CIGAR) and invariants
(reversal is an involution, query length of a mapped record = read length).
"""

from __future__ import annotations

OP_CODE = {"M": 0, "I": 1, "D": 2, "N": 3, "S": 4, "H": 5, "P": 6, "=": 7, "X": 8}
QUERY_OPS = {"M", "I", "S", "=", "X"}


def merge_cigar(ops):
    out = []
    for op, n in ops:
        if n <= 0:
            continue
        if out and out[-1][0] == op:
            out[-1] = (op, out[-1][1] + n)
        else:
            out.append((op, n))
    return out


def record_cigar(cigar, read_start, read_end, read_len):
    return merge_cigar([("S", read_start)] + list(cigar) + [("S", read_len - read_end)])


def bam_record(cigar, read_len, ref_start, ref_end, strand):
    """(BAM runs [len << 4 | code], flag) of one record."""
    mapped = bool(cigar) and read_len > 0 and ref_end > ref_start
    flag = 0
    cig = list(cigar)
    if mapped and strand == -1:
        flag |= 0x10
        cig = list(reversed(cig))
    if mapped and sum(n for op, n in cig if op in QUERY_OPS) != read_len:
        mapped = False
    if not mapped:
        return [], flag | 0x4
    return [(n << 4) | OP_CODE[op] for op, n in cig], flag

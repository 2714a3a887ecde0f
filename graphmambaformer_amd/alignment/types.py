"""Stage data contracts of the path (graphmambaformer/alignment/types.py).

Same names, fields and semantics as the reference: columnar ``AnchorSet``
(0-based half-open, forward-reference coordinates), ``Chain``,
``ExtensionResult`` and the CIGAR helpers (types.py:26-432).  Only what the
chaining and extension stages use is kept.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

CIGAR_OPS: tuple[str, ...] = ("=", "X", "I", "D", "S", "H", "M", "N")


@dataclass
class AnchorSet:
    """Seed anchors of one read as parallel arrays (types.py:26-79)."""

    read_pos: np.ndarray
    ref_pos: np.ndarray
    length: np.ndarray
    strand: np.ndarray
    node_id: np.ndarray
    source: np.ndarray
    score: np.ndarray
    read_len: int = 0
    ref_len: int = 0

    def __post_init__(self) -> None:
        n = len(self.read_pos)
        for name in ("ref_pos", "length", "strand", "node_id", "source", "score"):
            if len(getattr(self, name)) != n:
                raise ValueError(f"AnchorSet field {name!r} has inconsistent length")

    def __len__(self) -> int:
        return len(self.read_pos)

    @property
    def diagonal(self) -> np.ndarray:
        return self.ref_pos.astype(np.int64) - self.read_pos.astype(np.int64)

    @property
    def read_end(self) -> np.ndarray:
        return self.read_pos + self.length

    @property
    def ref_end(self) -> np.ndarray:
        return self.ref_pos + self.length

    @classmethod
    def from_lists(cls, read_pos, ref_pos, length, strand, node_id=None, source=None,
                   read_len: int = 0, ref_len: int = 0) -> "AnchorSet":
        n = len(read_pos)
        return cls(read_pos=np.asarray(read_pos, dtype=np.int64),
                   ref_pos=np.asarray(ref_pos, dtype=np.int64),
                   length=np.asarray(length, dtype=np.int64),
                   strand=np.asarray(strand, dtype=np.int8),
                   node_id=(np.asarray(node_id, dtype=np.int64) if node_id is not None
                            else np.full(n, -1, dtype=np.int64)),
                   source=(np.asarray(source, dtype=np.int8) if source is not None
                           else np.zeros(n, dtype=np.int8)),
                   score=np.full(n, np.nan, dtype=np.float32),
                   read_len=read_len, ref_len=ref_len)

    def take(self, idx) -> "AnchorSet":
        idx = np.asarray(idx, dtype=np.int64)
        return AnchorSet(read_pos=self.read_pos[idx], ref_pos=self.ref_pos[idx],
                         length=self.length[idx], strand=self.strand[idx],
                         node_id=self.node_id[idx], source=self.source[idx],
                         score=self.score[idx], read_len=self.read_len, ref_len=self.ref_len)

    def for_strand(self, strand: int):
        idx = np.flatnonzero(self.strand == strand)
        return self.take(idx), idx

    def sorted_by_ref(self):
        """Chaining DP order: reference end, then read end, stable (types.py:149-152)."""
        order = np.lexsort((self.read_end, self.ref_end))
        return self.take(order), order


@dataclass
class Chain:
    """A collinear run of anchors (types.py:224-266)."""

    anchor_idx: np.ndarray
    score: float
    strand: int
    read_start: int
    read_end: int
    ref_start: int
    ref_end: int
    neural_score: float = float("nan")
    is_primary: bool = False

    def __len__(self) -> int:
        return len(self.anchor_idx)

    @property
    def read_span(self) -> int:
        return self.read_end - self.read_start

    @property
    def ref_span(self) -> int:
        return self.ref_end - self.ref_start

    def coverage(self, read_len: int) -> float:
        return self.read_span / max(read_len, 1)


@dataclass
class ExtensionResult:
    """Base-level alignment of one chain (types.py:287-322)."""

    score: float
    cigar: list
    read_start: int
    read_end: int
    ref_start: int
    ref_end: int
    n_match: int = 0
    n_mismatch: int = 0
    n_insertion: int = 0
    n_deletion: int = 0

    @property
    def cigar_string(self) -> str:
        return "".join(f"{n}{op}" for op, n in self.cigar)

    @property
    def identity(self) -> float:
        total = self.n_match + self.n_mismatch + self.n_insertion + self.n_deletion
        return self.n_match / total if total else 0.0

    @property
    def edit_distance(self) -> int:
        return self.n_mismatch + self.n_insertion + self.n_deletion


def run_length_encode(ops: Sequence[str]) -> list:
    """Per-column op string -> [(op, count)] runs (types.py:396-410)."""
    out: list = []
    for op in ops:
        if out and out[-1][0] == op:
            out[-1] = (op, out[-1][1] + 1)
        else:
            out.append((op, 1))
    return out


def cigar_read_length(cigar) -> int:
    return sum(n for op, n in cigar if op in ("=", "X", "M", "I", "S"))


def cigar_ref_length(cigar) -> int:
    return sum(n for op, n in cigar if op in ("=", "X", "M", "D", "N"))


def merge_cigar(cigar) -> list:
    """Merge equal neighbours, drop empty runs (types.py:423-432)."""
    out: list = []
    for op, n in cigar:
        if n <= 0:
            continue
        if out and out[-1][0] == op:
            out[-1] = (op, out[-1][1] + n)
        else:
            out.append((op, n))
    return out


#: BAM CIGAR op codes (data/formats.py:52 of the reference)
BAM_OP = {"M": 0, "I": 1, "D": 2, "N": 3, "S": 4, "H": 5, "P": 6, "=": 7, "X": 8}


def cigar_to_bam(cigar) -> np.ndarray:
    """[(op, n)] -> uint32 array of n << 4 | code."""
    return np.array([(n << 4) | BAM_OP[op] for op, n in cigar], dtype=np.uint32)


def bam_to_cigar(runs) -> list:
    inv = {v: k for k, v in BAM_OP.items()}
    return [(inv[int(r) & 15], int(r) >> 4) for r in runs]




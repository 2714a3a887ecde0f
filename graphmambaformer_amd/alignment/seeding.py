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

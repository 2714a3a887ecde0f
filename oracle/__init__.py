"""CPU oracle for the banded affine SW + traceback and anchor-chaining path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package, and only as the
checker.  The product package ``graphmambaformer_amd`` never imports it and
fails loudly when its HIP library is missing.

Contents
--------
``oracle.py``         ctypes wrapper over ``gmfa_oracle.c`` (scalar C restatement
                      of extension.py:78-302 and chaining.py:120-180, 530-576).
``ref_semantics.py``  float32 numpy transcription of the reference's row loop
                      (sentinel -1e30, exclusive cummax E) used to pin the
                      int32 sentinel mapping of the C restatement.
``brute.py``          pure-Python brute force: unbanded Gotoh, O(n^2) chaining,
                      CIGAR rescoring -- the reference's own test oracles
                      (tests/test_alignment_stages.py:372-397, 977-996).

Parity status per function is listed in the module docstrings; every function
here is pinned by at least one check in ``tests/test_oracle.py``.
"""

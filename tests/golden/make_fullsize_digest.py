"""Write per-pair / per-read digests of the oracle's outputs over the whole
long-read (16,384 pairs, half_band 256), noisy (8,192 pairs, half_band 751)
and chaining (4,096 reads) workloads, so the GPU tests can compare every pair
and every read at full size instead of a sample.

    python tests/golden/make_fullsize_digest.py [--threads N]

Runs the CPU oracle (oracle/gmfa_oracle.c: extension.py:78-302 and
chaining.py:129-180, 530-617 restated) on all host cores; about 5 minutes on 8
cores and ~15 GB of memory at peak (the oracle keeps H/E/F of a pair, 1.8 GB for
the longest noisy pair).  Writes ``tests/golden/fullsize_digest.npz``:

* ``<cfg>_input``: blake2b-8 of the workload's queries, targets and band
  offsets (a check that the generator reproduced the same pairs);
* ``<cfg>_score / _qend / _tend / _rs / _ts``: int32 per pair (score, 1-based
  ends, 0-based starts);
* ``<cfg>_cigar``: uint64 per pair, blake2b-8 of the CIGAR as BAM runs
  (len << 4 | op, little-endian uint32; = 7, X 8, I 1, D 2);
* ``chains_input``, ``chains_nkept`` (int32 per read), ``chains_best``
  (float64 score of the primary, 0 if none) and ``chains_digest`` (uint64 per
  read: blake2b-8 over the kept chains in order, each as its float64 score,
  its member count and its members as anchor indices in DP order), for the
  anchors in the shuffled order of ``chains_input``.

Every value is the oracle's; nothing here runs the product.
"""

from __future__ import annotations

import argparse
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "fullsize_digest.npz")
OPCODE = {"=": 7, "X": 8, "I": 1, "D": 2}


def h8(*parts) -> int:
    m = hashlib.blake2b(digest_size=8)
    for p in parts:
        m.update(np.ascontiguousarray(p).tobytes())
    return int.from_bytes(m.digest(), "little")


def sw_input_digest(w) -> int:
    m = hashlib.blake2b(digest_size=8)
    for q, t in zip(w.queries, w.targets):
        m.update(np.int64(len(q)).tobytes())
        m.update(np.ascontiguousarray(q, np.uint8).tobytes())
        m.update(np.int64(len(t)).tobytes())
        m.update(np.ascontiguousarray(t, np.uint8).tobytes())
    m.update(np.ascontiguousarray(w.band_offset, np.int64).tobytes())
    return int.from_bytes(m.digest(), "little")


def cigar_digest(runs_u32: np.ndarray) -> int:
    return h8(np.ascontiguousarray(runs_u32, np.uint32))


def cigar_runs(cigar) -> np.ndarray:
    return np.array([(n << 4) | OPCODE[op] for op, n in cigar], np.uint32)


def chains_input_digest(w) -> int:
    return h8(np.asarray(w.read_end, np.int64), np.asarray(w.ref_end, np.int64),
              np.asarray(w.seg_off, np.int64))


def chain_digest(chains) -> int:
    """chains: [(score float, members int64 DP-order ids)] in kept order."""
    m = hashlib.blake2b(digest_size=8)
    for score, mem in chains:
        m.update(np.float64(score).tobytes())
        m.update(np.int64(len(mem)).tobytes())
        m.update(np.ascontiguousarray(mem, np.int64).tobytes())
    return int.from_bytes(m.digest(), "little")


def sw_digest(w, threads: int) -> dict:
    from oracle import oracle as O

    n = w.n_pairs
    cells = w.cells()
    order = np.argsort(-cells, kind="stable")  # longest first: the pool drains evenly
    out = {k: np.zeros(n, np.int32) for k in ("score", "qend", "tend", "rs", "ts")}
    out["cigar"] = np.zeros(n, np.uint64)
    chunk = 256
    t0 = time.time()
    for c0 in range(0, n, chunk):
        idx = order[c0:c0 + chunk]
        qs = [w.as_ref_codes(int(k))[0] for k in idx]
        ts = [w.as_ref_codes(int(k))[1] for k in idx]
        res = O.sw_align_pairs(qs, ts, np.asarray(w.band_offset)[idx].astype(np.int64),
                               w.half_band, w.match, w.mismatch, w.gap_open, w.gap_extend,
                               threads=threads)
        for k, e in zip(idx, res):
            out["score"][k] = e["score"]
            out["qend"][k] = e["read_end"]
            out["tend"][k] = e["ref_end"]
            out["rs"][k] = e["read_start"]
            out["ts"][k] = e["ref_start"]
            out["cigar"][k] = cigar_digest(cigar_runs(e["cigar"]))
        print(f"  {w.name}: {min(c0 + chunk, n)}/{n} pairs, {time.time() - t0:.0f} s", flush=True)
    return out


def chains_input(w):
    """The chaining workload as the device pipeline takes it: every read's
    anchors shuffled (seeded), as (read_pos, ref_pos, id) with id = the
    anchor's index in the read's DP order.  Reads stay contiguous."""
    seg_of = np.repeat(np.arange(w.n_reads), np.diff(w.seg_off))
    shuf = np.lexsort((np.random.default_rng(0).random(len(seg_of)), seg_of))
    rp = w.read_end[shuf] - 15
    fp = w.ref_end[shuf] - 15
    ids = (np.arange(len(seg_of)) - w.seg_off[seg_of])[shuf]
    return rp, fp, ids


def chains_reference(w, max_lookback: int, max_gap: int, reads: int | None = None):
    """Per read: the oracle's AffineChainer.chain on the read's anchors in the
    shuffled input order of chains_input (+ strand, span 15, reference-default
    float64 profile), members named by their DP-order ids.  The input order
    matters where anchors tie on (ref_end, read_end): the stable DP-order sort
    keeps them in input order, as the reference's np.lexsort does."""
    from oracle import oracle as O

    rp, fp, ids = chains_input(w)
    nk = np.zeros(w.n_reads, np.int32)
    best = np.zeros(w.n_reads, np.float64)
    dig = np.zeros(w.n_reads, np.uint64)
    for r in range(w.n_reads if reads is None else reads):
        a, b = int(w.seg_off[r]), int(w.seg_off[r + 1])
        kept = O.chain_read(rp[a:b], fp[a:b], np.full(b - a, 15), np.ones(b - a, np.int64),
                            max_lookback=max_lookback, max_gap=max_gap)
        nk[r] = len(kept)
        best[r] = kept[0]["score"] if kept else 0.0
        dig[r] = chain_digest([(c["score"], ids[a:b][np.asarray(c["anchor_idx"], np.int64)])
                               for c in kept])
    return nk, best, dig


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--only-chains", action="store_true",
                    help="recompute the chaining entries, keep the SW ones")
    args = ap.parse_args()
    sys.path.insert(0, ROOT)
    from graphmambaformer_amd import workloads as WL
    from oracle import oracle as O

    O.build()
    arrays = {}
    if args.only_chains:
        old = np.load(OUT)
        arrays = {k: old[k] for k in old.files if not k.startswith("chains_")}
    for name in ("sw_long", "sw_noisy"):
        if args.only_chains:
            break
        w = getattr(WL, name)()
        print(f"{name}: {w.n_pairs} pairs, {w.cells().sum():.3e} cells", flush=True)
        arrays[f"{name}_input"] = np.array([sw_input_digest(w)], np.uint64)
        for k, v in sw_digest(w, args.threads).items():
            arrays[f"{name}_{k}"] = v
    wc = WL.chains()
    print(f"chains: {wc.n_reads} reads", flush=True)
    nk, best, dig = chains_reference(wc, wc.max_lookback, wc.max_gap)
    arrays["chains_input"] = np.array([chains_input_digest(wc)], np.uint64)
    arrays["chains_nkept"] = nk
    arrays["chains_best"] = best
    arrays["chains_digest"] = dig
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)", flush=True)


if __name__ == "__main__":
    main()

"""Split launches (the longest pairs over several wavefronts, one band slice
each) against the plain one-pair-per-wavefront launch and the oracle, bit for
bit: scores, 1-based ends, CIGARs, starts and op counts.

The slices exchange one boundary per query row (the horizontal prefix to the
right, a ghost cell of the vertical input to the left), so the cases push
cells across every slice boundary: bands hanging over either end of the
target, targets shorter than the band, query lengths off the 8-row blocks,
long gaps (horizontal runs crossing boundaries), gap_open == gap_extend (no
horizontal continuation bit) and several scorings.  Every case prints its
name first."""

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
HB = 751  # W = 1503: the compiled split variants (2 x 12 x 64, 3 x 8 x 64, 4 x 6 x 64 over 24 x 64)


def _say(*a):
    print(*a, flush=True)


def _pair(rng, n, gap_rate=0.02):
    q = rng.integers(1, 5, size=n).astype(np.int16)
    t = q.tolist()
    for _ in range(int(rng.integers(0, max(2, int(n * gap_rate))))):
        p = int(rng.integers(0, max(len(t), 1)))
        k = int(rng.integers(0, 4))
        if k == 0:
            t[p:p + 1] = [int(rng.integers(1, 5))]
        elif k == 1:
            del t[p:p + int(rng.integers(1, 40))]  # long deletions: horizontal runs
        elif k == 2:
            t[p:p] = rng.integers(1, 5, size=int(rng.integers(1, 40))).tolist()
        else:
            t[p:p] = [int(rng.integers(1, 5))] * int(rng.integers(1, 4))
    pre = rng.integers(1, 5, size=int(rng.integers(0, 400))).tolist()
    t = np.array(pre + t + rng.integers(1, 5, size=int(rng.integers(0, 400))).tolist(), np.int16)
    return q, t, len(pre)


def _batch(rng, n_pairs, lo, hi):
    qs, ts, offs = [], [], []
    for p in range(n_pairs):
        n = int(rng.integers(lo, hi + 1))
        q, t, pre = _pair(rng, n)
        kind = p % 6
        if kind == 0:
            off = pre  # band centred on the true diagonal
        elif kind == 1:
            off = int(rng.integers(-HB, 0))  # band hanging off the left end
        elif kind == 2:
            off = len(t) - len(q) + int(rng.integers(0, HB))  # off the right end
        elif kind == 3:
            t = t[:int(rng.integers(1, 200))]  # target far shorter than the band
            off = int(rng.integers(-50, 50))
        elif kind == 4:
            off = pre + int(rng.integers(-600, 600))
        else:
            off = int(rng.integers(0, 2 * HB))
        qs.append(q)
        ts.append(t)
        offs.append(off)
    return qs, ts, np.array(offs, np.int32)


def _run(K, inp, scoring, split_rows, slices):
    lpt = torch.from_numpy(np.argsort(-inp.qlen_host, kind="stable").astype(np.int32)).to(DEV)
    f = K.sw_forward(inp, *scoring, HB, traceback=True, order=lpt, cigar=True,
                     split_rows=split_rows, split_slices=slices)
    torch.cuda.synchronize()
    w = f.walk
    return dict(score=f.score.cpu().numpy(), qend=f.query_end.cpu().numpy(),
                tend=f.target_end.cpu().numpy(), cigar=K.decode_cigars(w),
                rs=w.read_start.cpu().numpy(), ts=w.ref_start.cpu().numpy(),
                counts=w.counts.cpu().numpy())


def _same(a, b, label):
    for key in ("score", "qend", "tend", "rs", "ts", "counts"):
        bad = np.flatnonzero((a[key] != b[key]).reshape(len(a[key]), -1).any(axis=1))
        assert len(bad) == 0, (label, key, bad[:8].tolist())
    for p, (x, y) in enumerate(zip(a["cigar"], b["cigar"])):
        assert x == y, (label, "cigar", p)


def _oracle(qs, ts, offs, scoring, got, label, pick):
    exp = O.sw_align_pairs([qs[k] for k in pick], [ts[k] for k in pick],
                           offs[pick].astype(np.int64), HB, *scoring, threads=16)
    for e, k in zip(exp, pick):
        ctx = (label, int(k), len(qs[k]), len(ts[k]), int(offs[k]))
        assert int(got["score"][k]) == e["score"], ctx
        assert int(got["qend"][k]) == e["read_end"] and int(got["tend"][k]) == e["ref_end"], ctx
        assert got["cigar"][k] == e["cigar"], ctx
        assert int(got["rs"][k]) == e["read_start"] and int(got["ts"][k]) == e["ref_start"], ctx


def test_split_variants_compiled():
    from graphmambaformer_amd import kernels as K

    _say("case: split variants for half_band 751")
    assert sorted(K.split_variants(HB)) == [(6, 4), (8, 3), (12, 2)]
    assert K.split_variants(256) == []


@pytest.mark.parametrize("scoring", [(2, 4, 4, 2), (2, 4, 6, 2), (2, 4, 2, 2), (1, 3, 5, 1)])
@pytest.mark.parametrize("slices", [2, 3, 4])
def test_every_pair_split_vs_plain_and_oracle(scoring, slices):
    """split_rows = 1: every pair runs as slices."""
    from graphmambaformer_amd import kernels as K

    _say("case: all pairs split, slices", slices, "scoring", scoring)
    rng = np.random.default_rng(1000 * slices + sum(scoring))
    qs, ts, offs = _batch(rng, 60, 1, 2600)
    inp = K.upload_pairs(qs, ts, offs, DEV, alphabet="ref")
    plain = _run(K, inp, scoring, 0, 0)
    split = _run(K, inp, scoring, 1, slices)
    _same(plain, split, ("split", slices, scoring))
    _oracle(qs, ts, offs, scoring, split, ("oracle", slices, scoring), np.arange(0, 60, 3))


@pytest.mark.parametrize("slices", [2, 3, 4])
def test_mixed_batch_both_phases(slices):
    """Long pairs as slices, the rest one per wavefront, in the same launch
    (more long pairs than workgroups: each workgroup splits one, the later
    long pairs run on one wavefront)."""
    from graphmambaformer_amd import kernels as K

    _say("case: mixed split / plain batch, slices", slices)
    rng = np.random.default_rng(77 + slices)
    qs, ts, offs = _batch(rng, 1500, 200, 9000)
    inp = K.upload_pairs(qs, ts, offs, DEV, alphabet="ref")
    plain = _run(K, inp, (2, 4, 4, 2), 0, 0)
    split = _run(K, inp, (2, 4, 4, 2), 3000, slices)
    _same(plain, split, ("mixed", slices))
    _oracle(qs, ts, offs, (2, 4, 4, 2), split, ("mixed oracle", slices), np.arange(0, 1500, 100))


def test_split_request_with_n_masks_runs_plainly():
    """N masks are outside the split variants: the launch falls back to the
    plain kernel with the same outputs."""
    from graphmambaformer_amd import kernels as K

    _say("case: split request with N masks")
    rng = np.random.default_rng(5)
    qs, ts, offs = _batch(rng, 24, 100, 3000)
    for k in range(0, 24, 3):
        qs[k][int(rng.integers(0, len(qs[k])))] = 5
    inp = K.upload_pairs(qs, ts, offs, DEV, alphabet="ref")
    assert inp.has_n
    plain = _run(K, inp, (2, 4, 4, 2), 0, 0)
    split = _run(K, inp, (2, 4, 4, 2), 1, 4)
    _same(plain, split, "nmask")
    _oracle(qs, ts, offs, (2, 4, 4, 2), split, "nmask oracle", np.arange(24))

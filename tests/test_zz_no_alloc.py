"""gmfa_sw_forward_noalloc (static device scratch, no allocation) against
gmfa_sw_forward with the fused CIGAR walk, bit for bit, on calls in a row
whose inputs shrink, through the C ABI.  Every case prints its name first."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _say(*a):
    print(*a, flush=True)


def _pairs(rng, n, lo, hi, with_n):
    qs, ts, offs = [], [], []
    for _ in range(n):
        L = int(rng.integers(lo, hi + 1))
        q = rng.integers(1, 5, size=L).astype(np.int16)
        t = q.tolist()
        for _ in range(int(rng.integers(0, max(2, L // 40)))):
            p = int(rng.integers(0, len(t)))
            k = int(rng.integers(0, 3))
            if k == 0:
                t[p] = int(rng.integers(1, 5))
            elif k == 1 and len(t) > 3:
                del t[p:p + int(rng.integers(1, 4))]
            else:
                t[p:p] = [int(rng.integers(1, 5))] * int(rng.integers(1, 4))
        pre = int(rng.integers(0, 30))
        t = rng.integers(1, 5, size=pre).tolist() + t + rng.integers(1, 5, size=20).tolist()
        t = np.array(t, np.int16)
        if with_n:
            q[int(rng.integers(0, len(q)))] = 5
            t[int(rng.integers(0, len(t)))] = 5
        qs.append(q)
        ts.append(t)
        offs.append(pre)
    return qs, ts, np.array(offs, np.int32)


def _outputs(fwd):
    from graphmambaformer_amd import kernels as K

    w = fwd.walk
    return dict(score=fwd.score.cpu().numpy(), qend=fwd.query_end.cpu().numpy(),
                tend=fwd.target_end.cpu().numpy(), cigar=K.decode_cigars(w),
                rs=w.read_start.cpu().numpy(), ts=w.ref_start.cpu().numpy(),
                counts=w.counts.cpu().numpy())


def _same(a, b, label):
    for key in ("score", "qend", "tend", "rs", "ts", "counts"):
        assert np.array_equal(a[key], b[key]), (label, key)
    assert a["cigar"] == b["cigar"], (label, "cigar")


# (pairs, length range, half_band, N masks, tile): two calls in a row per
# group, the second smaller; the last group's pairs are long enough that the
# slabs fit fewer wavefronts than pairs (slabs reused pair after pair)
CASES = [
    [(700, (100, 2500), 64, False, None), (90, (50, 900), 64, False, None)],
    [(1024, (150, 150), 32, True, (5, 16)), (200, (20, 150), 32, True, (5, 16))],
    [(40, (20000, 60000), 751, False, None), (9, (5000, 30000), 751, False, None)],
]


@pytest.mark.parametrize("group", range(len(CASES)))
def test_noalloc_matches_forward(group):
    from graphmambaformer_amd import kernels as K

    rng = np.random.default_rng(100 + group)
    scoring = (2, 4, 4, 2)
    for call, (n, (lo, hi), hb, with_n, tile) in enumerate(CASES[group]):
        _say("case: noalloc group", group, "call", call, "pairs", n, "len", lo, hi, "hb", hb,
             "N", with_n, "tile", tile)
        qs, ts, offs = _pairs(rng, n, lo, hi, with_n)
        inp = K.upload_pairs(qs, ts, offs, DEV, alphabet="ref")
        t = K.resolve_tile(hb, *tile) if tile else None
        ref = _outputs(K.sw_forward(inp, *scoring, hb, traceback=True, tile=t, cigar=True))
        got = _outputs(K.sw_forward_noalloc(inp, *scoring, hb, tile=t))
        torch.cuda.synchronize()
        _same(ref, got, (group, call))
        # again with the scratch holding this call's planes
        got2 = _outputs(K.sw_forward_noalloc(inp, *scoring, hb, tile=t))
        _same(ref, got2, (group, call, "repeat"))


def test_noalloc_rejects_planes_beyond_the_scratch():
    from graphmambaformer_amd import _lib
    from graphmambaformer_amd import kernels as K

    _say("case: noalloc size limit")
    rng = np.random.default_rng(7)
    # 1,000,000 query bases at half_band 1023: 125,000 blocks of 2,048 plane
    # dwords, about 1 GB for one wavefront's slab
    q = rng.integers(1, 5, size=1_000_000).astype(np.int16)
    inp = K.upload_pairs([q], [q], np.array([0], np.int32), DEV, alphabet="ref")
    with pytest.raises(ValueError, match="static scratch"):
        K.sw_forward_noalloc(inp, 2, 4, 4, 2, 1023)
    assert _lib.GMFA_NOALLOC_SCRATCH_BYTES == 256 << 20

"""The standalone walk with one lane per pair (sw_traceback on batches that
pack at least 8 pairs of at most 1,024 rows into each of 8 wavefronts per
SIMD) on pairs that hold N bases, against the fused walk (a wavefront per
pair, in the forward launch) on the same planes and against the oracle
(extension.py:243-285: an N never matches, so N columns are mismatches).
The lane walk reads the N masks at 64-bit stream positions; this is the walk
the x-drop extension path takes for very large batches of short pairs whose
reads or reference hold N.  Every case prints its name first."""

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
N_PAIRS = 60_000
SC = (2, 4, 4, 2)


def _say(*a):
    print(*a, flush=True)


def _pairs(rng, n, hb):
    """Queries of 150-600 bp (codes 1..4), targets = the query with 2%
    substitutions and a few short indels between random flanks; every other
    query and every third target carries ~1% N (code 5)."""
    qs, ts, offs = [], [], []
    for k in range(n):
        L = int(rng.integers(150, 601))
        q = rng.integers(1, 5, size=L).astype(np.int16)
        t = q.copy()
        sub = rng.random(L) < 0.02
        t[sub] = rng.integers(1, 5, size=int(sub.sum()))
        t = t.tolist()
        for _ in range(int(rng.integers(0, 4))):
            p = int(rng.integers(0, len(t)))
            if rng.random() < 0.5:
                del t[p:p + int(rng.integers(1, 6))]
            else:
                t[p:p] = rng.integers(1, 5, size=int(rng.integers(1, 6))).tolist()
        pre = rng.integers(1, 5, size=int(rng.integers(0, min(hb, 120) + 1))).tolist()
        post = rng.integers(1, 5, size=int(rng.integers(0, 120))).tolist()
        t = np.array(pre + t + post, np.int16)
        if k % 2 == 0:
            q[rng.random(L) < 0.01] = 5
        if k % 3 == 0:
            t[rng.random(len(t)) < 0.01] = 5
        qs.append(q)
        ts.append(t)
        offs.append(len(pre) + int(rng.integers(-(hb // 2), hb // 2 + 1)))
    return qs, ts, np.array(offs, np.int32)


@pytest.mark.parametrize("hb", [48, 751])
def test_lane_walk_with_n_masks_matches_fused_walk_and_oracle(hb):
    from graphmambaformer_amd import kernels as K

    _say(f"case: lane walk with N masks, {N_PAIRS} pairs, half_band {hb}")
    cus = torch.cuda.get_device_properties(DEV).multi_processor_count
    assert N_PAIRS > 7 * 8 * 4 * cus, "the batch must be large enough for the lane walk"
    rng = np.random.default_rng(4000 + hb)
    qs, ts, offs = _pairs(rng, N_PAIRS, hb)
    inp = K.upload_pairs(qs, ts, offs, DEV, alphabet="ref")
    assert inp.has_n
    fused = K.sw_forward(inp, *SC, hb, traceback=True, cigar=True)
    torch.cuda.synchronize()
    _say("  standalone lane walk over the same planes")
    lane = K.sw_traceback(inp, fused)
    torch.cuda.synchronize()
    for x, y, what in ((fused.walk.n_cigar, lane.n_cigar, "n_cigar"),
                       (fused.walk.read_start, lane.read_start, "read_start"),
                       (fused.walk.ref_start, lane.ref_start, "ref_start"),
                       (fused.walk.counts, lane.counts, "counts")):
        bad = torch.nonzero((x != y).reshape(N_PAIRS, -1).any(dim=1)).flatten()
        assert bad.numel() == 0, (hb, what, bad[:8].tolist())
    # every pair's runs, compared on the device (same output layout)
    assert torch.equal(fused.walk.cigar_off, lane.cigar_off)
    pos = torch.arange(lane.cigar.numel(), device=lane.cigar.device)
    owner = torch.searchsorted(lane.cigar_off, pos, right=True) - 1
    written = pos < lane.cigar_off[owner] + lane.n_cigar.to(torch.int64)[owner]
    diff = written & (fused.walk.cigar != lane.cigar)
    assert not bool(diff.any()), (hb, "cigar", torch.unique(owner[diff])[:8].tolist())
    cl = K.decode_cigars(lane)
    # an oracle sample weighted to pairs with N on both sides
    pick = np.unique(np.concatenate([np.arange(0, N_PAIRS, 6)[:24],
                                     rng.choice(N_PAIRS, size=16, replace=False)]))
    _say(f"  oracle sample of {len(pick)} pairs")
    exp = O.sw_align_pairs([qs[k] for k in pick], [ts[k] for k in pick],
                           offs[pick].astype(np.int64), hb, *SC, threads=8)
    score = fused.score.cpu().numpy()
    qe, te = fused.query_end.cpu().numpy(), fused.target_end.cpu().numpy()
    rs, tss = lane.read_start.cpu().numpy(), lane.ref_start.cpu().numpy()
    n_with_n = 0
    for k, e in zip(pick, exp):
        ctx = (hb, int(k), len(qs[k]), len(ts[k]), int(offs[k]))
        n_with_n += bool((qs[k] == 5).any() and (ts[k] == 5).any())
        assert int(score[k]) == e["score"], ctx
        assert int(qe[k]) == e["read_end"] and int(te[k]) == e["ref_end"], ctx
        assert cl[k] == e["cigar"], ctx
        assert int(rs[k]) == e["read_start"] and int(tss[k]) == e["ref_start"], ctx
    assert n_with_n >= 4, "the sample holds pairs with N in the query and the target"

"""A batch step captured in a HIP graph (kernels.SwGraph): every replay,
with eager work and synchronizations between replays, gives exactly the
eager pass's scores, ends and CIGARs; pairs rewritten in place between
replays (same lengths) give that batch's eager results; and those match the
oracle."""

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
SC = (2, 4, 4, 2)


def _say(*a):
    print(*a, flush=True)


def _batch(rng, lens, hb, off):
    from graphmambaformer_amd import workloads as WL

    qs, ts = [], []
    for L in lens:
        t = rng.integers(0, 4, size=int(L) + 2 * off, dtype=np.uint8)
        q = WL.mutate(rng, t[off:off + int(L)], 0.02, 0.5, 2)
        qs.append(q)
        ts.append(t)
    return qs, ts


def _outputs(fwd):
    from graphmambaformer_amd import kernels as K

    torch.cuda.synchronize()
    return (fwd.score.cpu().numpy().copy(), fwd.query_end.cpu().numpy().copy(),
            fwd.target_end.cpu().numpy().copy(), K.decode_cigars(fwd.walk),
            fwd.walk.read_start.cpu().numpy().copy(), fwd.walk.ref_start.cpu().numpy().copy())


def _same(a, b, label):
    for x, y, name in zip(a, b, ("score", "query_end", "target_end", "cigar", "read_start",
                                 "ref_start")):
        if isinstance(x, list):
            assert x == y, (label, name)
        else:
            assert np.array_equal(x, y), (label, name)


@pytest.mark.parametrize("hb,off", [(32, 32), (256, 0)])
def test_graph_replays_match_eager_and_oracle(hb, off):
    from graphmambaformer_amd import kernels as K

    _say(f"case: graph replays, half_band {hb}")
    rng = np.random.default_rng(21 + hb)
    lens = rng.integers(100, 1200 if hb > 32 else 300, size=96)
    qs, ts = _batch(rng, lens, hb, off)
    boff = np.full(len(qs), off, np.int32)
    inp = K.upload_pairs(qs, ts, boff, DEV, alphabet="2bit")
    tile = K.resolve_tile(hb)
    eager = _outputs(K.sw_forward(inp, *SC, hb, traceback=True, tile=tile, cigar=True))

    # the same batch against the oracle (a spread of pairs)
    pick = np.arange(0, len(qs), 7)
    exp = O.sw_align_pairs([qs[k].astype(np.int16) + 1 for k in pick],
                           [ts[k].astype(np.int16) + 1 for k in pick],
                           boff[pick].astype(np.int64), hb, *SC, threads=8)
    for j, k in enumerate(pick):
        e = exp[j]
        assert eager[0][k] == e["score"] and eager[3][k] == e["cigar"], k
        assert eager[1][k] == e["read_end"] and eager[2][k] == e["ref_end"], k

    g = K.SwGraph(inp, SC, hb, tile=tile, cigar=True)
    other_q, other_t = _batch(np.random.default_rng(5), lens[:40], hb, off)
    other = K.upload_pairs(other_q, other_t, boff[:40], DEV, alphabet="2bit")
    for rep in range(4):
        # the program's own eager work between replays
        K.sw_forward(other, *SC, hb, traceback=True, tile=tile, cigar=True)
        x = torch.randn(4096, device=DEV).cumsum(0)
        torch.cuda.synchronize()
        _same(_outputs(g.replay()), eager, f"replay {rep}")
        assert torch.isfinite(x).all()

    # new pairs with exactly the same query / target lengths (substitutions
    # only), written into the captured inputs
    r2 = np.random.default_rng(99)
    ts2 = [r2.integers(0, 4, size=len(t), dtype=np.uint8) for t in ts]
    qs2 = []
    for q, t in zip(qs, ts2):
        q2 = t[off:off + len(q)].copy() if off + len(q) <= len(t) else \
            r2.integers(0, 4, size=len(q), dtype=np.uint8)
        q2 = np.resize(q2, len(q)).astype(np.uint8)
        flip = r2.random(len(q2)) < 0.02
        q2[flip] = (q2[flip] + 1) % 4
        qs2.append(q2)
    inp2 = K.upload_pairs(qs2, ts2, boff, DEV, alphabet="2bit")
    eager2 = _outputs(K.sw_forward(inp2, *SC, hb, traceback=True, tile=tile, cigar=True))
    inp.qseq.copy_(inp2.qseq)
    inp.tseq.copy_(inp2.tseq)
    torch.cuda.synchronize()
    for rep in range(2):
        _same(_outputs(g.replay()), eager2, f"new pairs, replay {rep}")


@pytest.mark.parametrize("slices", [2, 4])
def test_graph_of_a_split_launch_matches_eager(slices):
    """A split launch (the longest pairs on the library's side stream, forked
    from and joined back into the captured stream) captured in a HIP graph:
    every replay, with eager work between replays, equals the eager split
    launch and the plain launch."""
    from graphmambaformer_amd import kernels as K

    _say(f"case: graph of a split launch, {slices} slices")
    rng = np.random.default_rng(40 + slices)
    lens = np.concatenate([rng.integers(3000, 9000, size=20), rng.integers(100, 2000, size=200)])
    qs, ts = _batch(rng, lens, 751, 0)
    boff = np.zeros(len(qs), np.int32)
    inp = K.upload_pairs(qs, ts, boff, DEV, alphabet="2bit")
    tile = K.resolve_tile(751)
    order = torch.from_numpy(np.argsort(-inp.qlen_host, kind="stable").astype(np.int32)).to(DEV)
    rows = int(np.sort(inp.qlen_host)[::-1][11])  # the 12 longest pairs split
    plain = _outputs(K.sw_forward(inp, *SC, 751, traceback=True, order=order, tile=tile,
                                  cigar=True))
    eager = _outputs(K.sw_forward(inp, *SC, 751, traceback=True, order=order, tile=tile,
                                  cigar=True, split_rows=rows, split_slices=slices))
    _same(eager, plain, "eager split vs plain")
    g = K.SwGraph(inp, SC, 751, order=order, tile=tile, cigar=True, split_rows=rows,
                  split_slices=slices)
    for rep in range(4):
        K.sw_forward(inp, *SC, 751, traceback=True, order=order, tile=tile, cigar=True,
                     workspace=K.make_workspace(inp, tile))
        x = torch.randn(4096, device=DEV).cumsum(0)
        torch.cuda.synchronize()
        _same(_outputs(g.replay()), plain, f"split replay {rep}")
        assert torch.isfinite(x).all()

"""The long-read (16,384 pairs, half_band 256) and noisy (8,192 pairs, half_band
751) batches at their full sizes on the GPU, checked through properties that
hold at any size, over every pair (tests/alignment_checks.py: span
consumption, '=' / 'X' columns against the bases, affine rescoring equal to
the score), with results that do not depend on the work order (longest-first
vs natural), and every pair bit-for-bit against the oracle's outputs over the whole batch
(score, ends, starts and a CIGAR digest per pair, tests/golden/
fullsize_digest.npz, written by tests/golden/make_fullsize_digest.py), for
the plain launch and, on the noisy batch, for split launches (the longest
pairs over 2, 3 or 4 wavefronts).  The chaining configuration is checked read by
read: the DP against the oracle, and the chain anchor sets of the device
pipeline (sort, DP, backtrack, select) against the oracle's digest."""

import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.alignment_checks import check_alignments
from tests.golden import make_fullsize_digest as D

DIGEST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "fullsize_digest.npz")

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _say(*a):
    print(*a, flush=True)


def _run(K, w, inp, order, workspace=None, split_rows=0, split_slices=0):
    fwd = K.sw_forward(inp, w.match, w.mismatch, w.gap_open, w.gap_extend, w.half_band,
                       traceback=True, order=order, cigar=True, workspace=workspace,
                       split_rows=split_rows, split_slices=split_slices)
    torch.cuda.synchronize()
    return fwd


def _against_digest(config, fwd, label):
    """Every pair's score, ends, starts and CIGAR against the oracle's digest."""
    d = np.load(DIGEST)
    tbk = fwd.walk
    got = {"score": fwd.score, "qend": fwd.query_end, "tend": fwd.target_end,
           "rs": tbk.read_start, "ts": tbk.ref_start}
    for key, t in got.items():
        g = t.cpu().numpy()
        bad = np.flatnonzero(g != d[f"{config}_{key}"])
        assert len(bad) == 0, (label, key, len(bad), bad[:8].tolist())
    nc = tbk.n_cigar.cpu().numpy().astype(np.int64)
    off = tbk.cigar_off.cpu().numpy()
    runs = tbk.cigar.cpu().numpy().view(np.uint32)
    exp = d[f"{config}_cigar"]
    bad = [p for p in range(len(nc))
           if D.cigar_digest(runs[off[p]:off[p] + nc[p]]) != int(exp[p])]
    assert not bad, (label, "cigar", len(bad), bad[:8])


def _properties(w, inp, fwd, label):
    """Every pair through the size-independent checks of alignment_checks."""
    tbk = fwd.walk
    ql = torch.tensor([len(q) for q in w.queries], dtype=torch.long, device=DEV)
    tl = torch.tensor([len(t) for t in w.targets], dtype=torch.long, device=DEV)
    return check_alignments(
        label, q_codes=torch.from_numpy(np.concatenate(w.queries).astype(np.uint8)).to(DEV),
        q_base=torch.cumsum(ql, 0) - ql,
        t_codes=torch.from_numpy(np.concatenate(w.targets).astype(np.uint8)).to(DEV),
        t_base=torch.cumsum(tl, 0) - tl, score=fwd.score, qend=fwd.query_end,
        tend=fwd.target_end, rs=tbk.read_start, ts=tbk.ref_start, n_cigar=tbk.n_cigar,
        cigar_off=tbk.cigar_off, cigar=tbk.cigar, counts=tbk.counts,
        scoring=(w.match, w.mismatch, w.gap_open, w.gap_extend))


@pytest.mark.parametrize("config", ["sw_long", "sw_noisy"])
def test_full_batch_every_pair_vs_oracle(config):
    from graphmambaformer_amd import kernels as K
    from graphmambaformer_amd import workloads as WL

    _say(f"case: full-size {config}")
    w = getattr(WL, config)()
    d = np.load(DIGEST)
    assert D.sw_input_digest(w) == int(d[f"{config}_input"][0]), \
        f"{config}: the generator did not reproduce the digest's pairs"
    inp = K.upload_pairs(w.queries, w.targets, w.band_offset, DEV, alphabet="2bit")
    cells = w.cells()
    lpt = torch.from_numpy(np.argsort(-cells, kind="stable").astype(np.int32)).to(DEV)
    a = _run(K, w, inp, lpt)
    _against_digest(config, a, (config, "plain"))
    cols = _properties(w, inp, a, config)
    _say(f"  {config}: {w.n_pairs} pairs, {cells.sum():.3e} cells, every pair equals the "
         f"oracle, {cols} aligned columns ok")
    # the standalone walk over the same planes (sw_long: a lane per pair, two
    # pairs per wavefront; sw_noisy: a wavefront per pair) equals the fused walk
    _say(f"  {config}: standalone walk vs fused walk")
    sep = K.sw_traceback(inp, a)
    torch.cuda.synchronize()
    for x, y in ((a.walk.n_cigar, sep.n_cigar), (a.walk.read_start, sep.read_start),
                 (a.walk.ref_start, sep.ref_start), (a.walk.counts, sep.counts)):
        assert torch.equal(x, y), config
    assert K.decode_cigars(a.walk) == K.decode_cigars(sep), config
    del sep
    ws = {"tb": a.tb, "tb_off": a.tb_off}
    b = _run(K, w, inp, None, workspace=ws)
    _against_digest(config, b, (config, "natural order"))
    del b
    if config == "sw_noisy":
        for slices in (2, 3, 4):
            _say(f"  {config}: split launch, {slices} slices")
            c = _run(K, w, inp, lpt, workspace=ws, split_rows=40000, split_slices=slices)
            _against_digest(config, c, (config, "split", slices))
            del c


def test_full_chaining_config_vs_oracle():
    """The chaining configuration at full size (4,096 reads, ~Poisson(400)
    anchors, lookback 50, max_gap 5,000): every read's f and parent, integer
    profile and the reference-default float64 profile, equal the oracle's."""
    from graphmambaformer_amd import kernels as K
    from graphmambaformer_amd import workloads as WL

    _say("case: full-size chaining")
    w = WL.chains()
    d = torch.device(DEV)
    re_, fe_ = torch.from_numpy(w.read_end).to(d), torch.from_numpy(w.ref_end).to(d)
    so = torch.from_numpy(w.seg_off).to(d)
    f32, p32 = K.chain_dp(re_, fe_, torch.from_numpy(w.weight).to(d), so, w.max_lookback,
                          w.max_gap, gap_open=w.gap_open, gap_extend=w.gap_extend)
    lut_h = O.penalty_lut(6.0, 0.05, 0.5, w.max_gap)
    w64 = w.weight.astype(np.float64)
    f64, p64 = K.chain_dp(re_, fe_, torch.from_numpy(w64).to(d), so, w.max_lookback, w.max_gap,
                          pen_lut=torch.from_numpy(lut_h).to(d))
    f32, p32 = f32.cpu().numpy(), p32.cpu().numpy()
    f64, p64 = f64.cpu().numpy(), p64.cpu().numpy()
    for s in range(w.n_reads):
        a, b = w.seg_off[s], w.seg_off[s + 1]
        ef, ep = O.chain_dp_i32(w.read_end[a:b], w.ref_end[a:b], w.weight[a:b], w.max_lookback,
                                w.max_gap, w.gap_open, w.gap_extend)
        assert np.array_equal(f32[a:b], ef) and np.array_equal(p32[a:b], ep), s
        ef, ep = O.chain_dp_f64(w.read_end[a:b], w.ref_end[a:b], w64[a:b], w.max_lookback,
                                w.max_gap, lut_h)
        assert np.array_equal(f64[a:b], ef) and np.array_equal(p64[a:b], ep), ("f64", s)
    _say(f"  chaining: {w.n_reads} reads, {len(w.read_end)} anchors ok")


def test_full_chaining_anchor_sets_vs_oracle():
    """Chain anchor sets at the chaining configuration's full size (4,096 reads,
    ~1.64 M anchors, unsorted on the device): the device pipeline (sort into
    DP order, float64 DP with the reference defaults 6 / 0.05 / 0.5, claimed-
    walk backtrack, primary/secondary select) against the oracle's
    AffineChainer.chain per read (digest over every kept chain's score and
    members, in order)."""
    from graphmambaformer_amd import kernels as K
    from graphmambaformer_amd import workloads as WL
    from graphmambaformer_amd.alignment.chaining import penalty_table
    from graphmambaformer_amd.config import ChainingConfig

    _say("case: full-size chain anchor sets")
    w = WL.chains()
    d = np.load(DIGEST)
    assert D.chains_input_digest(w) == int(d["chains_input"][0])
    cfg = ChainingConfig(max_lookback=w.max_lookback, max_gap=w.max_gap)
    n = w.n_reads
    rp, fp, ids = D.chains_input(w)  # every read's anchors shuffled
    read_seg = np.stack([np.arange(n), np.full(n, -1)], axis=1).astype(np.int32)
    res = K.chain_pipeline(rp + 15, fp + 15, rp, fp, np.full(len(rp), 15.0), ids, w.seg_off,
                           np.ones(n, np.int32), read_seg, w.seg_off[:-1], cfg.max_lookback,
                           cfg.max_gap, penalty_table(cfg, torch.device(DEV)),
                           cfg.min_chain_anchors, cfg.min_chain_score, cfg.secondary_score_ratio,
                           cfg.secondary_overlap, cfg.max_chains, DEV)
    assert np.array_equal(res.n_kept, d["chains_nkept"])
    for r in range(n):
        chains = []
        for x in range(int(res.n_kept[r])):
            rec = int(res.kept[r, x])
            o, c = res.mem[rec]
            chains.append((float(res.score[rec]), res.members[o:o + c]))
        assert D.chain_digest(chains) == int(d["chains_digest"][r]), r
        if chains:
            assert chains[0][0] == float(d["chains_best"][r]), r
    _say(f"  chain anchor sets: {n} reads, {int(res.n_kept.sum())} kept chains equal the oracle")

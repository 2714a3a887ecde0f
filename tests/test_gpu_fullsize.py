"""The long-read (16,384 pairs, half_band 256) and noisy (8,192 pairs, half_band
751) batches at their full sizes on the GPU, checked through properties that
hold at any size, over every pair:

* each CIGAR consumes exactly the query span [read_start, query_end) and the
  window span [ref_start, target_end) it reports, and its op counts match;
* every '=' / 'X' column agrees with the actual bases (checked on the device);
* rescoring each CIGAR with the affine scheme (a gap of length L costs
  open + (L-1) * extend, extension.py:154-156) gives exactly the reported
  score, and a positive-score local alignment begins and ends with '=';
* the results do not depend on the work order (longest-first vs natural);

plus a stratified sample of pairs checked bit-for-bit against the oracle."""

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _say(*a):
    print(*a, flush=True)


def _run(K, w, inp, order):
    fwd = K.sw_forward(inp, w.match, w.mismatch, w.gap_open, w.gap_extend, w.half_band,
                       traceback=True, order=order, cigar=True)
    torch.cuda.synchronize()
    return fwd


def _properties(w, inp, fwd, label):
    n = w.n_pairs
    tbk = fwd.walk
    score = fwd.score.long()
    qe, te = fwd.query_end.long(), fwd.target_end.long()
    rs, ts = tbk.read_start.long(), tbk.ref_start.long()
    nc = tbk.n_cigar.long()
    assert (score >= 0).all() and (nc >= 0).all(), label
    total = int(nc.sum())
    pair = torch.arange(n, device=DEV)
    owner = torch.repeat_interleave(pair, nc)
    first = torch.cumsum(nc, 0) - nc
    within = torch.arange(total, device=DEV) - first[owner]
    runs = tbk.cigar[tbk.cigar_off.long()[owner] + within].long() & 0xFFFFFFFF
    op, ln = runs & 15, runs >> 4
    assert (ln >= 1).all(), label
    assert torch.isin(op, torch.tensor([1, 2, 7, 8], device=DEV)).all(), label
    is_eq, is_x, is_i, is_d = op == 7, op == 8, op == 1, op == 2
    zero = torch.zeros(n, dtype=torch.long, device=DEV)

    def per_pair(v):
        return zero.clone().index_add_(0, owner, v)

    q_cons = torch.where(is_eq | is_x | is_i, ln, 0)
    t_cons = torch.where(is_eq | is_x | is_d, ln, 0)
    assert torch.equal(per_pair(q_cons), qe - rs), label
    assert torch.equal(per_pair(t_cons), te - ts), label
    counts = tbk.counts.long()
    for k, m in enumerate((is_eq, is_x, is_i, is_d)):
        assert torch.equal(per_pair(torch.where(m, ln, 0)), counts[:, k]), (label, k)
    # rescoring
    ma, mi, go, ge = w.match, w.mismatch, w.gap_open, w.gap_extend
    contrib = torch.where(is_eq, ma * ln, torch.where(is_x, -mi * ln,
                                                      -(go + (ln - 1) * ge)))
    assert torch.equal(per_pair(contrib), score), label
    pos = score > 0
    last = first + nc - 1
    assert (op[first[pos]] == 7).all() and (op[last[pos]] == 7).all(), label
    assert (nc[~pos] == 0).all(), label
    # '=' / 'X' columns against the bases
    gq = torch.cumsum(q_cons, 0) - q_cons
    gt = torch.cumsum(t_cons, 0) - t_cons
    q_run = gq - gq[first[owner]]      # query offset of each run inside its alignment
    t_run = gt - gt[first[owner]]
    ql = torch.tensor([len(q) for q in w.queries], dtype=torch.long, device=DEV)
    tl = torch.tensor([len(t) for t in w.targets], dtype=torch.long, device=DEV)
    q_base = torch.cumsum(ql, 0) - ql
    t_base = torch.cumsum(tl, 0) - tl
    qcodes = torch.from_numpy(np.concatenate(w.queries).astype(np.uint8)).to(DEV)
    tcodes = torch.from_numpy(np.concatenate(w.targets).astype(np.uint8)).to(DEV)
    diag = torch.nonzero(is_eq | is_x).squeeze(1)
    dl = ln[diag]
    col_run = torch.repeat_interleave(diag, dl)
    col_k = torch.arange(int(dl.sum()), device=DEV) - (torch.cumsum(dl, 0) - dl).repeat_interleave(dl)
    p = owner[col_run]
    qi = q_base[p] + rs[p] + q_run[col_run] + col_k
    ti = t_base[p] + ts[p] + t_run[col_run] + col_k
    same = qcodes[qi] == tcodes[ti]
    assert torch.equal(same, op[col_run] == 7), label
    return int(dl.sum())


def _oracle_sample(w, fwd, n_sample, label):
    cells = w.cells()
    srt = np.argsort(cells, kind="stable")
    pick = srt[np.linspace(0, len(srt) - 1, n_sample).astype(int)]
    qs = [w.as_ref_codes(int(k))[0] for k in pick]
    ts = [w.as_ref_codes(int(k))[1] for k in pick]
    exp = O.sw_align_pairs(qs, ts, np.asarray(w.band_offset)[pick].astype(np.int64), w.half_band,
                           w.match, w.mismatch, w.gap_open, w.gap_extend, threads=16)
    from graphmambaformer_amd import kernels as K

    cig = K.decode_cigars(fwd.walk)
    score = fwd.score.cpu().numpy()
    qe, te = fwd.query_end.cpu().numpy(), fwd.target_end.cpu().numpy()
    rs, ts_ = fwd.walk.read_start.cpu().numpy(), fwd.walk.ref_start.cpu().numpy()
    for e, k in zip(exp, pick):
        ctx = (label, int(k), int(cells[k]))
        assert int(score[k]) == e["score"] and cig[k] == e["cigar"], ctx
        assert int(qe[k]) == e["read_end"] and int(te[k]) == e["ref_end"], ctx
        assert int(rs[k]) == e["read_start"] and int(ts_[k]) == e["ref_start"], ctx


@pytest.mark.parametrize("config", ["sw_long", "sw_noisy"])
def test_full_batch_properties_and_oracle_sample(config):
    from graphmambaformer_amd import kernels as K
    from graphmambaformer_amd import workloads as WL

    _say(f"case: full-size {config}")
    w = getattr(WL, config)()
    inp = K.upload_pairs(w.queries, w.targets, w.band_offset, DEV, alphabet="2bit")
    cells = w.cells()
    lpt = torch.from_numpy(np.argsort(-cells, kind="stable").astype(np.int32)).to(DEV)
    a = _run(K, w, inp, lpt)
    cols = _properties(w, inp, a, config)
    _say(f"  {config}: {w.n_pairs} pairs, {cells.sum():.3e} cells, {cols} aligned columns ok")
    b = _run(K, w, inp, None)
    for x, y in ((a.score, b.score), (a.query_end, b.query_end), (a.target_end, b.target_end),
                 (a.walk.n_cigar, b.walk.n_cigar), (a.walk.read_start, b.walk.read_start),
                 (a.walk.ref_start, b.walk.ref_start), (a.walk.counts, b.walk.counts)):
        assert torch.equal(x, y), config
    assert K.decode_cigars(a.walk) == K.decode_cigars(b.walk), config
    del b
    # the standalone walk over the same planes (a lane per pair, several
    # pairs per wavefront at these batch sizes) equals the fused walk
    _say(f"  {config}: standalone walk vs fused walk")
    sep = K.sw_traceback(inp, a)
    torch.cuda.synchronize()
    for x, y in ((a.walk.n_cigar, sep.n_cigar), (a.walk.read_start, sep.read_start),
                 (a.walk.ref_start, sep.ref_start), (a.walk.counts, sep.counts)):
        assert torch.equal(x, y), config
    assert K.decode_cigars(a.walk) == K.decode_cigars(sep), config
    _oracle_sample(w, a, 32, config)


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

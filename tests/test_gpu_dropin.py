"""GPU tests of the drop-in API (same signatures as the reference's
extension.py / chaining.py / accel/cuda_kernels.py), against the reference's
own golden vectors and known answers."""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.golden.load import fhex, load

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _say(*a):
    print(*a, flush=True)


def test_kat_tier_functions():
    """tests/test_accel.py:288-371 of the reference, on the HIP tier."""
    from graphmambaformer_amd.accel import banded_sw, chain_dp
    from graphmambaformer_amd.alignment import encode_bases

    _say("case: tier KATs")
    q = torch.as_tensor(encode_bases("ACGT"), dtype=torch.int8, device=DEV)[None]
    t = torch.as_tensor(encode_bases("TTACGTAA"), dtype=torch.int8, device=DEV)[None]
    s, qe, te = banded_sw(q, t, torch.tensor([4], device=DEV), torch.tensor([8], device=DEV),
                          half_band=2, match_score=2.0, mismatch_penalty=4.0, gap_open=6.0,
                          gap_extend=2.0, band_offset=torch.tensor([2], device=DEV))
    assert float(s[0]) == 8.0 and int(qe[0]) == 4 and int(te[0]) == 6
    anchors = torch.tensor([[[10, 10, 0], [20, 20, 0], [30, 31, 0]]], dtype=torch.int32,
                           device=DEV)
    f, p = chain_dp(anchors, torch.tensor([3], dtype=torch.int32, device=DEV), lookback=64,
                    max_gap=5000, gap_open=6.0, gap_extend=0.05, log_coeff=0.5,
                    weights=torch.tensor([[10.5, 9.25, 8.75]], dtype=torch.float32, device=DEV))
    assert np.allclose(f[0].cpu().numpy(), [10.5, 19.75, 21.95], atol=1e-5)
    assert p[0].cpu().numpy().tolist() == [-1, 0, 1]


def test_banded_sw_drop_in_identical_and_full_band():
    """test_alignment_stages.py:999-1048 of the reference on the drop-in."""
    from graphmambaformer_amd.alignment import banded_affine_sw_batch, traceback_banded
    from graphmambaformer_amd.config import ExtensionConfig
    from oracle import brute

    _say("case: identical / full band")
    cfg = ExtensionConfig()
    rng = np.random.default_rng(3)
    seq = rng.integers(1, 5, size=40).astype(np.int16)
    r = banded_affine_sw_batch(torch.as_tensor(seq, device=DEV)[None],
                               torch.as_tensor(seq, device=DEV)[None], torch.tensor([40]),
                               torch.tensor([40]), cfg, half_band=8, return_matrices=True)
    assert float(r.score[0]) == 80.0
    ext = traceback_banded(r, 0, seq, seq, cfg)
    assert ext.cigar == [("=", 40)] and ext.edit_distance == 0
    for trial in range(6):
        n = int(rng.integers(20, 45))
        a = rng.integers(1, 5, size=n).astype(np.int16)
        b = a.copy()
        for _ in range(2):
            b[int(rng.integers(0, n))] = int(rng.integers(1, 5))
        r = banded_affine_sw_batch(torch.as_tensor(a, device=DEV)[None],
                                   torch.as_tensor(b, device=DEV)[None], torch.tensor([n]),
                                   torch.tensor([n]), cfg, half_band=n)
        assert float(r.score[0]) == brute.full_gotoh_score(a, b, 2, 4, 6, 2), trial


def test_banded_sw_drop_in_golden_ties_xdrop():
    from graphmambaformer_amd.alignment import banded_affine_sw_batch, traceback_banded
    from graphmambaformer_amd.config import ExtensionConfig

    for n, case in enumerate(load("sw_ties.json")):
        _say("case: drop-in golden", n)
        ma, mi, go, ge = case["scoring"]
        cfg = ExtensionConfig(match_score=float(ma), mismatch_penalty=float(mi),
                              gap_open=float(go), gap_extend=float(ge), x_drop=case["x_drop"])
        q = torch.tensor(case["query"], dtype=torch.int16, device=DEV)
        t = torch.tensor(case["target"], dtype=torch.int16, device=DEV)
        r = banded_affine_sw_batch(q, t, torch.tensor(case["qlen"]), torch.tensor(case["tlen"]),
                                   cfg, half_band=case["half_band"],
                                   band_offset=torch.tensor(case["off"]), return_matrices=True)
        for b, exp in enumerate(case["out"]):
            assert float(r.score[b]) == exp["score"], (n, b)
            assert int(r.query_end[b]) == exp["query_end"], (n, b)
            assert int(r.target_end[b]) == exp["target_end"], (n, b)
            ext = traceback_banded(r, b, None, None, cfg)
            assert [list(x) for x in ext.cigar] == [list(x) for x in exp["cigar"]], (n, b)
            assert ext.read_start == exp["read_start"] and ext.ref_start == exp["ref_start"]


def test_extension_engine_golden():
    from graphmambaformer_amd.alignment import AnchorSet, Chain, ExtensionEngine
    from graphmambaformer_amd.config import ExtensionConfig

    for n, case in enumerate(load("extend.json")):
        _say("case: ExtensionEngine golden", n)
        a = case["anchors"]
        anchors = AnchorSet.from_lists(a["read_pos"], a["ref_pos"], a["length"], a["strand"],
                                       read_len=len(case["read"]), ref_len=len(case["ref"]))
        chains = [Chain(anchor_idx=np.array(c["anchor_idx"]), score=50.0, strand=c["strand"],
                        read_start=c["read_start"], read_end=c["read_end"],
                        ref_start=c["ref_start"], ref_end=c["ref_end"]) for c in case["chains"]]
        cfg = ExtensionConfig(half_band=case["cfg"]["half_band"], flank=case["cfg"]["flank"],
                              x_drop=case["cfg"]["x_drop"])
        out = ExtensionEngine(cfg, device=DEV).extend_chains(case["read"], chains, anchors,
                                                             case["ref"])
        for got, exp in zip(out, case["out"]):
            assert got.score == exp["score"], n
            assert [list(x) for x in got.cigar] == [list(x) for x in exp["cigar"]], n
            assert (got.read_start, got.read_end, got.ref_start, got.ref_end) == (
                exp["read_start"], exp["read_end"], exp["ref_start"], exp["ref_end"]), n


def test_chain_dp_numpy_and_batched_golden():
    from graphmambaformer_amd.alignment import chain_dp_batched, chain_dp_numpy
    from graphmambaformer_amd.config import ChainingConfig

    data = load("chain_dp.json")
    _say("case: chain_dp_numpy golden")
    for n, case in enumerate(data["cases"]):
        cfg = ChainingConfig(max_lookback=case["max_lookback"], max_gap=case["max_gap"])
        f, p = chain_dp_numpy(np.array(case["read_end"]), np.array(case["ref_end"]),
                              fhex(case["weight"]), cfg)
        assert np.array_equal(f, fhex(case["f"])), n
        assert p.tolist() == case["parent"], n
    # ragged batch (test_alignment_stages.py:502-517): rows equal the per-read DP
    cfg = ChainingConfig(max_lookback=4)
    re_ = torch.tensor([[10, 20, 30], [8, 18, 0]], dtype=torch.float32)
    fe = torch.tensor([[10, 21, 32], [8, 20, 0]], dtype=torch.float32)
    w = torch.tensor([[10.5, 9.25, 8.75], [7.5, 8.25, 0.0]])
    f, p = chain_dp_batched(re_, fe, w, torch.tensor([3, 2]), cfg)
    assert np.allclose(f[0].numpy(), [10.5, 13.2, 15.4]) and p[0].tolist() == [-1, 0, 1]
    assert np.isclose(float(f[1, 1]), 7.5 + 8.25 - (6.1 + 0.5 * np.log2(3.0)))
    assert p[1].tolist() == [-1, 0, -1] and float(f[1, 2]) == 0.0


def test_chain_dp_batched_rows_vs_oracle():
    """A padded (B, A) batch on the device (the accelerator tier's shape,
    chaining.py:183-258): every row equals the oracle's float64 DP on that
    read (f rounded to float32), padded slots 0 / -1, with and without a
    (B, A, L) graph bonus."""
    from graphmambaformer_amd import workloads
    from graphmambaformer_amd.alignment import chain_dp_batched
    from graphmambaformer_amd.config import ChainingConfig

    _say("case: chain_dp_batched rows vs oracle")
    w = workloads.chains(n_reads=40, seed=9)
    cfg = ChainingConfig(max_lookback=64, max_gap=5000)
    B = w.n_reads
    cnt = np.diff(w.seg_off)
    A = int(cnt.max()) + 3
    rng = np.random.default_rng(9)
    wt = rng.integers(5, 30, size=len(w.read_end)).astype(np.float64)
    re_ = np.zeros((B, A), np.float32)
    fe = np.zeros((B, A), np.float64)
    ww = np.zeros((B, A), np.float64)
    for b in range(B):
        a, z = w.seg_off[b], w.seg_off[b + 1]
        re_[b, :z - a] = w.read_end[a:z]
        fe[b, :z - a] = w.ref_end[a:z]
        ww[b, :z - a] = wt[a:z]
    bonus = rng.normal(0.0, 1.5, size=(B, A, cfg.max_lookback))
    lut = O.penalty_lut(cfg.gap_open, cfg.gap_extend, cfg.log_coeff, cfg.max_gap)
    for use_bonus in (False, True):
        f, p = chain_dp_batched(torch.from_numpy(re_).to(DEV), torch.from_numpy(fe).to(DEV),
                                torch.from_numpy(ww).to(DEV), torch.from_numpy(cnt).to(DEV), cfg,
                                bonus=torch.from_numpy(bonus).to(DEV) if use_bonus else None)
        f, p = f.cpu().numpy(), p.cpu().numpy()
        for b in range(B):
            a, z = w.seg_off[b], w.seg_off[b + 1]
            ef, ep = O.chain_dp_f64(w.read_end[a:z], w.ref_end[a:z], wt[a:z], cfg.max_lookback,
                                    cfg.max_gap, lut,
                                    bonus=bonus[b, :z - a] if use_bonus else None)
            assert np.array_equal(f[b, :z - a], ef.astype(np.float32)), (use_bonus, b)
            assert np.array_equal(p[b, :z - a], ep), (use_bonus, b)
            assert not f[b, z - a:].any() and (p[b, z - a:] == -1).all(), (use_bonus, b)


def test_affine_chainer_golden():
    from graphmambaformer_amd.alignment import AffineChainer, AnchorSet
    from graphmambaformer_amd.config import ChainingConfig

    cases = load("chainer.json")
    _say("case: AffineChainer golden", len(cases))
    sets, cfgs = [], []
    for case in cases:
        sets.append(AnchorSet.from_lists(case["read_pos"], case["ref_pos"], case["length"],
                                         case["strand"], read_len=case["read_len"]))
        cfgs.append(ChainingConfig(max_lookback=case["max_lookback"],
                                   min_chain_score=case["min_chain_score"]))
    for n, (case, anchors, cfg) in enumerate(zip(cases, sets, cfgs)):
        got = AffineChainer(cfg).chain(anchors)
        exp = case["chains"]
        assert len(got) == len(exp), n
        for g, e in zip(got, exp):
            assert g.anchor_idx.tolist() == e["anchor_idx"], n
            assert g.score == float.fromhex(e["score"]), n
            assert (g.strand, g.read_start, g.read_end, g.ref_start, g.ref_end, g.is_primary) == (
                e["strand"], e["read_start"], e["read_end"], e["ref_start"], e["ref_end"],
                e["is_primary"]), n
    # a batch of reads sharing one launch gives the same chains as read by read
    same = [c for c in cfgs if c.max_lookback == 64 and c.min_chain_score == 20.0]
    idx = [k for k, c in enumerate(cfgs) if c in same]
    batch = AffineChainer(cfgs[idx[0]]).chain_batch([sets[k] for k in idx])
    for k, chains in zip(idx, batch):
        single = AffineChainer(cfgs[k]).chain(sets[k])
        assert [(c.anchor_idx.tolist(), c.score) for c in chains] == \
            [(c.anchor_idx.tolist(), c.score) for c in single]


def test_invalid_inputs_raise():
    from graphmambaformer_amd.alignment import banded_affine_sw_batch
    from graphmambaformer_amd.config import ExtensionConfig

    _say("case: invalid inputs")
    q = torch.tensor([[1, 2, 0, 4]], dtype=torch.int16, device=DEV)  # code 0 is not a base
    with pytest.raises(ValueError):
        banded_affine_sw_batch(q, q, torch.tensor([4]), torch.tensor([4]), ExtensionConfig(),
                               half_band=2)
    q = torch.tensor([[1, 2, 3, 4]], dtype=torch.int16, device=DEV)
    with pytest.raises(ValueError):
        banded_affine_sw_batch(q, q, torch.tensor([4]), torch.tensor([4]),
                               ExtensionConfig(match_score=1.5), half_band=2)


def _batch_case(seed, n_reads, x_drop, with_n):
    """Reads sampled from one reference (both strands, decoy chains, anchor
    spreads that give different half-bands per read)."""
    from graphmambaformer_amd import workloads as WL
    from graphmambaformer_amd.alignment import AnchorSet, Chain

    rng = np.random.default_rng(seed)
    ref = rng.integers(0, 4, size=300_000).astype(np.uint8)
    if with_n:
        ref[1000:1040] = 4
    alpha = np.array(list("ACGTN"))
    comp = np.array([3, 2, 1, 0, 4])
    ref_s = "".join(alpha[ref])
    reads, chains, anchors, ora = [], [], [], []
    for r in range(n_reads):
        L = int(rng.integers(150, 2500))
        pos = int(rng.integers(0, len(ref) - L - 10)) if r % 7 else 900
        seg = WL.mutate(rng, np.minimum(ref[pos:pos + L], 3), 0.03, 0.5, 3)
        strand = 1 if r % 2 == 0 else -1
        fwd_codes = seg if strand > 0 else comp[seg[::-1]]
        reads.append("".join(alpha[fwd_codes]))
        spread = int(rng.integers(0, 30))
        dg = np.array([pos, pos + spread, pos + spread // 2], np.int64)
        a = AnchorSet.from_lists([0, 50, 100], (dg + [0, 50, 100]).tolist(), [15, 15, 15],
                                 [strand] * 3, read_len=len(seg), ref_len=len(ref))
        cs = [dict(strand=strand, read_start=0, ref_start=pos, anchor_idx=[0, 1, 2])]
        if r % 3 == 0:  # a decoy chain somewhere else, other strand
            d = int(rng.integers(0, len(ref) - L))
            cs.append(dict(strand=-strand, read_start=10, ref_start=d, anchor_idx=[2]))
        chains.append([Chain(anchor_idx=np.array(c["anchor_idx"]), score=30.0,
                             strand=c["strand"], read_start=c["read_start"], read_end=L,
                             ref_start=c["ref_start"], ref_end=c["ref_start"] + L) for c in cs])
        anchors.append(a)
        ora.append((cs, a.diagonal))
    return ref_s, reads, chains, anchors, ora


@pytest.mark.parametrize("x_drop,with_n", [(600.0, False), (25.0, True), (0.0, True)])
def test_extension_engine_batch_vs_oracle(x_drop, with_n):
    """extend_batch over many reads (one launch per half-band, per-read x-drop
    on the device, windows as offsets into the resident reference) equals the
    oracle's per-read extend_chains."""
    from graphmambaformer_amd.alignment import ExtensionEngine
    from graphmambaformer_amd.config import ExtensionConfig

    _say("case: extend_batch", x_drop, with_n)
    ref_s, reads, chains, anchors, ora = _batch_case(int(x_drop) + 5, 36, x_drop, with_n)
    cfg = ExtensionConfig(half_band=16, max_half_band=40, flank=60, x_drop=x_drop)
    eng = ExtensionEngine(cfg, device=DEV)
    got = eng.extend_batch(reads, chains, anchors, ref_s)
    for r, (cs, diag) in enumerate(ora):
        exp = O.extend_chains(reads[r], cs, diag, ref_s, half_band=16, max_half_band=40,
                              flank=60, max_window=cfg.max_window, ma=2, mi=4, go=6, ge=2,
                              x_drop=x_drop)
        assert len(got[r]) == len(exp), r
        for g, e in zip(got[r], exp):
            ctx = (r, x_drop)
            assert g.score == e["score"], ctx
            assert [list(x) for x in g.cigar] == [list(x) for x in e["cigar"]], ctx
            assert (g.read_start, g.read_end, g.ref_start, g.ref_end) == (
                e["read_start"], e["read_end"], e["ref_start"], e["ref_end"]), ctx
    # single-read calls agree with the batch
    one = eng.extend_chains(reads[3], chains[3], anchors[3], ref_s)
    assert [(x.score, x.cigar) for x in one] == [(x.score, x.cigar) for x in got[3]]


def _anchor_read(rng, n_anchors, read_len=8000, with_ties=False):
    """Anchors of one read: collinear runs on both strands, decoys, and (with
    ties) duplicated anchors and equal spacings so equal f values occur."""
    rp, fp, ln, st = [], [], [], []
    for strand in (1, -1):
        for _ in range(int(rng.integers(1, 4))):
            start = int(rng.integers(0, 40_000_000))
            k = int(rng.integers(2, max(3, n_anchors // 3)))
            pos = np.sort(rng.integers(0, read_len - 20, size=k))
            jitter = rng.integers(-2, 3, size=k) if not with_ties else np.zeros(k, np.int64)
            rp += pos.tolist()
            fp += (start + pos + jitter).tolist()
            ln += [15] * k
            st += [strand] * k
    nd = max(1, n_anchors // 10)
    rp += rng.integers(0, read_len - 20, size=nd).tolist()
    fp += rng.integers(0, 40_000_000, size=nd).tolist()
    ln += rng.integers(11, 20, size=nd).tolist()
    st += rng.choice([1, -1], size=nd).tolist()
    if with_ties and len(rp) > 4:
        for x in rng.integers(0, len(rp), size=4):
            rp.append(rp[x]), fp.append(fp[x]), ln.append(ln[x]), st.append(st[x])
    return (np.array(rp), np.array(fp), np.array(ln), np.array(st))


@pytest.mark.parametrize("min_score,lookback,ratio,max_chains",
                         [(20.0, 64, 0.6, 8), (0.0, 17, 0.0, 64), (200.0, 50, 0.6, 8),
                          (10.0, 33, 0.1, 3), (20.0, 200, 0.6, 8)])
def test_affine_chainer_device_pipeline_vs_oracle(min_score, lookback, ratio, max_chains):
    """Sort, DP, backtrack and select all on the device for a batch of reads
    (ragged, both strands, ties, one read with 20,000 anchors) equal the
    oracle's per-read chaining."""
    from graphmambaformer_amd.alignment import AffineChainer, AnchorSet
    from graphmambaformer_amd.config import ChainingConfig

    _say("case: device chaining pipeline", min_score, lookback)
    rng = np.random.default_rng(int(min_score) + lookback)
    reads = [_anchor_read(rng, int(rng.integers(1, 400)), with_ties=(r % 3 == 0))
             for r in range(120)]
    reads.append(_anchor_read(rng, 20_000))
    reads.append((np.zeros(0, np.int64),) * 4)
    cfg = ChainingConfig(max_lookback=lookback, min_chain_score=min_score,
                         secondary_score_ratio=ratio, max_chains=max_chains)
    sets = [AnchorSet.from_lists(rp, fp, ln, st, read_len=8000) for rp, fp, ln, st in reads]
    got = AffineChainer(cfg).chain_batch(sets)
    for r, (rp, fp, ln, st) in enumerate(reads):
        exp = O.chain_read(rp, fp, ln, st, max_lookback=lookback, min_chain_score=min_score,
                           secondary_score_ratio=ratio, max_chains=max_chains)
        assert len(got[r]) == len(exp), r
        for g, e in zip(got[r], exp):
            assert g.anchor_idx.tolist() == e["anchor_idx"].tolist(), r
            assert g.score == e["score"], r
            assert (g.strand, g.read_start, g.read_end, g.ref_start, g.ref_end, g.is_primary) == (
                e["strand"], e["read_start"], e["read_end"], e["ref_start"], e["ref_end"],
                e["is_primary"]), r


def test_tier_kmer_lookup_and_wfa_distance():
    """cuda_kernels.kmer_lookup / wfa_distance signatures: the KAT of
    tests/test_accel.py:336-345 (ACGTAC vs ACGTTAC -> 1), random pairs vs the
    oracle, and table lookups vs the oracle's KmerTable."""
    from graphmambaformer_amd.accel import kmer_lookup, wfa_distance
    from oracle import seeding as S
    from oracle import wfa as W

    _say("case: tier kmer_lookup / wfa_distance")
    q = torch.tensor(O.encode("ACGTAC"), dtype=torch.int8, device=DEV)[None]
    t = torch.tensor(O.encode("ACGTTAC"), dtype=torch.int8, device=DEV)[None]
    d = wfa_distance(q, t, torch.tensor([6], device=DEV), torch.tensor([7], device=DEV), 4)
    assert int(d[0]) == 1
    rng = np.random.default_rng(21)
    qs = [rng.integers(1, 6, size=int(rng.integers(0, 80))) for _ in range(40)]
    ts = [np.concatenate([x[:len(x) // 2], rng.integers(1, 5, size=int(rng.integers(0, 6))),
                          x[len(x) // 2:]]) for x in qs]
    M = max(len(x) for x in qs) + 1
    N = max(len(x) for x in ts) + 1
    Q = np.full((40, M), -1, np.int8)
    T = np.full((40, N), -1, np.int8)
    for b in range(40):
        Q[b, :len(qs[b])] = qs[b]
        T[b, :len(ts[b])] = ts[b]
    d = wfa_distance(torch.from_numpy(Q).to(DEV), torch.from_numpy(T).to(DEV),
                     torch.tensor([len(x) for x in qs]), torch.tensor([len(x) for x in ts]), 3)
    for b in range(40):
        e = W.wfa_align(qs[b], ts[b], 3)
        assert int(d[b]) == (-1 if e is None else e[0]), b
    ref = rng.integers(1, 5, size=20000).astype(np.int8)
    tab = S.build_index(ref, 13, 6, 0)
    qk = np.concatenate([tab.keys[::3], rng.integers(0, 1 << 26, size=500)]).astype(np.int64)
    st, ct = kmer_lookup(torch.from_numpy(tab.keys).to(DEV),
                         torch.from_numpy(tab.offsets.astype(np.int32)).to(DEV),
                         torch.from_numpy(qk).to(DEV))
    es, ec = tab.lookup(qk)
    assert np.array_equal(st.cpu().numpy(), es.astype(np.int32))
    assert np.array_equal(ct.cpu().numpy(), ec.astype(np.int32))

"""GPU parity of the batched wavefront aligner (gmfa_wfa_distance / align)
and of ExtensionEngine's WFA branch against the oracle and the reference's
golden vectors, exact."""

import numpy as np
import pytest

from oracle import oracle as O
from oracle import wfa as W
from tests.golden.load import load

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _say(*a):
    print(*a, flush=True)


def test_wfa_golden_align():
    from graphmambaformer_amd.alignment import WavefrontAligner
    from graphmambaformer_amd.config import ExtensionConfig

    cases = load("wfa.json")["align"]
    _say("case: wfa golden", len(cases))
    for maxd in sorted({c["max_distance"] for c in cases}):
        sub = [c for c in cases if c["max_distance"] == maxd]
        wfa = WavefrontAligner(ExtensionConfig(algorithm="wfa", wfa_max_distance=maxd), DEV)
        got = wfa.align_batch([O.encode(c["q"]) for c in sub], [O.encode(c["t"]) for c in sub])
        for n, (c, g) in enumerate(zip(sub, got)):
            if c["distance"] is None:
                assert g is None, (maxd, n)
            else:
                assert g[0] == c["distance"], (maxd, n, c["q"], c["t"])
                assert [list(x) for x in g[1]] == c["cigar"], (maxd, n)
    with pytest.raises(RuntimeError):
        WavefrontAligner(ExtensionConfig(wfa_max_distance=2), DEV).align(
            O.encode("ACGTACGTAC"), O.encode("TTTTTTTTTT"))


def test_wfa_batch_vs_oracle_long():
    from graphmambaformer_amd import workloads as WL
    from graphmambaformer_amd.alignment import WavefrontAligner
    from graphmambaformer_amd.config import ExtensionConfig

    _say("case: wfa batch long")
    rng = np.random.default_rng(9)
    qs, ts = [], []
    for p in range(64):
        L = int(rng.integers(1, 3000))
        t = rng.integers(0, 4, size=L).astype(np.uint8)
        q = WL.mutate(rng, t, [0.0, 0.01, 0.03, 0.1][p % 4], 0.5, 3)
        q1, t1 = q.astype(np.int8) + 1, t.astype(np.int8) + 1
        if p % 5 == 0 and len(q1) > 3:
            q1[rng.integers(0, len(q1), size=2)] = 5
            t1[rng.integers(0, len(t1), size=2)] = 5
        qs.append(q1)
        ts.append(t1)
    maxd = 150
    got = WavefrontAligner(ExtensionConfig(wfa_max_distance=maxd), DEV).align_batch(qs, ts)
    for p, (q, t, g) in enumerate(zip(qs, ts, got)):
        e = W.wfa_align(q, t, maxd)
        if e is None:
            assert g is None, p
        else:
            assert g is not None and g[0] == e[0], p
            assert g[1] == e[1], p


def test_extension_engine_wfa_golden():
    from graphmambaformer_amd.alignment import AnchorSet, Chain, ExtensionEngine
    from graphmambaformer_amd.config import ExtensionConfig

    cases = load("wfa.json")["extend"]
    _say("case: ExtensionEngine wfa golden", len(cases))
    reads, chains, anchors, cfgs = [], [], [], []
    for case in cases:
        c = case["chain"]
        s = c["strand"]
        anchors.append(AnchorSet.from_lists([0, 20], [c["ref_start"], c["ref_start"] + 20],
                                            [15, 15], [s, s], read_len=len(case["read"]),
                                            ref_len=len(case["ref"])))
        chains.append([Chain(anchor_idx=np.array([0, 1]), score=40.0, strand=s,
                             read_start=c["read_start"], read_end=c["read_end"],
                             ref_start=c["ref_start"], ref_end=c["ref_end"])])
        reads.append(case["read"])
    for maxd in sorted({c["max_distance"] for c in cases}):
        idx = [k for k, c in enumerate(cases) if c["max_distance"] == maxd]
        cfg = ExtensionConfig(algorithm="wfa", wfa_max_distance=maxd, half_band=24, flank=40)
        got = ExtensionEngine(cfg, device=DEV).extend_batch(
            [reads[k] for k in idx], [chains[k] for k in idx], [anchors[k] for k in idx],
            cases[0]["ref"])
        for k, g in zip(idx, got):
            e = cases[k]["out"][0]
            assert g[0].score == e["score"], k
            assert [list(x) for x in g[0].cigar] == e["cigar"], k
            assert (g[0].read_start, g[0].read_end, g[0].ref_start, g[0].ref_end) == (
                e["read_start"], e["read_end"], e["ref_start"], e["ref_end"]), k


def test_cigar_bam_records_vs_oracle():
    """gmfa_cigar_bam on random CIGARs (clips, strands, length mismatches)."""
    import torch

    from graphmambaformer_amd import kernels as K
    from oracle import bam as B

    _say("case: BAM CIGAR records")
    rng = np.random.default_rng(12)
    ops = ["=", "X", "I", "D"]
    code = {"=": 7, "X": 8, "I": 1, "D": 2}
    recs = []
    for r in range(500):
        cig = []
        for _ in range(int(rng.integers(0, 8))):
            op = ops[int(rng.integers(0, 4))]
            n = int(rng.integers(1, 20))
            if cig and cig[-1][0] == op:
                cig[-1] = (op, cig[-1][1] + n)
            else:
                cig.append((op, n))
        qcons = sum(n for op, n in cig if op in "=XI")
        rs = int(rng.integers(0, 5))
        re_ = rs + qcons
        L = re_ + int(rng.integers(0, 5))
        if r % 9 == 0 and re_ > 0:
            L = re_ - 1  # read end past the read: the right clip vanishes, length rule fails
        fs = int(rng.integers(0, 1000))
        fe = fs + sum(n for op, n in cig if op in "=XD")
        if r % 13 == 0:
            fe = fs  # empty reference span: unmapped
        recs.append((cig, rs, re_, L, fs, fe, 1 if r % 2 else -1))
    runs, offs, nc = [], [0], []
    for cig, *_ in recs:
        runs += [(n << 4) | code[op] for op, n in cig]
        nc.append(len(cig))
        offs.append(len(runs))
    dev = torch.device(DEV)
    t = lambda a, dt: torch.tensor(a, dtype=dt, device=dev)  # noqa: E731
    out, d_off, out_n, flag = K.cigar_bam(
        t(runs or [0], torch.int64).to(torch.int32), t(offs[:-1], torch.int64),
        t(nc, torch.int32), t([x[1] for x in recs], torch.int32),
        t([x[2] for x in recs], torch.int32), t([x[3] for x in recs], torch.int32),
        t([x[4] for x in recs], torch.int64), t([x[5] for x in recs], torch.int64),
        t([x[6] for x in recs], torch.int32), np.array(nc))
    out = out.cpu().numpy().view(np.uint32)
    d_off, out_n, flag = d_off.cpu().numpy(), out_n.cpu().numpy(), flag.cpu().numpy()
    n_unmapped = 0
    for r, (cig, rs, re_, L, fs, fe, st) in enumerate(recs):
        full = B.record_cigar(cig, rs, re_, L)
        exp_runs, exp_flag = B.bam_record(full, L, fs, fe, st)
        got = out[d_off[r]:d_off[r] + out_n[r]].tolist()
        assert got == exp_runs and int(flag[r]) == exp_flag, r
        n_unmapped += bool(exp_flag & 4)
    assert 0 < n_unmapped < len(recs)

"""Stages 1-3 end to end on the GPU (seeding -> chaining -> extension -> BAM
CIGAR records), restating the checks of the reference's pipeline verification
(scripts/verify_alignment_pipeline.py:167-272): the primary chain lands on the
true locus and covers the read, a chimera 2.5 kb apart joins at max_gap 5000
and splits at 500, chain members are collinear, CIGARs consume the read and
the reference span, identity at 3% errors stays above 85%, an exact substring
aligns with no edit, and WFA distances are exact on the known cases."""

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
BASES = "ACGT"


def _say(*a):
    print(*a, flush=True)


def _mutate(rng, s, rate):
    out = []
    for c in s:
        r = rng.random()
        if r < rate / 3:
            out.append(BASES[int(rng.integers(0, 4))])
        elif r < 2 * rate / 3:
            continue
        elif r < rate:
            out.append(c)
            out.append(BASES[int(rng.integers(0, 4))])
        else:
            out.append(c)
    return "".join(out)


def _anchors_both_strands(index, reads):
    """Forward reads and their reverse complements through one seeding batch;
    anchors of the reverse complement carry strand -1 in its own coordinates,
    as the reference's _seed_one_orientation does (seeding.py:1080-1104)."""
    from graphmambaformer_amd.alignment import AnchorSet, encode_bases, reverse_complement_codes

    codes = []
    for r in reads:
        f = encode_bases(r)
        codes += [f, reverse_complement_codes(f)]
    hits = index.query_batch(codes)
    out = []
    for k, r in enumerate(reads):
        parts = []
        for h, strand in ((hits[2 * k], 1), (hits[2 * k + 1], -1)):
            parts.append((h[0], h[1], h[2], np.full(len(h[0]), strand)))
        rp = np.concatenate([p[0] for p in parts])
        fp = np.concatenate([p[1] for p in parts])
        ln = np.concatenate([p[2] for p in parts])
        st = np.concatenate([p[3] for p in parts])
        out.append(AnchorSet.from_lists(rp, fp, ln, st, read_len=len(r)))
    return out


def test_pipeline_stages_on_the_gpu():
    from graphmambaformer_amd import kernels as K
    from graphmambaformer_amd.alignment import (AffineChainer, ExtensionEngine, GPUKmerIndex,
                                                WavefrontAligner, encode_bases)
    from graphmambaformer_amd.config import ChainingConfig, ExtensionConfig

    _say("case: pipeline stages 1-3")
    rng = np.random.default_rng(7)
    ref = "".join(BASES[int(x)] for x in rng.integers(0, 4, size=6000))
    index = GPUKmerIndex(encode_bases(ref), k=15, window=10, device=DEV)
    chainer = AffineChainer(ChainingConfig())
    read = _mutate(rng, ref[1500:1750], 0.02)
    chimera = ref[300:450] + ref[3000:3150]
    read3 = _mutate(rng, ref[900:1150], 0.03)
    exact = ref[2000:2200]
    rc = "".join({"A": "T", "C": "G", "G": "C", "T": "A"}[c] for c in reversed(ref[4000:4300]))
    reads = [read, chimera, read3, exact, rc]
    anchors = _anchors_both_strands(index, reads)

    chains = chainer.chain_batch(anchors)
    best = chains[0][0]
    assert best.is_primary and abs(best.ref_start - 1500) < 60, best.ref_start
    assert best.coverage(len(read)) > 0.5
    for c in chains[0]:
        pos = anchors[0].read_pos[c.anchor_idx]
        assert np.all(np.diff(pos) >= 0), "chain anchors not in read order"
    joined = chains[1]
    assert len(joined) == 1 and joined[0].read_start < 150 < joined[0].read_end, \
        [(c.read_start, c.read_end) for c in joined]
    split = AffineChainer(ChainingConfig(max_gap=500)).chain(anchors[1])
    assert len(split) >= 2, len(split)
    assert chains[4][0].strand == -1 and abs(chains[4][0].ref_start - 4000) < 60

    ext = ExtensionEngine(ExtensionConfig(), device=DEV)
    res = ext.extend_batch([reads[2], reads[3], reads[4]], [chains[2], chains[3], chains[4]],
                           [anchors[2], anchors[3], anchors[4]], ref)
    b3 = res[0][0]
    assert sum(n for op, n in b3.cigar if op in "=XIS") == len(read3)
    assert sum(n for op, n in b3.cigar if op in "=XD") == b3.ref_end - b3.ref_start
    assert b3.identity > 0.85, b3.identity
    assert res[1][0].edit_distance == 0, res[1][0].cigar
    assert res[2][0].edit_distance == 0 and res[2][0].ref_start == 4000

    # the BAM records of those three: mapped, reverse flag on the last one
    runs, offs, nc = [], [0], []
    code = {"=": 7, "X": 8, "I": 1, "D": 2, "S": 4}
    for r in res:
        e = r[0]
        core = [(op, n) for op, n in e.cigar if op != "S"]
        runs += [(n << 4) | code[op] for op, n in core]
        nc.append(len(core))
        offs.append(len(runs))
    t = lambda a, dt: torch.tensor(a, dtype=dt, device=DEV)  # noqa: E731
    lens = [len(reads[2]), len(reads[3]), len(reads[4])]
    out, d_off, out_n, flag = K.cigar_bam(
        t(runs, torch.int32), t(offs[:-1], torch.int64), t(nc, torch.int32),
        t([r[0].read_start for r in res], torch.int32), t([r[0].read_end for r in res], torch.int32),
        t(lens, torch.int32), t([r[0].ref_start for r in res], torch.int64),
        t([r[0].ref_end for r in res], torch.int64), t([1, 1, -1], torch.int32), np.array(nc))
    flag = flag.cpu().numpy()
    assert flag.tolist() == [0, 0, 0x10], flag

    wfa = WavefrontAligner(ExtensionConfig(algorithm="wfa"), DEV)
    cases = {("ACGTACGTACGT", "ACGTACGTACGT"): 0, ("ACGTACGTACGT", "ACGTTCGTACGT"): 1,
             ("ACGTACGTACGT", "ACGTACGACGT"): 1, ("ACGTACGTACGT", "ACGTACGTTACGT"): 1}
    got = wfa.align_batch([O.encode(a) for a, _ in cases], [O.encode(b) for _, b in cases])
    for ((a, b), d), (dist, cig) in zip(cases.items(), got):
        assert dist == d and sum(n for op, n in cig if op in "=XI") == len(a), (a, b, cig)

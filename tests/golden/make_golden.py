"""Write golden vectors for the parts of the path only the reference can pin.

Run in the build container only (it imports the reference package from
/root/reference, which never travels to a GPU host):

    python tests/golden/make_golden.py

It writes ``tests/golden/*.json``: inputs and the reference's outputs, nothing
else.  Cases target what the brute-force and known-answer checks leave open:

* sw_ties.json      banded_affine_sw_batch + traceback_banded
                    (extension.py:78-302) on tie-rich inputs -- homopolymer
                    and tandem-repeat indels, equal maxima across rows and
                    columns -- and multi-pair batches with the batch-wide
                    x-drop (extension.py:190-193).
* chain_dp.json     chain_dp_numpy (chaining.py:129-180) with the default
                    non-integer profile (6, 0.05, 0.5; float64 + np.log2),
                    plus the penalty LUT as computed here.
* chainer.json      AffineChainer.chain (chaining.py:620-642): strand split,
                    DP order, _backtrack and select, incl. equal-f ties.
* extend.json       ExtensionEngine.extend_chains (extension.py:511-641):
                    per-read shared half-band geometry, soft clips, lifts.
* wfa.json          WavefrontAligner.align (extension.py:308-463) incl. N
                    bases, empty sequences, ties and the over-budget signal,
                    and ExtensionEngine's WFA branch (extension.py:643-730).
* seeding.json      hash64 / pack_kmers / minimizer_mask (seeding.py:83-199)
                    and whole GPUKmerIndex.query outputs (seeding.py:720-798,
                    its portable lookup path) incl. N bases, repeats, max_occ.

    python tests/golden/make_golden.py [file.json ...]   # default: all files
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def _imports():
    sys.path.insert(0, REF)
    import torch  # noqa: F401
    from graphmambaformer.alignment.chaining import AffineChainer, chain_dp_numpy
    from graphmambaformer.alignment.extension import (
        ExtensionEngine,
        banded_affine_sw_batch,
        traceback_banded,
    )
    from graphmambaformer.alignment.types import AnchorSet, Chain
    from graphmambaformer.config import ChainingConfig, ExtensionConfig

    return dict(AffineChainer=AffineChainer, chain_dp_numpy=chain_dp_numpy,
                ExtensionEngine=ExtensionEngine, banded_affine_sw_batch=banded_affine_sw_batch,
                traceback_banded=traceback_banded, AnchorSet=AnchorSet, Chain=Chain,
                ChainingConfig=ChainingConfig, ExtensionConfig=ExtensionConfig)


def _tie_rich_pair(rng, length):
    """Read built from homopolymer runs and short tandem repeats; target = read
    with indels inside runs/repeats (where gap placement is a tie)."""
    parts = []
    while sum(len(p) for p in parts) < length:
        kind = rng.integers(0, 3)
        if kind == 0:
            parts.append([int(rng.integers(1, 5))] * int(rng.integers(3, 9)))
        elif kind == 1:
            unit = list(rng.integers(1, 5, size=int(rng.integers(2, 4))))
            parts.append([int(x) for x in unit] * int(rng.integers(2, 5)))
        else:
            parts.append([int(x) for x in rng.integers(1, 5, size=int(rng.integers(1, 6)))])
    read = [b for p in parts for b in p][:length]
    tgt = list(read)
    for _ in range(int(rng.integers(1, 4))):
        pos = int(rng.integers(1, len(tgt) - 1))
        op = rng.integers(0, 3)
        if op == 0:
            del tgt[pos:pos + int(rng.integers(1, 4))]
        elif op == 1:
            tgt[pos:pos] = [tgt[pos]] * int(rng.integers(1, 4))
        else:
            tgt[pos] = int(rng.integers(1, 5))
    flank_l = [int(x) for x in rng.integers(1, 5, size=int(rng.integers(0, 8)))]
    flank_r = [int(x) for x in rng.integers(1, 5, size=int(rng.integers(0, 8)))]
    return read, flank_l + tgt + flank_r, len(flank_l)


def sw_cases(R, rng):
    import torch

    cases = []
    scorings = [(2, 4, 6, 2), (2, 4, 4, 2), (1, 1, 2, 1), (2, 4, 2, 2), (3, 2, 5, 1)]
    for k in range(60):
        ma, mi, go, ge = scorings[k % len(scorings)]
        B = int(rng.integers(1, 5))
        pairs = [_tie_rich_pair(rng, int(rng.integers(8, 60))) for _ in range(B)]
        hb = int(rng.integers(2, 12))
        xdrop = [0.0, 0.0, 5.0, 9.5, 20.0][k % 5]
        with_n = k % 7 == 3
        M = max(len(p[0]) for p in pairs)
        N = max(len(p[1]) for p in pairs)
        q = np.full((B, M), -1, np.int16)
        t = np.full((B, N), -1, np.int16)
        for b, (rd, tg, fl) in enumerate(pairs):
            q[b, :len(rd)] = rd
            t[b, :len(tg)] = tg
            if with_n:
                q[b, int(rng.integers(0, len(rd)))] = 5
                t[b, int(rng.integers(0, len(tg)))] = 5
        qlen = [len(p[0]) for p in pairs]
        tlen = [len(p[1]) for p in pairs]
        off = [int(p[2]) + int(rng.integers(-2, 3)) for p in pairs]
        cfg = R["ExtensionConfig"](match_score=float(ma), mismatch_penalty=float(mi),
                                   gap_open=float(go), gap_extend=float(ge), x_drop=xdrop)
        res = R["banded_affine_sw_batch"](torch.as_tensor(q), torch.as_tensor(t),
                                          torch.tensor(qlen), torch.tensor(tlen), cfg,
                                          half_band=hb, band_offset=torch.tensor(off),
                                          return_matrices=True)
        outs = []
        for b in range(B):
            ext = R["traceback_banded"](res, b, q[b], t[b], cfg)
            outs.append(dict(score=float(res.score[b]), query_end=int(res.query_end[b]),
                             target_end=int(res.target_end[b]), cigar=ext.cigar,
                             read_start=ext.read_start, ref_start=ext.ref_start))
        cases.append(dict(query=q.tolist(), target=t.tolist(), qlen=qlen, tlen=tlen,
                          off=off, half_band=hb, scoring=[ma, mi, go, ge], x_drop=xdrop,
                          out=outs))
    return cases


def chain_cases(R, rng):
    cases = []
    cfg = R["ChainingConfig"]()
    for k in range(40):
        n = int(rng.integers(1, 160))
        lookback = [64, 4, 16, 1, 50][k % 5]
        max_gap = [5000, 200, 5000, 60][k % 4]
        c = R["ChainingConfig"](max_lookback=lookback, max_gap=max_gap)
        read_end = np.sort(rng.integers(15, 3000, size=n)).astype(np.int64)
        ref_end = np.sort(rng.integers(15, 3300, size=n)).astype(np.int64)
        # equal keys and collinear runs create ties
        if k % 3 == 0:
            ref_end = read_end + rng.integers(-3, 4, size=n)
        if k % 2 == 0:
            weight = rng.integers(5, 40, size=n).astype(np.float64)
        else:
            weight = rng.uniform(5, 40, size=n)
        order = np.lexsort((read_end, ref_end))
        read_end, ref_end, weight = read_end[order], ref_end[order], weight[order]
        f, parent = R["chain_dp_numpy"](read_end, ref_end, weight, c)
        cases.append(dict(read_end=read_end.tolist(), ref_end=ref_end.tolist(),
                          weight=[float(x).hex() for x in weight], max_lookback=lookback,
                          max_gap=max_gap, gap_open=c.gap_open, gap_extend=c.gap_extend,
                          log_coeff=c.log_coeff, f=[float(x).hex() for x in f],
                          parent=parent.tolist()))
    g = np.arange(cfg.max_gap, dtype=np.int64)
    lut = np.where(g > 0, cfg.gap_open + cfg.gap_extend * g + cfg.log_coeff * np.log2(g + 1.0),
                   0.0)
    return dict(cases=cases, default_lut=[float(x).hex() for x in lut])


def chainer_cases(R, rng):
    cases = []
    for k in range(30):
        n = int(rng.integers(0, 120))
        read_len = 2000
        loci = rng.integers(0, 20000, size=3)
        read_pos, ref_pos, strand = [], [], []
        for a in range(n):
            s = 1 if rng.random() < 0.6 else -1
            rp = int(rng.integers(0, read_len - 15))
            if rng.random() < 0.8:
                base = int(loci[a % 3])
                ref_pos.append(base + rp + int(rng.integers(-2, 3)))
            else:
                ref_pos.append(int(rng.integers(0, 22000)))
            read_pos.append(rp)
            strand.append(s)
        length = rng.choice([15, 15, 21], size=n)
        anchors = R["AnchorSet"].from_lists(read_pos=read_pos, ref_pos=ref_pos,
                                            length=length, strand=strand,
                                            read_len=read_len, ref_len=22000)
        cfg = R["ChainingConfig"](max_lookback=[64, 8, 50][k % 3],
                                  min_chain_score=[20.0, 1.0][k % 2])
        chains = R["AffineChainer"](cfg).chain(anchors)
        cases.append(dict(read_pos=list(map(int, read_pos)), ref_pos=list(map(int, ref_pos)),
                          length=list(map(int, length)), strand=list(map(int, strand)),
                          read_len=read_len, max_lookback=cfg.max_lookback,
                          min_chain_score=cfg.min_chain_score,
                          chains=[dict(anchor_idx=c.anchor_idx.tolist(),
                                       score=float(c.score).hex(), strand=c.strand,
                                       read_start=c.read_start, read_end=c.read_end,
                                       ref_start=c.ref_start, ref_end=c.ref_end,
                                       is_primary=bool(c.is_primary)) for c in chains]))
    return cases


def extend_cases(R, rng):
    cases = []
    bases = "ACGT"
    for k in range(16):
        ref = "".join(rng.choice(list(bases), size=3000))
        start = int(rng.integers(200, 2500))
        L = int(rng.integers(60, 220))
        read = list(ref[start:start + L])
        for _ in range(int(rng.integers(0, 5))):
            p = int(rng.integers(1, len(read) - 1))
            r = rng.integers(0, 3)
            if r == 0:
                read[p] = bases[int(rng.integers(0, 4))]
            elif r == 1:
                del read[p]
            else:
                read.insert(p, bases[int(rng.integers(0, 4))])
        if k % 4 == 1:
            read[int(rng.integers(0, len(read)))] = "N"
        read = "".join(read)
        strand = 1
        if k % 3 == 2:
            comp = {"A": "T", "C": "G", "G": "C", "T": "A", "N": "N"}
            read = "".join(comp[c] for c in reversed(read))
            strand = -1
        nch = int(rng.integers(1, 4))
        chains = []
        for c in range(nch):
            shift = 0 if c == 0 else int(rng.integers(-40, 40))
            rs = int(rng.integers(0, 20))
            re_ = min(L, rs + 40)
            idx = np.arange(int(rng.integers(1, 4)), dtype=np.int64)
            chains.append(R["Chain"](anchor_idx=idx, score=50.0, strand=strand,
                                     read_start=rs, read_end=re_,
                                     ref_start=start + rs + shift,
                                     ref_end=start + re_ + shift))
        anchors = R["AnchorSet"].from_lists(
            read_pos=[0, 20, 40], ref_pos=[start, start + 20 + int(rng.integers(0, 6)),
                                           start + 40 + int(rng.integers(0, 9))],
            length=[15, 15, 15], strand=[strand] * 3, read_len=len(read), ref_len=len(ref))
        cfg = R["ExtensionConfig"](half_band=int(rng.integers(8, 40)),
                                   flank=int(rng.integers(10, 60)),
                                   x_drop=[600.0, 0.0, 30.0, 12.0][k % 4])
        out = R["ExtensionEngine"](cfg, device="cpu").extend_chains(read, chains, anchors, ref)
        cases.append(dict(ref=ref, read=read,
                          anchors=dict(read_pos=[0, 20, 40],
                                       ref_pos=anchors.ref_pos.tolist(),
                                       length=[15, 15, 15], strand=[strand] * 3),
                          chains=[dict(anchor_idx=c.anchor_idx.tolist(), strand=c.strand,
                                       read_start=c.read_start, read_end=c.read_end,
                                       ref_start=c.ref_start, ref_end=c.ref_end)
                                  for c in chains],
                          cfg=dict(half_band=cfg.half_band, flank=cfg.flank,
                                   x_drop=cfg.x_drop),
                          out=[dict(score=e.score, cigar=e.cigar, read_start=e.read_start,
                                    read_end=e.read_end, ref_start=e.ref_start,
                                    ref_end=e.ref_end) for e in out]))
    return cases


def seeding_cases(rng):
    sys.path.insert(0, REF)
    from graphmambaformer.alignment import seeding as SD

    bases = np.array(list("ACGTN"))
    out = {}
    vals = [0, 1, 2, 3, (1 << 30) - 1, (1 << 62) - 1] + \
        [int(x) for x in rng.integers(0, 1 << 62, size=40, dtype=np.int64)]
    out["hash64"] = {"in": vals, "out": [int(x) for x in SD.hash64(np.array(vals, np.int64))]}
    masks = []
    for n, (k, w) in enumerate([(5, 1), (7, 4), (15, 10), (21, 5), (11, 16), (3, 3)]):
        L = int(rng.integers(20, 200))
        codes = rng.integers(0, 4, size=L)
        if n % 2:
            codes[rng.integers(0, L, size=2)] = 4
        seq = "".join(bases[codes])
        packed, valid = SD.pack_kmers(SD.encode_bases(seq), k)
        sel = valid & SD.minimizer_mask(SD.hash64(packed), w)
        masks.append({"seq": seq, "k": k, "w": w, "packed": [int(x) for x in packed],
                      "valid": np.flatnonzero(valid).tolist(),
                      "selected": np.flatnonzero(sel).tolist()})
    out["kmers"] = masks
    queries = []
    for n, (k, w, occ) in enumerate([(15, 10, 0), (11, 5, 3), (7, 4, 0), (19, 8, 0)]):
        unit = rng.integers(0, 4, size=300)
        codes = np.concatenate([rng.integers(0, 4, size=6000), np.tile(unit[:37], 40),
                                rng.integers(0, 4, size=6000), unit, unit,
                                rng.integers(0, 4, size=3000)])
        codes[rng.integers(0, len(codes), size=30)] = 4
        ref = "".join(bases[codes])
        idx = SD.GPUKmerIndex(SD.encode_bases(ref), k=k, window=w, max_occ=occ, device="cpu")
        reads = []
        for r in range(6):
            if r == 5:
                read = ref[100:100 + k - 1]  # shorter than k
            else:
                s0 = int(rng.integers(0, len(ref) - 1500))
                read = list(ref[s0:s0 + int(rng.integers(200, 1500))])
                for _ in range(int(rng.integers(0, 12))):
                    read[int(rng.integers(0, len(read)))] = "ACGTN"[int(rng.integers(0, 5))]
                read = "".join(read)
                if r == 3:
                    read = "".join({"A": "T", "C": "G", "G": "C", "T": "A"}.get(c, "N")
                                   for c in reversed(read))
            rp, fp, ln = idx.query(SD.encode_bases(read))
            reads.append({"read": read, "read_pos": rp.tolist(), "ref_pos": fp.tolist(),
                          "length": ln.tolist()})
        queries.append({"ref": ref, "k": k, "window": w, "max_occ": occ,
                        "n_keys": int(len(idx.keys)), "reads": reads})
    out["queries"] = queries
    return out


def wfa_cases(rng):
    sys.path.insert(0, REF)
    from graphmambaformer.alignment.extension import ExtensionEngine, WavefrontAligner
    from graphmambaformer.alignment.seeding import encode_bases
    from graphmambaformer.alignment.types import AnchorSet, Chain
    from graphmambaformer.config import ExtensionConfig

    bases = "ACGT"
    out = {"align": [], "extend": []}

    def mutate(a, n_ed, alphabet="ACGT"):
        b = list(a)
        for _ in range(n_ed):
            if not b:
                b.append(alphabet[int(rng.integers(0, len(alphabet)))])
                continue
            pos = int(rng.integers(0, len(b)))
            act = int(rng.integers(0, 3))
            if act == 0:
                b[pos] = alphabet[int(rng.integers(0, len(alphabet)))]
            elif act == 1:
                b.pop(pos)
            else:
                b.insert(pos, alphabet[int(rng.integers(0, len(alphabet)))])
        return "".join(b)

    pairs = [("", ""), ("", "ACG"), ("ACGT", ""), ("AN", "AN"), ("ACGT", "TGCA"),
             ("A", "A"), ("A", "C"), ("NNNN", "NNNN"), ("ACGTN", "ACGTA")]
    for n in range(120):
        L = int(rng.integers(1, 60)) if n < 90 else int(rng.integers(200, 900))
        a = "".join(bases[int(x)] for x in rng.integers(0, 4, size=L))
        if n % 9 == 0:
            a = "".join("ACGTN"[int(x)] for x in rng.integers(0, 5, size=L))
        if n % 7 == 0:  # homopolymer / tandem content, ties in the traceback
            a = "".join(bases[int(x)] * int(rng.integers(1, 6)) for x in rng.integers(0, 4, size=L))
        b = mutate(a, int(rng.integers(0, max(2, L // 8))), "ACGTN" if n % 11 == 0 else "ACGT")
        pairs.append((a, b))
    for n in range(10):  # unrelated: over a small budget
        pairs.append(("".join(bases[int(x)] for x in rng.integers(0, 4, size=40)),
                      "".join(bases[int(x)] for x in rng.integers(0, 4, size=40))))
    for n, (a, b) in enumerate(pairs):
        maxd = 6 if n >= len(pairs) - 10 else 4096
        wfa = WavefrontAligner(ExtensionConfig(algorithm="wfa", wfa_max_distance=maxd))
        try:
            d, cig = wfa.align(encode_bases(a), encode_bases(b))
            out["align"].append({"q": a, "t": b, "max_distance": maxd, "distance": int(d),
                                 "cigar": [[op, int(k)] for op, k in cig]})
        except RuntimeError:
            out["align"].append({"q": a, "t": b, "max_distance": maxd, "distance": None,
                                 "cigar": None})
    # ExtensionEngine WFA branch, incl. the banded fallback past the budget
    ref = "".join(bases[int(x)] for x in rng.integers(0, 4, size=4000))
    for n in range(12):
        s0 = int(rng.integers(0, 3000))
        L = int(rng.integers(50, 700))
        read = mutate(ref[s0:s0 + L], int(rng.integers(0, max(2, L // 30))))
        strand = 1
        if n % 4 == 3:
            read = "".join({"A": "T", "C": "G", "G": "C", "T": "A"}.get(c, "N")
                           for c in reversed(read))
            strand = -1
        maxd = 3 if n % 5 == 4 else 4096
        cfg = ExtensionConfig(algorithm="wfa", wfa_max_distance=maxd, half_band=24, flank=40)
        anchors = AnchorSet.from_lists([0, 20], [s0, s0 + 20], [15, 15], [strand, strand],
                                       read_len=len(read), ref_len=len(ref))
        chains = [Chain(anchor_idx=np.array([0, 1]), score=40.0, strand=strand, read_start=0,
                        read_end=int(rng.integers(30, len(read) + 1)), ref_start=s0,
                        ref_end=s0 + int(rng.integers(30, L + 5)))]
        res = ExtensionEngine(cfg).extend_chains(read, chains, anchors, ref)
        out["extend"].append({
            "ref": ref, "read": read, "max_distance": maxd,
            "chain": dict(strand=strand, read_start=chains[0].read_start,
                          read_end=chains[0].read_end, ref_start=chains[0].ref_start,
                          ref_end=chains[0].ref_end),
            "out": [dict(score=float(e.score), cigar=[[op, int(k)] for op, k in e.cigar],
                         read_start=e.read_start, read_end=e.read_end, ref_start=e.ref_start,
                         ref_end=e.ref_end) for e in res]})
    return out


def main():
    want = set(sys.argv[1:])
    R = _imports()
    rng = np.random.default_rng(20261203)
    out = {
        "sw_ties.json": lambda: sw_cases(R, rng),
        "chain_dp.json": lambda: chain_cases(R, rng),
        "chainer.json": lambda: chainer_cases(R, rng),
        "extend.json": lambda: extend_cases(R, rng),
        "seeding.json": lambda: seeding_cases(np.random.default_rng(20261204)),
        "wfa.json": lambda: wfa_cases(np.random.default_rng(20261205)),
    }
    for name, make in out.items():
        data = make()  # in file order: every file draws from the shared stream
        if want and name not in want:
            continue
        with open(os.path.join(HERE, name), "w") as fh:
            json.dump(data, fh, separators=(",", ":"))
        print(name, os.path.getsize(os.path.join(HERE, name)), "bytes")


if __name__ == "__main__":
    main()

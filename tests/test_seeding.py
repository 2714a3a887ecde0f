"""CPU pins of the seeding oracle (oracle/seeding.py): the reference's golden
vectors, the reference's own seeding tests restated, brute force and the
hash's invertibility."""

import numpy as np

from oracle import seeding as S
from tests.golden.load import load

BASES = "ACGT"


def test_hash64_golden_and_invertible():
    g = load("seeding.json")["hash64"]
    got = S.hash64(np.array(g["in"], np.int64))
    assert [int(x) for x in got] == g["out"]
    for v, h in zip(g["in"], g["out"]):
        assert S.hash64_inverse(h) == v


def test_pack_and_minimizers_golden():
    for case in load("seeding.json")["kmers"]:
        packed, valid = S.pack_kmers(S.encode(case["seq"]), case["k"])
        assert packed.tolist() == case["packed"]
        assert np.flatnonzero(valid).tolist() == case["valid"]
        sel = valid & S.minimizer_mask(S.hash64(packed), case["w"])
        assert np.flatnonzero(sel).tolist() == case["selected"]


def test_kmer_query_golden():
    for case in load("seeding.json")["queries"]:
        tab = S.build_index(S.encode(case["ref"]), case["k"], case["window"], case["max_occ"])
        assert len(tab.keys) == case["n_keys"]
        for r in case["reads"]:
            rp, fp, ln = S.kmer_query(tab, S.encode(r["read"]), case["window"])
            assert rp.tolist() == r["read_pos"]
            assert fp.tolist() == r["ref_pos"]
            assert ln.tolist() == r["length"]


def test_minimizer_mask_explicit_windows_and_ties():
    """test_alignment_stages.py:128-139 restated, plus ties: a position is
    selected iff it attains the minimum of some window containing it."""
    rng = np.random.default_rng(1)
    for window in (1, 3, 5, 10, 64):
        for hi in (1000, 4):  # 4: many ties
            h = rng.integers(0, hi, size=70).astype(np.uint64)
            mask = S.minimizer_mask(h, window)
            w = max(1, min(window, len(h)))
            exp = np.zeros(len(h), bool)
            for s in range(len(h) - w + 1):
                blk = h[s:s + w]
                exp[s:s + w] |= blk == blk.min()
            assert np.array_equal(mask, exp), (window, hi)


def test_packing_round_trip_and_n_windows():
    """test_alignment_stages.py:142-160 restated."""
    rng = np.random.default_rng(2)
    seq = "".join(rng.choice(list(BASES), size=50))
    k = 8
    packed, valid = S.pack_kmers(S.encode(seq), k)
    assert len(packed) == len(seq) - k + 1 and valid.all()
    for pos, v in enumerate(packed):
        digits = [(int(v) >> (2 * (k - 1 - j))) & 3 for j in range(k)]
        assert "".join(BASES[d] for d in digits) == seq[pos:pos + k]
    with_n = seq[:10] + "N" + seq[11:]
    p2, v2 = S.pack_kmers(S.encode(with_n), k)
    assert not v2[max(0, 10 - k + 1):11].any() and v2[:3].all()
    assert (p2[~v2] == 0).all()


def test_kmer_query_brute_force():
    """Dict index over every selected reference k-mer; the query returns, per
    selected read position in order, every reference position of its k-mer in
    reference order."""
    rng = np.random.default_rng(3)
    unit = rng.integers(1, 5, size=25)
    ref = np.concatenate([rng.integers(1, 5, size=3000), np.tile(unit, 20),
                          rng.integers(1, 5, size=2000)]).astype(np.int8)
    ref[rng.integers(0, len(ref), size=15)] = 5
    for k, w, occ in ((9, 5, 0), (13, 1, 0), (6, 3, 4)):
        tab = S.build_index(ref, k, w, occ)
        packed, valid = S.pack_kmers(ref, k)
        sel = valid & S.minimizer_mask(S.hash64(packed), w)
        d = {}
        for p in np.flatnonzero(sel):
            d.setdefault(int(packed[p]), []).append(int(p))
        if occ:
            d = {key: v for key, v in d.items() if len(v) <= occ}
        for _ in range(5):
            s0 = int(rng.integers(0, len(ref) - 600))
            read = ref[s0:s0 + 500].copy()
            read[rng.integers(0, 500, size=6)] = rng.integers(1, 6, size=6)
            rp, fp, ln = S.kmer_query(tab, read, w)
            qp, qv = S.pack_kmers(read, k)
            qs = qv & S.minimizer_mask(S.hash64(qp), w)
            exp = [(int(p), q) for p in np.flatnonzero(qs) for q in d.get(int(qp[p]), [])]
            assert list(zip(rp.tolist(), fp.tolist())) == exp
            assert (ln == k).all()


def test_product_host_functions_match_oracle():
    """The library's host-side pack / hash / minimizer / table build (used to
    build the device index) equal the oracle on random inputs."""
    from graphmambaformer_amd.alignment import seeding as P

    rng = np.random.default_rng(4)
    v = rng.integers(0, 1 << 62, size=500, dtype=np.int64)
    assert np.array_equal(P.hash64(v), S.hash64(v))
    for _ in range(60):
        L = int(rng.integers(1, 300))
        c = rng.integers(-1, 7, size=L).astype(np.int8)
        k = int(rng.integers(1, 25))
        a, b = P.pack_kmers(c, k), S.pack_kmers(c, k)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        h = rng.integers(0, 6, size=L).astype(np.uint64)
        w = int(rng.integers(1, 40))
        assert np.array_equal(P.minimizer_mask(h, w), S.minimizer_mask(h, w))
    ref = rng.integers(1, 5, size=30000).astype(np.int8)
    ref[::997] = 5
    ref[5000:5600] = np.tile(ref[:20], 30)
    for k, w, occ in ((11, 5, 0), (15, 10, 3), (7, 1, 0)):
        pk, vd = P.pack_kmers(ref, k)
        t1 = P.KmerTable.build(pk, vd & P.minimizer_mask(P.hash64(pk), w), k, max_occ=occ)
        t2 = S.build_index(ref, k, w, occ)
        assert np.array_equal(t1.keys, t2.keys) and np.array_equal(t1.offsets, t2.offsets)
        assert np.array_equal(t1.positions, t2.positions)
        q = np.concatenate([t2.keys[::7], rng.integers(-5, 1 << (2 * k), size=200)])
        assert all(np.array_equal(x, y) for x, y in zip(t1.lookup(q), t2.lookup(q)))

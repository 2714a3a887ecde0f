"""Pins the CPU oracle before anything is checked against it.

Each check is chosen so a plausible slip (a dropped term, a wrong sign or
index, a transposed operand) fails at least one of them:

* known answers from the reference's own tests (KAT-SW1, KAT-CH1, KAT-CH2),
* closed forms (identical sequences score 2*len with CIGAR len=),
* brute force (unbanded Gotoh when the band covers the matrix; O(n^2)
  chaining; lexicographic first best cell),
* invariants (CIGAR rescoring equals the score, consumed spans),
* the float32 -1e30 transcription (sentinel mapping, matrices cell by cell),
* golden vectors written from the reference itself (ties, x-drop, default
  non-integer chaining profile, chainer selection, extension geometry).
"""

import numpy as np
import pytest

from oracle import brute
from oracle import ref_semantics as rs
from tests.golden.load import fhex, load

pytestmark = pytest.mark.usefixtures("oracle_lib")


def O():
    from oracle import oracle

    return oracle


def enc(s):
    return O().encode(s)


# ----------------------------------------------------------------------------
# banded SW
# ----------------------------------------------------------------------------
def test_kat_sw1():
    """tests/test_accel.py:306-334 of the reference: ACGT vs TTACGTAA."""
    q, t = enc("ACGT"), enc("TTACGTAA")
    r = O().sw_banded(q[None], t[None], [4], [8], [2], 2, 2, 4, 6, 2, store=True)
    assert (int(r["score"][0]), int(r["query_end"][0]), int(r["target_end"][0])) == (8, 4, 6)
    tb = O().sw_traceback(r, 0, q, t, 2, 4, 6, 2)
    assert tb["cigar"] == [("=", 4)]
    assert (tb["read_start"], tb["ref_start"]) == (0, 2)


def test_identical_sequences():
    """Closed form: score 2*len, CIGAR len= (test_alignment_stages.py:1029-1048)."""
    rng = np.random.default_rng(11)
    for L in (1, 7, 40, 133):
        s = rng.integers(1, 5, size=L).astype(np.int16)
        r = O().sw_banded(s[None], s[None], [L], [L], [0], 8, 2, 4, 6, 2, store=True)
        assert int(r["score"][0]) == 2 * L
        tb = O().sw_traceback(r, 0, s, s, 2, 4, 6, 2)
        assert tb["cigar"] == [("=", L)]
        assert tb["read_start"] == 0 and tb["ref_start"] == 0


@pytest.mark.parametrize("scoring", [(2, 4, 6, 2), (2, 4, 4, 2), (1, 3, 5, 2), (2, 4, 2, 2)])
def test_full_band_equals_unbanded_gotoh(scoring):
    """Band covering the whole matrix == textbook Gotoh (test_alignment_stages.py:977-1026)."""
    ma, mi, go, ge = scoring
    rng = np.random.default_rng(5)
    for _ in range(12):
        n = int(rng.integers(5, 30))
        m = int(rng.integers(5, 30))
        q = rng.integers(1, 5, size=n).astype(np.int16)
        t = rng.integers(1, 5, size=m).astype(np.int16)
        hb = max(n, m) + 2
        for off in (0, 3, -2):
            r = O().sw_banded(q[None], t[None], [n], [m], [off], hb, ma, mi, go, ge)
            exp, ei, ej = brute.first_best_cell(q, t, ma, mi, go, ge)
            assert int(r["score"][0]) == int(exp)
            assert (int(r["query_end"][0]), int(r["target_end"][0])) == (ei, ej)


def _random_pairs(rng, count, with_n=False):
    pairs = []
    for _ in range(count):
        n = int(rng.integers(3, 50))
        q = rng.integers(1, 5, size=n).astype(np.int16)
        t = q.copy().tolist()
        for _ in range(int(rng.integers(0, 4))):
            p = int(rng.integers(0, len(t)))
            k = rng.integers(0, 3)
            if k == 0:
                t[p] = int(rng.integers(1, 5))
            elif k == 1 and len(t) > 2:
                del t[p]
            else:
                t.insert(p, int(rng.integers(1, 5)))
        t = rng.integers(1, 5, size=int(rng.integers(0, 6))).tolist() + t
        t = np.array(t, np.int16)
        if with_n:
            q[int(rng.integers(0, n))] = 5
            t[int(rng.integers(0, len(t)))] = 5
        pairs.append((q, t))
    return pairs


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_int32_matches_float32_transcription(seed):
    """The ONEG saturating mapping reproduces float32 -1e30 cell by cell."""
    rng = np.random.default_rng(seed)
    scorings = [(2, 4, 6, 2), (2, 4, 4, 2), (2, 4, 2, 2), (3, 1, 3, 1)]
    for k in range(6):
        ma, mi, go, ge = scorings[k % 4]
        pairs = _random_pairs(rng, int(rng.integers(1, 5)), with_n=(k % 2 == 1))
        B = len(pairs)
        M = max(len(p[0]) for p in pairs)
        N = max(len(p[1]) for p in pairs)
        q = np.full((B, M), -1, np.int16)
        t = np.full((B, N), -1, np.int16)
        for b, (a, c) in enumerate(pairs):
            q[b, :len(a)] = a
            t[b, :len(c)] = c
        ql = [len(p[0]) for p in pairs]
        tl = [len(p[1]) for p in pairs]
        off = rng.integers(-3, 6, size=B).tolist()
        hb = int(rng.integers(1, 9))
        xd = [0.0, 7.0, 15.5][k % 3]
        ri = O().sw_banded(q, t, ql, tl, off, hb, ma, mi, go, ge, x_drop=xd, store=True)
        rf = rs.sw_rows_f32(q, t, ql, tl, off, hb, ma, mi, go, ge, xdrop=xd)
        assert np.array_equal(ri["score"], rf["score"].astype(np.int32))
        assert np.array_equal(ri["query_end"], rf["query_end"])
        assert np.array_equal(ri["target_end"], rf["target_end"])
        for X in ("H", "E", "F"):
            a = ri[X].astype(np.float64)
            a[a <= -(1 << 29)] = -1e30
            b = rf[X].astype(np.float64)
            b[b <= -1e29] = -1e30
            assert np.array_equal(a, b), X
        for b in range(B):
            ti = O().sw_traceback(ri, b, q[b], t[b], ma, mi, go, ge)
            cf, i0, j0 = rs.traceback_f32(rf, b, q[b], t[b], ma, mi, go, ge)
            assert ti["cigar"] == cf and ti["read_start"] == i0 and ti["ref_start"] == j0


def test_cigar_rescoring_and_spans():
    rng = np.random.default_rng(9)
    for (q, t) in _random_pairs(rng, 60):
        for ma, mi, go, ge in ((2, 4, 6, 2), (2, 4, 4, 2), (1, 2, 3, 1)):
            off = int(rng.integers(-2, 6))
            r = O().sw_banded(q[None], t[None], [len(q)], [len(t)], [off], 6, ma, mi, go, ge,
                              store=True)
            tb = O().sw_traceback(r, 0, q, t, ma, mi, go, ge)
            s, i, j = brute.rescore_cigar(tb["cigar"], q, t, tb["read_start"],
                                          tb["ref_start"], ma, mi, go, ge)
            assert s == int(r["score"][0])
            assert i == tb["read_end"] and j == tb["ref_end"]


def test_empty_and_degenerate_inputs():
    q = np.zeros((1, 0), np.int16)
    t = enc("ACGT")[None]
    r = O().sw_banded(np.full((1, 1), -1, np.int16), t, [0], [4], [0], 2, 2, 4, 6, 2)
    assert int(r["score"][0]) == 0 and int(r["query_end"][0]) == 0
    r = O().sw_banded(enc("ACGT")[None], np.full((1, 1), -1, np.int16), [4], [0], [0], 2, 2, 4,
                      6, 2)
    assert int(r["score"][0]) == 0
    # all-N: nothing matches, score 0
    n5 = np.full((1, 6), 5, np.int16)
    r = O().sw_banded(n5, n5, [6], [6], [0], 3, 2, 4, 6, 2)
    assert int(r["score"][0]) == 0
    del q


def test_golden_sw_ties_and_xdrop():
    """Reference outputs on tie-rich inputs and batch-wide x-drop."""
    for case in load("sw_ties.json"):
        ma, mi, go, ge = case["scoring"]
        q = np.array(case["query"], np.int16)
        t = np.array(case["target"], np.int16)
        r = O().sw_banded(q, t, case["qlen"], case["tlen"], case["off"], case["half_band"],
                          ma, mi, go, ge, x_drop=case["x_drop"], store=True)
        for b, exp in enumerate(case["out"]):
            assert int(r["score"][b]) == int(exp["score"])
            assert int(r["query_end"][b]) == exp["query_end"]
            assert int(r["target_end"][b]) == exp["target_end"]
            tb = O().sw_traceback(r, b, q[b], t[b], ma, mi, go, ge)
            assert [list(x) for x in tb["cigar"]] == [list(x) for x in exp["cigar"]]
            assert tb["read_start"] == exp["read_start"]
            assert tb["ref_start"] == exp["ref_start"]


def test_golden_extend_chains():
    """ExtensionEngine.extend_chains geometry, soft clips and lifts."""
    for case in load("extend.json"):
        a = case["anchors"]
        diags = np.array(a["ref_pos"]) - np.array(a["read_pos"])
        cfg = case["cfg"]
        out = O().extend_chains(case["read"], case["chains"], diags, case["ref"],
                                half_band=cfg["half_band"], flank=cfg["flank"],
                                x_drop=cfg["x_drop"])
        assert len(out) == len(case["out"])
        for got, exp in zip(out, case["out"]):
            assert got["score"] == int(exp["score"])
            assert [list(x) for x in got["cigar"]] == [list(x) for x in exp["cigar"]]
            for k in ("read_start", "read_end", "ref_start", "ref_end"):
                assert got[k] == exp[k], k


# ----------------------------------------------------------------------------
# chaining
# ----------------------------------------------------------------------------
DEFAULT_LUT = None


def lut_default():
    global DEFAULT_LUT
    if DEFAULT_LUT is None:
        DEFAULT_LUT = O().penalty_lut(6.0, 0.05, 0.5, 5000)
    return DEFAULT_LUT


def test_kat_ch1():
    """tests/test_accel.py:347-370 of the reference, derived by hand."""
    f, p = O().chain_dp_f64([10, 20, 30], [10, 20, 31], [10.5, 9.25, 8.75], 64, 5000,
                            lut_default())
    assert np.allclose(f, [10.5, 19.75, 21.95], rtol=0, atol=1e-12)
    assert p.tolist() == [-1, 0, 1]


def test_kat_ch2():
    """tests/test_alignment_stages.py:502-517 (lookback 4), rows derived by hand."""
    f, p = O().chain_dp_f64([10, 20, 30], [10, 21, 32], [10.5, 9.25, 8.75], 4, 5000,
                            lut_default())
    assert np.allclose(f, [10.5, 13.2, 15.4], rtol=0, atol=1e-12) and p.tolist() == [-1, 0, 1]
    f, p = O().chain_dp_f64([8, 18], [8, 20], [7.5, 8.25], 4, 5000, lut_default())
    assert np.isclose(f[1], 7.5 + 8.25 - (6.1 + 0.5 * np.log2(3.0)), rtol=0, atol=1e-12)
    assert p.tolist() == [-1, 0]


def test_penalty_lut_matches_committed():
    """The LUT computed here equals the one the reference computed (bitwise)."""
    got = lut_default()
    exp = fhex(load("chain_dp.json")["default_lut"])
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("profile", ["int", "float"])
def test_chain_dp_matches_bruteforce(profile):
    rng = np.random.default_rng(17)
    for _ in range(25):
        n = int(rng.integers(1, 60))
        re_ = np.sort(rng.integers(1, 600, size=n)).astype(np.int64)
        fe = np.sort(rng.integers(1, 600, size=n)).astype(np.int64)
        L = int(rng.integers(1, 20))
        G = int(rng.integers(30, 700))
        if profile == "int":
            w = rng.integers(5, 40, size=n).astype(np.int32)
            f, p = O().chain_dp_i32(re_, fe, w, L, G, 6, 1)
            ef, ep = brute.chain_dp_bruteforce(re_, fe, w.tolist(), L, G, brute.int_penalty(6, 1))
            assert f.tolist() == ef and p.tolist() == ep
        else:
            w = rng.integers(5, 40, size=n).astype(np.float64)
            f, p = O().chain_dp_f64(re_, fe, w, L, G, O().penalty_lut(6.0, 0.05, 0.5, G))
            ef, ep = brute.chain_dp_bruteforce(re_, fe, w.tolist(), L, G,
                                               brute.float_penalty(6.0, 0.05, 0.5))
            assert np.allclose(f, ef, rtol=0, atol=1e-9)
            for i in range(n):
                if p[i] != ep[i]:
                    assert np.isclose(f[i], ef[i])


def test_golden_chain_dp_default_profile():
    """chain_dp_numpy float64 outputs reproduced bit for bit."""
    data = load("chain_dp.json")
    for case in data["cases"]:
        lut = O().penalty_lut(case["gap_open"], case["gap_extend"], case["log_coeff"],
                              case["max_gap"])
        f, p = O().chain_dp_f64(case["read_end"], case["ref_end"], fhex(case["weight"]),
                                case["max_lookback"], case["max_gap"], lut)
        assert np.array_equal(f, fhex(case["f"]))
        assert p.tolist() == case["parent"]


def test_golden_chainer():
    """AffineChainer.chain: strands, DP order, backtrack, select."""
    for case in load("chainer.json"):
        got = O().chain_read(case["read_pos"], case["ref_pos"], case["length"], case["strand"],
                             max_lookback=case["max_lookback"],
                             min_chain_score=case["min_chain_score"])
        exp = case["chains"]
        assert len(got) == len(exp)
        for g, e in zip(got, exp):
            assert g["anchor_idx"].tolist() == e["anchor_idx"]
            assert g["score"] == float.fromhex(e["score"])
            for k in ("strand", "read_start", "read_end", "ref_start", "ref_end", "is_primary"):
                assert g[k] == e[k], k


def test_backtrack_hand_case():
    """Claimed-walk partial chain: f = [5, 9, 12, 7], parents 0<-1<-2, 1<-3.
    Visit 2 (12): chain {0,1,2} score 12.  Visit 3 (7): stops at claimed 1, score
    7 - 9 = -2 < min score -> dropped."""
    chains = O().chain_backtrack([5.0, 9.0, 12.0, 7.0], [-1, 0, 1, 1], 1, -5.0)
    assert [(m.tolist(), s) for m, s in chains] == [([0, 1, 2], 12.0), ([3], -2.0)]
    chains = O().chain_backtrack([5.0, 9.0, 12.0, 7.0], [-1, 0, 1, 1], 1, 0.0)
    assert [(m.tolist(), s) for m, s in chains] == [([0, 1, 2], 12.0)]


def test_fullsize_digest_matches_the_oracle_on_a_sample():
    """The committed every-pair digest (tests/golden/fullsize_digest.npz) is the
    oracle's: the generators reproduce its inputs, and a stratified sample of
    pairs / reads recomputed here gives the same entries."""
    import os

    from graphmambaformer_amd import workloads as WL
    from tests.golden import make_fullsize_digest as D

    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "fullsize_digest.npz"))
    for name in ("sw_long", "sw_noisy"):
        w = getattr(WL, name)()
        assert D.sw_input_digest(w) == int(d[f"{name}_input"][0]), name
        cells = w.cells()
        srt = np.argsort(cells, kind="stable")
        pick = srt[np.linspace(0, len(srt) // 2, 6).astype(int)]  # short to medium pairs
        qs = [w.as_ref_codes(int(k))[0] for k in pick]
        ts = [w.as_ref_codes(int(k))[1] for k in pick]
        exp = O().sw_align_pairs(qs, ts, np.asarray(w.band_offset)[pick].astype(np.int64),
                               w.half_band, w.match, w.mismatch, w.gap_open, w.gap_extend,
                               threads=4)
        for e, k in zip(exp, pick):
            assert int(d[f"{name}_score"][k]) == e["score"], (name, k)
            assert int(d[f"{name}_qend"][k]) == e["read_end"], (name, k)
            assert int(d[f"{name}_rs"][k]) == e["read_start"], (name, k)
            assert int(d[f"{name}_cigar"][k]) == D.cigar_digest(D.cigar_runs(e["cigar"])), (name, k)
    wc = WL.chains()
    assert D.chains_input_digest(wc) == int(d["chains_input"][0])
    nk, best, dig = D.chains_reference(wc, wc.max_lookback, wc.max_gap, reads=40)
    assert np.array_equal(nk[:40], d["chains_nkept"][:40])
    assert np.array_equal(dig[:40], d["chains_digest"][:40])


def test_align_batch_matches_per_pair_driver():
    """oracle_sw_align_batch (the C-threaded CPU comparator) gives the same
    results as the per-pair driver sw_align_pairs."""
    rng = np.random.default_rng(11)
    qs, ts, offs = [], [], []
    for _ in range(40):
        n = int(rng.integers(1, 400))
        q = rng.integers(1, 5, size=n).astype(np.int16)
        t = np.concatenate([rng.integers(1, 5, size=int(rng.integers(0, 30))).astype(np.int16), q,
                            rng.integers(1, 5, size=int(rng.integers(0, 30))).astype(np.int16)])
        t[rng.integers(0, len(t), size=max(1, len(t) // 20))] = rng.integers(1, 5)
        qs.append(q)
        ts.append(t)
        offs.append(int(rng.integers(-5, 40)))
    for hb in (8, 32, 100):
        exp = O().sw_align_pairs(qs, ts, offs, hb, 2, 4, 4, 2)
        got = O().sw_align_batch(qs, ts, offs, hb, 2, 4, 4, 2, threads=4)
        for p, e in enumerate(exp):
            assert int(got["score"][p]) == e["score"], (hb, p)
            assert int(got["read_end"][p]) == e["read_end"], (hb, p)
            assert int(got["ref_end"][p]) == e["ref_end"], (hb, p)
            assert int(got["read_start"][p]) == e["read_start"], (hb, p)
            assert int(got["ref_start"][p]) == e["ref_start"], (hb, p)
            assert int(got["n_cols"][p]) == sum(n for _, n in e["cigar"]), (hb, p)

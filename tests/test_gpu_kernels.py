"""GPU parity of the HIP kernels against the CPU oracle (bit-exact).

Called through the C ABI (graphmambaformer_amd.kernels -> libgmfa.so).  Every
case prints its name first so a failing run names it.
"""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.golden.load import fhex, load

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def K():
    from graphmambaformer_amd import kernels

    return kernels


def _say(*a):
    print(*a, flush=True)


def _random_pair(rng, n, with_n=False):
    q = rng.integers(1, 5, size=n).astype(np.int16)
    t = q.tolist()
    for _ in range(int(rng.integers(0, max(2, n // 15)))):
        p = int(rng.integers(0, len(t)))
        k = rng.integers(0, 3)
        if k == 0:
            t[p] = int(rng.integers(1, 5))
        elif k == 1 and len(t) > 3:
            del t[p:p + int(rng.integers(1, 4))]
        else:
            t[p:p] = [int(rng.integers(1, 5))] * int(rng.integers(1, 4))
    t = rng.integers(1, 5, size=int(rng.integers(0, 20))).tolist() + t
    t = t + rng.integers(1, 5, size=int(rng.integers(0, 20))).tolist()
    t = np.array(t, np.int16)
    if with_n:
        for _ in range(2):
            q[int(rng.integers(0, len(q)))] = 5
            t[int(rng.integers(0, len(t)))] = 5
    return q, t


def _gpu_align(qs, ts, offs, hb, scoring, tile=None, traceback=True, fused=False):
    k = K()
    inp = k.upload_pairs(qs, ts, offs, DEV, alphabet="ref")
    t = k.resolve_tile(hb, *tile) if tile else None
    fwd = k.sw_forward(inp, *scoring, hb, traceback=traceback, tile=t, cigar=fused)
    out = dict(score=fwd.score.cpu().numpy(), qend=fwd.query_end.cpu().numpy(),
               tend=fwd.target_end.cpu().numpy())
    if traceback:
        tbk = fwd.walk if fused else k.sw_traceback(inp, fwd)
        out["cigar"] = k.decode_cigars(tbk)
        out["rs"] = tbk.read_start.cpu().numpy()
        out["ts"] = tbk.ref_start.cpu().numpy()
        out["counts"] = tbk.counts.cpu().numpy()
    return out


def _check_against_oracle(qs, ts, offs, hb, scoring, got, label):
    ma, mi, go, ge = scoring
    exp = O.sw_align_pairs(qs, ts, offs, hb, ma, mi, go, ge, threads=8)
    for p, e in enumerate(exp):
        ctx = (label, p, len(qs[p]), len(ts[p]), int(offs[p]))
        assert int(got["score"][p]) == e["score"], ctx
        assert int(got["qend"][p]) == e["read_end"], ctx
        assert int(got["tend"][p]) == e["ref_end"], ctx
        if "cigar" in got:
            assert got["cigar"][p] == e["cigar"], ctx
            assert int(got["rs"][p]) == e["read_start"], ctx
            assert int(got["ts"][p]) == e["ref_start"], ctx
            assert list(got["counts"][p]) == [e["n_match"], e["n_mismatch"], e["n_insertion"],
                                               e["n_deletion"]], ctx


def test_library_loads_on_device():
    _say("case: library loads on device")
    k = K()
    assert len(k.list_tiles()) > 0
    assert torch.cuda.is_available()


@pytest.mark.parametrize("scoring", [(2, 4, 4, 2), (2, 4, 6, 2), (2, 4, 2, 2), (1, 3, 5, 1)])
def test_sw_random_pairs_default_tile(scoring):
    rng = np.random.default_rng(sum(scoring))
    for hb in (2, 7, 32, 64, 100):
        _say("case: default tile", scoring, "hb", hb)
        qs, ts, offs = [], [], []
        for p in range(24):
            q, t = _random_pair(rng, int(rng.integers(1, 260)), with_n=(p % 5 == 0))
            qs.append(q)
            ts.append(t)
            offs.append(int(rng.integers(-6, 25)))
        got = _gpu_align(qs, ts, offs, hb, scoring, fused=(hb % 2 == 0))
        _check_against_oracle(qs, ts, offs, hb, scoring, got, f"hb{hb}")


def test_sw_every_compiled_tile():
    rng = np.random.default_rng(77)
    scoring = (2, 4, 4, 2)
    for C, G in K().list_tiles():
        cap = C * G
        # exact fit (if odd), partial last lane, and a band far narrower than the tile;
        # a shape compiled for one band only (1 x 64 + the scalar cell: W = 65) runs
        # that band
        hbs = sorted({(cap - 1) // 2, max(1, (cap - 1) // 2 - 3), max(1, cap // 4), cap // 2})
        served = []
        for hb in hbs:
            try:
                K().resolve_tile(hb, C, G)
                served.append(hb)
            except (RuntimeError, ValueError):
                continue
        assert served, (C, G)
        for hb in served:
            qs, ts, offs = [], [], []
            for p in range(12):
                q, t = _random_pair(rng, int(rng.integers(1, 200)), with_n=(p % 4 == 1))
                qs.append(q)
                ts.append(t)
                offs.append(int(rng.integers(-3, 30)))
            # both walks for a one-band shape, else alternating by band
            for fused in ((False, True) if len(served) == 1 else (hb % 2 == 1,)):
                _say("case: tile", C, G, "hb", hb, "fused", fused)
                got = _gpu_align(qs, ts, offs, hb, scoring, tile=(C, G), fused=fused)
                _check_against_oracle(qs, ts, offs, hb, scoring, got, f"tile{C}x{G}hb{hb}")


@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.parametrize("hb", [32, 256, 751])
def test_sw_specialised_band_tiles(hb, fused):
    """Bands with a compiled exact / fixed-remainder tile (65 = 5x13, 513 = 9x57,
    1503 = 24x63 with 15 live registers in the last lane), incl. N codes and
    pairs shorter than the band."""
    rng = np.random.default_rng(hb)
    qs, ts, offs = [], [], []
    for p in range(10):
        n = int(rng.integers(1, 3 * hb + 40)) if p % 3 else int(rng.integers(1, 30))
        q, t = _random_pair(rng, n, with_n=(p % 4 == 2))
        qs.append(q)
        ts.append(t)
        offs.append(int(rng.integers(-5, 25)))
    _say("case: specialised tile hb", hb, K().resolve_tile(hb).cells_per_lane)
    got = _gpu_align(qs, ts, offs, hb, (2, 4, 4, 2), fused=fused)
    _check_against_oracle(qs, ts, offs, hb, (2, 4, 4, 2), got, f"spec{hb}")


@pytest.mark.parametrize("tile", [(24, 64), (32, 64)])
def test_sw_band_1503_on_every_tile_that_holds_it(tile):
    """half_band 751 (the noisy batch's band) on each compiled tile a length
    bin's autotune can pick for it, with raised priority for the longer half
    of the pairs (prio_rows), against the oracle."""
    rng = np.random.default_rng(1503 + tile[0])
    qs, ts, offs = [], [], []
    for p in range(16):
        n = int(rng.integers(1, 4000)) if p % 3 else int(rng.integers(1, 60))
        q, t = _random_pair(rng, n)
        qs.append(q)
        ts.append(t)
        offs.append(int(rng.integers(-800, 800)) if p % 2 else int(rng.integers(-5, 25)))
    _say("case: band 1503 on tile", tile)
    k = K()
    inp = k.upload_pairs(qs, ts, offs, DEV, alphabet="ref")
    lpt = torch.from_numpy(np.argsort(-inp.qlen_host, kind="stable").astype(np.int32)).to(DEV)
    fwd = k.sw_forward(inp, 2, 4, 4, 2, 751, traceback=True, order=lpt, cigar=True,
                       tile=k.resolve_tile(751, *tile),
                       prio_rows=int(np.median(inp.qlen_host)))
    tbk = fwd.walk
    got = dict(score=fwd.score.cpu().numpy(), qend=fwd.query_end.cpu().numpy(),
               tend=fwd.target_end.cpu().numpy(), cigar=k.decode_cigars(tbk),
               rs=tbk.read_start.cpu().numpy(), ts=tbk.ref_start.cpu().numpy(),
               counts=tbk.counts.cpu().numpy())
    _check_against_oracle(qs, ts, offs, 751, (2, 4, 4, 2), got, f"w1503_{tile}")


@pytest.mark.parametrize("fused", [False, True])
def test_sw_scalar_cell_tile_matches_oracle_and_9x64(fused):
    """W = 513 as 8 x 64 lanes plus the scalar cell (diagonal 512 on the scalar
    unit, plane lane 64): every output equals the oracle's and the 9 x 64
    tile's, row keys included; pairs shorter than the band, N codes, band
    offsets that push diagonal 512 in and out of the matrix."""
    rng = np.random.default_rng(513)
    k = K()
    qs, ts, offs = [], [], []
    for p in range(24):
        n = int(rng.integers(1, 1400)) if p % 3 else int(rng.integers(1, 40))
        q, t = _random_pair(rng, n, with_n=(p % 4 == 2))
        qs.append(q)
        ts.append(t)
        offs.append(int(rng.integers(-300, 300)) if p % 2 else int(rng.integers(-5, 25)))
    _say("case: scalar-cell tile 8x64+1, fused", fused)
    t8 = k.resolve_tile(256, 8, 64)
    assert (t8.cells_per_lane, t8.lanes_stored) == (8, 65)
    got = _gpu_align(qs, ts, offs, 256, (2, 4, 4, 2), tile=(8, 64), fused=fused)
    _check_against_oracle(qs, ts, offs, 256, (2, 4, 4, 2), got, "xs513")
    inp = k.upload_pairs(qs, ts, offs, DEV, alphabet="ref")
    keys = []
    for C, G in ((8, 64), (9, 64)):
        f = k.sw_forward(inp, 2, 4, 4, 2, 256, traceback=True, rowkeys=True,
                         tile=k.resolve_tile(256, C, G))
        keys.append((f.score.cpu().numpy(), f.rowkey.cpu().numpy(), f.rowkey_off_host))
    assert np.array_equal(keys[0][0], keys[1][0])
    assert np.array_equal(keys[0][2], keys[1][2])
    assert np.array_equal(keys[0][1], keys[1][1])


@pytest.mark.parametrize("scoring", [(2, 4, 4, 2), (2, 4, 2, 2), (3, 2, 9, 1)])
def test_sw_skew_tile_edges(scoring):
    """The skewed tile (2 diagonals per lane, cells along anti-diagonals) on
    bands from one diagonal to its 125-diagonal limit: pairs of 1 base, pairs
    shorter than the band, offsets that put the band's edge far left or right
    of the matrix, N codes; both walks, against the oracle."""
    rng = np.random.default_rng(9 + sum(scoring))
    for hb in (0, 1, 4, 31, 32, 62):
        qs, ts, offs = [], [], []
        for p in range(20):
            n = 1 if p == 0 else (int(rng.integers(1, 12)) if p % 4 == 1 else
                                  int(rng.integers(1, 400)))
            q, t = _random_pair(rng, n, with_n=(p % 5 == 2))
            qs.append(q)
            ts.append(t)
            offs.append(int(rng.integers(-hb - 20, hb + 40)))
        for fused in (False, True):
            _say("case: skew tile", scoring, "hb", hb, "fused", fused)
            got = _gpu_align(qs, ts, offs, hb, scoring, tile=(2, 63), fused=fused)
            _check_against_oracle(qs, ts, offs, hb, scoring, got, f"skew{hb}")


def test_sw_skew_fused_walk_paths():
    """The skewed tile's fused walk on both sides of its limits: sequence
    windows loaded with the forward pass (pairs up to ~62 words) or by the
    walk, CIGARs of up to 64 runs kept in a register or spilled to the output
    and reversed there, walk steps of up to 64 cells; against the oracle."""
    rng = np.random.default_rng(31)
    scoring = (2, 4, 4, 2)
    for hb in (16, 32, 62):
        qs, ts, offs = [], [], []
        for n in (150, 191, 192, 193, 367, 368, 369, 990, 1300):
            q = rng.integers(1, 5, size=n).astype(np.int16)
            t = q.copy()
            # a substitution every ~4 bases: well over 64 runs
            sub = rng.choice(n, size=max(1, n // 4), replace=False)
            t[sub] = (t[sub] % 4) + 1
            qs.append(q)
            ts.append(t)
            offs.append(hb)
        for n in (150, 368, 369, 1300):
            q, t = _random_pair(rng, n)
            qs.append(q)
            ts.append(t)
            offs.append(int(rng.integers(0, hb + 1)))
        _say("case: skew fused walk paths hb", hb)
        got = _gpu_align(qs, ts, offs, hb, scoring, tile=(2, 63), fused=True)
        _check_against_oracle(qs, ts, offs, hb, scoring, got, f"skewpaths{hb}")


@pytest.mark.parametrize("tile", [(2, 63), (1, 64), (8, 64), (5, 16)])
def test_sw_grid_smaller_than_batch_claims_the_rest(tile):
    """A persistent grid capped below the batch (max_waves): each wavefront's
    first item is its workgroup index, the rest come from the work counter
    after the grid's; a processing order is read.  Two launches in a row on
    one workspace (the counter must be back at zero after the first), both
    against the oracle."""
    k = K()
    rng = np.random.default_rng(71 + tile[0])
    hb = 32 if tile != (8, 64) else 100
    qs, ts, offs = [], [], []
    for p in range(300):
        q, t = _random_pair(rng, int(rng.integers(1, 250)))
        qs.append(q)
        ts.append(t)
        offs.append(int(rng.integers(0, hb + 10)))
    inp = k.upload_pairs(qs, ts, offs, DEV, alphabet="ref")
    tl = k.resolve_tile(hb, *tile)
    ws = k.make_workspace(inp, tl, cigar=True)
    cells = np.array([len(q) for q in qs])
    order = torch.from_numpy(np.argsort(-cells, kind="stable").astype(np.int32)).to(DEV)
    for rep in range(2):
        _say("case: capped grid, tile", tile, "launch", rep)
        fwd = k.sw_forward(inp, 2, 4, 4, 2, hb, traceback=True, tile=tl, workspace=ws,
                           cigar=True, order=order, max_waves=37)
        got = dict(score=fwd.score.cpu().numpy(), qend=fwd.query_end.cpu().numpy(),
                   tend=fwd.target_end.cpu().numpy(), cigar=k.decode_cigars(fwd.walk),
                   rs=fwd.walk.read_start.cpu().numpy(), ts=fwd.walk.ref_start.cpu().numpy(),
                   counts=fwd.walk.counts.cpu().numpy())
        _check_against_oracle(qs, ts, offs, hb, (2, 4, 4, 2), got, f"capped{tile}-{rep}")
        assert int(ws["counter"].abs().sum()) == 0, "work counter left at zero"


def test_sw_short_config_full():
    """1,024 x 150 bp pairs, hb 32 at offset 32, scoring 2/-4/-4/-2 (whole config)."""
    from graphmambaformer_amd import workloads

    _say("case: sw_short full config")
    w = workloads.sw_short()
    qs = [w.as_ref_codes(k)[0] for k in range(w.n_pairs)]
    ts = [w.as_ref_codes(k)[1] for k in range(w.n_pairs)]
    sc = (w.match, w.mismatch, w.gap_open, w.gap_extend)
    got = _gpu_align(qs, ts, w.band_offset, w.half_band, sc, fused=True)
    _check_against_oracle(qs, ts, w.band_offset, w.half_band, sc, got, "sw_short")


def test_sw_long_pairs_sample():
    """A few long pairs at band half-width 256 (long-read configuration shape)."""
    from graphmambaformer_amd import workloads

    _say("case: sw_long sample")
    rng = np.random.default_rng(2)
    qs, ts = [], []
    for L in (5000, 6100, 7777):
        tmpl = rng.integers(0, 4, size=L).astype(np.uint8)
        qs.append(workloads.mutate(rng, tmpl, 0.005, 0.7, 2).astype(np.int16) + 1)
        ts.append(tmpl.astype(np.int16) + 1)
    got = _gpu_align(qs, ts, [0, 0, 0], 256, (2, 4, 4, 2), fused=True)
    _check_against_oracle(qs, ts, [0, 0, 0], 256, (2, 4, 4, 2), got, "sw_long")


def test_golden_sw_ties_with_batch_xdrop():
    """Reference golden vectors, incl. batch-wide x-drop via per-row keys."""
    k = K()
    for n, case in enumerate(load("sw_ties.json")):
        _say("case: golden sw_ties", n)
        ma, mi, go, ge = case["scoring"]
        qs = [np.array(r[:L], np.int16) for r, L in zip(case["query"], case["qlen"])]
        ts = [np.array(r[:L], np.int16) for r, L in zip(case["target"], case["tlen"])]
        inp = k.upload_pairs(qs, ts, case["off"], DEV, alphabet="ref")
        hb = case["half_band"]
        fwd = k.sw_forward(inp, ma, mi, go, ge, hb, traceback=True, rowkeys=True)
        sc, qe, te = O.xdrop_from_rowkeys(fwd.rowkey.cpu().numpy(), fwd.rowkey_off_host,
                                          inp.qlen_host, case["x_drop"], hb, inp.boff_host)
        dev = inp.device
        # the device x-drop (the product path) agrees with the host restatement
        dsc, dqe, dte = k.sw_xdrop(fwd, inp, np.array([0, inp.n_pairs]), case["x_drop"], hb)
        assert np.array_equal(dsc.cpu().numpy(), sc), n
        assert np.array_equal(dqe.cpu().numpy(), qe), n
        assert np.array_equal(dte.cpu().numpy(), te), n
        tbk = k.sw_traceback(inp, fwd, torch.from_numpy(sc).to(dev), torch.from_numpy(qe).to(dev),
                             torch.from_numpy(te).to(dev))
        cig = k.decode_cigars(tbk)
        rs = tbk.read_start.cpu().numpy()
        tsr = tbk.ref_start.cpu().numpy()
        for b, exp in enumerate(case["out"]):
            assert int(sc[b]) == int(exp["score"]), (n, b)
            assert int(qe[b]) == exp["query_end"], (n, b)
            assert int(te[b]) == exp["target_end"], (n, b)
            assert [list(x) for x in cig[b]] == [list(x) for x in exp["cigar"]], (n, b)
            assert int(rs[b]) == exp["read_start"] and int(tsr[b]) == exp["ref_start"], (n, b)


def test_chain_dp_f64_golden():
    k = K()
    data = load("chain_dp.json")
    cases = data["cases"]
    _say("case: chain f64 golden", len(cases))
    for n, case in enumerate(cases):
        lut = O.penalty_lut(case["gap_open"], case["gap_extend"], case["log_coeff"],
                            case["max_gap"])
        d = torch.device(DEV)
        re_ = torch.tensor(case["read_end"], dtype=torch.int64, device=d)
        fe = torch.tensor(case["ref_end"], dtype=torch.int64, device=d)
        w = torch.tensor(fhex(case["weight"]), dtype=torch.float64, device=d)
        seg = torch.tensor([0, len(case["read_end"])], dtype=torch.int64, device=d)
        f, par = k.chain_dp(re_, fe, w, seg, case["max_lookback"], case["max_gap"],
                            pen_lut=torch.from_numpy(lut).to(d))
        assert np.array_equal(f.cpu().numpy(), fhex(case["f"])), n
        assert par.cpu().numpy().tolist() == case["parent"], n


def test_chain_dp_i32_segments_vs_oracle():
    from graphmambaformer_amd import workloads

    _say("case: chain i32 segments")
    k = K()
    w = workloads.chains(n_reads=96, seed=1)
    d = torch.device(DEV)
    f, par = k.chain_dp(torch.from_numpy(w.read_end).to(d), torch.from_numpy(w.ref_end).to(d),
                        torch.from_numpy(w.weight).to(d), torch.from_numpy(w.seg_off).to(d),
                        w.max_lookback, w.max_gap, gap_open=w.gap_open, gap_extend=w.gap_extend)
    f = f.cpu().numpy()
    par = par.cpu().numpy()
    for s in range(w.n_reads):
        a, b = w.seg_off[s], w.seg_off[s + 1]
        ef, ep = O.chain_dp_i32(w.read_end[a:b], w.ref_end[a:b], w.weight[a:b], w.max_lookback,
                                w.max_gap, w.gap_open, w.gap_extend)
        assert np.array_equal(f[a:b], ef), s
        assert np.array_equal(par[a:b], ep), s


def test_chain_dp_coordinate_ranges_vs_oracle():
    """Reads whose coordinates sit far from zero (rebased to 32 bits in the
    kernel) and reads whose anchors span more than 2^30 (the 64-bit form),
    read_end and ref_end alike, in both profiles against the oracle."""
    from graphmambaformer_amd import workloads

    _say("case: chain coordinate ranges")
    k = K()
    w = workloads.chains(n_reads=32, seed=5)
    re_, fe = w.read_end.copy(), w.ref_end.copy()
    for s in range(w.n_reads):
        a, b = w.seg_off[s], w.seg_off[s + 1]
        if s % 4 == 1:
            fe[a:b] += 1 << 40
            re_[a:b] += 1 << 33
        elif s % 4 == 2:
            fe[b - 1] += 3 << 30  # the last anchor (largest ref_end) far away
        elif s % 4 == 3:
            re_[a] += 5 << 30
    d = torch.device(DEV)
    so = torch.from_numpy(w.seg_off).to(d)
    tre, tfe = torch.from_numpy(re_).to(d), torch.from_numpy(fe).to(d)
    f, par = k.chain_dp(tre, tfe, torch.from_numpy(w.weight).to(d), so, w.max_lookback, w.max_gap,
                        gap_open=w.gap_open, gap_extend=w.gap_extend)
    w64 = np.random.default_rng(5).integers(5, 30, size=len(re_)).astype(np.float64)
    lut = O.penalty_lut(6, 0.05, 0.5, w.max_gap)
    f64, p64 = k.chain_dp(tre, tfe, torch.from_numpy(w64).to(d), so, w.max_lookback, w.max_gap,
                          pen_lut=torch.from_numpy(lut).to(d))
    f, par, f64, p64 = f.cpu().numpy(), par.cpu().numpy(), f64.cpu().numpy(), p64.cpu().numpy()
    for s in range(w.n_reads):
        a, b = w.seg_off[s], w.seg_off[s + 1]
        ef, ep = O.chain_dp_i32(re_[a:b], fe[a:b], w.weight[a:b], w.max_lookback, w.max_gap,
                                w.gap_open, w.gap_extend)
        assert np.array_equal(f[a:b], ef) and np.array_equal(par[a:b], ep), s
        ef, ep = O.chain_dp_f64(re_[a:b], fe[a:b], w64[a:b], w.max_lookback, w.max_gap, lut)
        assert np.array_equal(f64[a:b], ef) and np.array_equal(p64[a:b], ep), s


@pytest.mark.parametrize("lookback", [65, 100, 128, 129, 200, 256])
def test_chain_dp_wide_lookback_vs_oracle(lookback):
    """Lookbacks beyond one wavefront (2 or 4 windows of 64 predecessors):
    both profiles against the oracle, read by read.  Equal weights make many
    ties, which must go to the smallest predecessor index; the float64
    profile adds random weights and a graph bonus (chaining.py:129-180)."""
    from graphmambaformer_amd import workloads

    _say("case: chain wide lookback", lookback)
    k = K()
    w = workloads.chains(n_reads=48, seed=lookback)
    d = torch.device(DEV)
    re_, fe = torch.from_numpy(w.read_end).to(d), torch.from_numpy(w.ref_end).to(d)
    so = torch.from_numpy(w.seg_off).to(d)
    f, par = k.chain_dp(re_, fe, torch.from_numpy(w.weight).to(d), so, lookback, w.max_gap,
                        gap_open=w.gap_open, gap_extend=w.gap_extend)
    f, par = f.cpu().numpy(), par.cpu().numpy()
    rng = np.random.default_rng(lookback)
    w64 = rng.integers(5, 30, size=len(w.read_end)).astype(np.float64)
    bonus = rng.normal(0.0, 2.0, size=(len(w.read_end), lookback))
    lut = O.penalty_lut(6, 0.05, 0.5, w.max_gap)
    f64, p64 = k.chain_dp(re_, fe, torch.from_numpy(w64).to(d), so, lookback, w.max_gap,
                          pen_lut=torch.from_numpy(lut).to(d),
                          bonus=torch.from_numpy(np.ascontiguousarray(bonus)).to(d))
    f64, p64 = f64.cpu().numpy(), p64.cpu().numpy()
    for s in range(w.n_reads):
        a, b = w.seg_off[s], w.seg_off[s + 1]
        ef, ep = O.chain_dp_i32(w.read_end[a:b], w.ref_end[a:b], w.weight[a:b], lookback,
                                w.max_gap, w.gap_open, w.gap_extend)
        assert np.array_equal(f[a:b], ef), s
        assert np.array_equal(par[a:b], ep), s
        ef, ep = O.chain_dp_f64(w.read_end[a:b], w.ref_end[a:b], w64[a:b], lookback, w.max_gap,
                                lut, bonus=bonus[a:b])
        assert np.array_equal(f64[a:b], ef), s
        assert np.array_equal(p64[a:b], ep), s


def test_sw_edge_cases_vs_oracle():
    """Empty query or target, a band of one diagonal, bands lying entirely
    beside the target, all-N sequences, one-base pairs."""
    rng = np.random.default_rng(31)
    qs, ts, offs = [], [], []
    base = rng.integers(1, 5, size=120).astype(np.int16)

    def add(q, t, o):
        qs.append(np.asarray(q, np.int16))
        ts.append(np.asarray(t, np.int16))
        offs.append(int(o))

    add([], base[:50], 0)
    add(base[:50], [], 0)
    add([], [], 0)
    add(base[:1], base[:1], 0)
    add(base[:1], base[1:2], 0)
    add(base[:60], base[:60], 500)    # band far right of the target
    add(base[:60], base[:60], -500)   # band far left
    add(np.full(40, 5), np.full(40, 5), 0)
    add(base[:80], np.concatenate([base[:40], np.full(5, 5), base[40:80]]), 3)
    for hb in (0, 1, 3, 64):
        _say("case: SW edge cases hb", hb)
        for fused in (False, True):
            got = _gpu_align(qs, ts, offs, hb, (2, 4, 4, 2), fused=fused)
            _check_against_oracle(qs, ts, offs, hb, (2, 4, 4, 2), got, f"edge{hb}")


@pytest.mark.parametrize("with_n", [False, True])
def test_walk_window_edges_vs_oracle(with_n):
    """The wavefront walk reads its planes through a 4-block x 16-diagonal
    window (plus the 4 blocks above) and its sequences through 64-word
    windows.  Cases at the windows' edges: one gap of 12-40 bases per pair
    (the walk leaves the window sideways), pairs of 3-60 rows (the window
    ahead reaches past block 0), pairs packed back to back so that sequence
    offsets fall at every position inside a 16-base word, and N bases (the
    N-mask kernels and their mask windows).  Fused walk and standalone walk
    against the oracle, bit for bit."""
    rng = np.random.default_rng(77 + int(with_n))
    qs, ts, offs = [], [], []
    for k in range(64):
        n = int(rng.integers(3, 61)) if k % 4 == 0 else int(rng.integers(100, 700))
        q = rng.integers(1, 5, size=n).astype(np.int16)
        t = q.tolist()
        g = int(rng.integers(12, 41))
        p = int(rng.integers(0, n))
        if k % 2:
            t[p:p] = rng.integers(1, 5, size=g).tolist()  # reference insertion: D run
        else:
            del t[p:p + g]  # reference deletion: I run
        for _ in range(int(rng.integers(0, 6))):
            if t:
                t[int(rng.integers(0, len(t)))] = int(rng.integers(1, 5))
        t = rng.integers(1, 5, size=int(rng.integers(0, 17))).tolist() + t
        t = np.array(t, np.int16)
        if with_n and k % 3 == 0:
            q[int(rng.integers(0, n))] = 5
            if len(t):
                t[int(rng.integers(0, len(t)))] = 5
        qs.append(q)
        ts.append(t)
        offs.append(int(rng.integers(-8, 9)))
    for hb in (48, 200):
        for fused in (False, True):
            _say("case: walk window edges, N", with_n, "hb", hb, "fused", fused)
            got = _gpu_align(qs, ts, offs, hb, (2, 4, 4, 2), fused=fused)
            _check_against_oracle(qs, ts, offs, hb, (2, 4, 4, 2), got, f"win{hb}{fused}")


def test_xdrop_small_groups_vs_host_restatement():
    """Groups of 1-16 pairs (the row-parallel x-drop kernel: 64 rows per tile,
    several tiles per pair, stops inside and at the edges of tiles) against
    the host restatement of the x-drop over the same row keys, and against
    the oracle's x-drop for a few groups."""
    k = K()
    rng = np.random.default_rng(43)
    qs, ts, offs = [], [], []
    for p in range(160):
        n = int(rng.integers(1, 420))
        q, t = _random_pair(rng, n)
        if p % 3 == 0:  # a diverging tail makes the x-drop fire mid-pair
            t = np.concatenate([t[: n // 2], rng.integers(1, 5, size=n).astype(np.int16)])
        qs.append(q)
        ts.append(t)
        offs.append(int(rng.integers(0, 20)))
    inp = k.upload_pairs(qs, ts, offs, DEV, alphabet="ref")
    hb = 24
    fwd = k.sw_forward(inp, 2, 4, 4, 2, hb, traceback=True, rowkeys=True)
    keys = fwd.rowkey.cpu().numpy()
    sizes = []
    while sum(sizes) < len(qs):
        sizes.append(min(int(rng.integers(1, 17)), len(qs) - sum(sizes)))
    go = np.concatenate([[0], np.cumsum(sizes)])
    assert go.max() == len(qs) and max(sizes) <= 16
    for xd in (0.0, 3.0, 10.0, 40.0, 1e9):
        _say("case: x-drop small groups", len(sizes), xd)
        sc, qe, te = k.sw_xdrop(fwd, inp, go, xd, hb)
        esc, eqe, ete = O.xdrop_from_rowkeys(keys, fwd.rowkey_off_host, inp.qlen_host, xd, hb,
                                             inp.boff_host, go)
        assert np.array_equal(sc.cpu().numpy(), esc), xd
        assert np.array_equal(qe.cpu().numpy(), eqe), xd
        assert np.array_equal(te.cpu().numpy(), ete), xd
    for a, b in list(zip(go[:-1], go[1:]))[:6]:
        for xd in (3.0, 10.0):
            M = max(len(x) for x in qs[a:b])
            N = max(len(x) for x in ts[a:b])
            Q = np.full((b - a, M), -1, np.int16)
            T = np.full((b - a, N), -1, np.int16)
            for x in range(a, b):
                Q[x - a, :len(qs[x])] = qs[x]
                T[x - a, :len(ts[x])] = ts[x]
            res = O.sw_banded(Q, T, [len(x) for x in qs[a:b]], [len(x) for x in ts[a:b]],
                              offs[a:b], hb, 2, 4, 4, 2, x_drop=xd)
            sc, qe, te = k.sw_xdrop(fwd, inp, go, xd, hb)
            assert np.array_equal(sc.cpu().numpy()[a:b], res["score"]), (a, b, xd)
            assert np.array_equal(qe.cpu().numpy()[a:b], res["query_end"]), (a, b, xd)
            assert np.array_equal(te.cpu().numpy()[a:b], res["target_end"]), (a, b, xd)


def test_xdrop_wide_groups_vs_oracle():
    """Groups above 64 pairs (the 256-thread x-drop kernel) and mixed group
    sizes in one call, against the oracle's x-drop on the same pairs."""
    k = K()
    rng = np.random.default_rng(41)
    qs, ts, offs = [], [], []
    for p in range(300):
        q, t = _random_pair(rng, int(rng.integers(20, 160)))
        qs.append(q)
        ts.append(t)
        offs.append(int(rng.integers(0, 20)))
    inp = k.upload_pairs(qs, ts, offs, DEV, alphabet="ref")
    hb = 24
    fwd = k.sw_forward(inp, 2, 4, 4, 2, hb, traceback=True, rowkeys=True)
    for groups in ([0, 300], [0, 10, 200, 201, 300], [0, 64, 65, 300]):
        for xd in (6.0, 30.0, 0.0):
            _say("case: x-drop groups", groups, xd)
            go = np.array(groups)
            sc, qe, te = k.sw_xdrop(fwd, inp, go, xd, hb)
            for a, b in zip(go[:-1], go[1:]):
                M = max(len(x) for x in qs[a:b])
                N = max(len(x) for x in ts[a:b])
                Q = np.full((b - a, M), -1, np.int16)
                T = np.full((b - a, N), -1, np.int16)
                for x in range(a, b):
                    Q[x - a, :len(qs[x])] = qs[x]
                    T[x - a, :len(ts[x])] = ts[x]
                res = O.sw_banded(Q, T, [len(x) for x in qs[a:b]], [len(x) for x in ts[a:b]],
                                  offs[a:b], hb, 2, 4, 4, 2, x_drop=xd)
                assert np.array_equal(sc.cpu().numpy()[a:b], res["score"]), (groups, xd)
                assert np.array_equal(qe.cpu().numpy()[a:b], res["query_end"]), (groups, xd)
                assert np.array_equal(te.cpu().numpy()[a:b], res["target_end"]), (groups, xd)

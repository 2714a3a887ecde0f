"""GPU tests of continuous serving (gmfa_serve_run): every request of a mixed
stream completes, and its score, ends and CIGAR are bit-identical to the
oracle's on the same pair, whatever the arrival pattern or the class mix
(alone vs mixed, paced vs all-at-once)."""

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
SC = (2, 4, 4, 2)


def _say(*a):
    print(*a, flush=True)


def _small_mix(seed=11):
    """A 2 Mbp reference and three small pools (short / long / noisy bands)."""
    from graphmambaformer_amd import workloads as WL

    rng = np.random.default_rng(seed)
    ref = rng.integers(0, 4, size=2_000_000, dtype=np.uint8)
    pools = []
    pos = rng.integers(0, len(ref) - 214, size=48)
    qs = [WL.mutate(rng, ref[p + 32:p + 182], 0.0, 0.0, 3, sub_rate=0.01, indel_rate=0.002)
          for p in pos]
    pools.append(("short", 32, qs, pos, np.full(48, 214), np.full(48, 32)))
    lens = rng.integers(800, 4000, size=12)
    pos = np.array([rng.integers(0, len(ref) - L) for L in lens])
    qs = [WL.mutate(rng, ref[p:p + L], 0.005, 0.7, 2) for p, L in zip(pos, lens)]
    pools.append(("long", 256, qs, pos, lens, np.zeros(12)))
    lens = rng.integers(300, 3000, size=8)
    pos = np.array([rng.integers(0, len(ref) - L) for L in lens])
    qs = [WL.mutate(rng, ref[p:p + L], 0.06, 0.6, 3, homopolymer_weight=2.0)
          for p, L in zip(pos, lens)]
    pools.append(("noisy", 751, qs, pos, lens, np.zeros(8)))
    return ref, pools


def _expected(ref, pool):
    name, hb, qs, pos, lens, offs = pool
    q1 = [q.astype(np.int16) + 1 for q in qs]
    t1 = [ref[p:p + L].astype(np.int16) + 1 for p, L in zip(pos, lens)]
    return O.sw_align_pairs(q1, t1, np.asarray(offs, np.int64), hb, *SC, threads=8)


def _build(ref_dev, pools_dev, picks, arrivals, slots=(16, 16, 8)):
    from graphmambaformer_amd import serving as S

    out = []
    for c, (pool, pk, arr) in enumerate(zip(pools_dev, picks, arrivals)):
        out.append(S.make_class(f"class{c}", ref_dev, pool, pk, arr, SC, slots[c]))
    return out


def _check(cls, exp, label):
    from graphmambaformer_amd import kernels as K

    assert np.all(cls.complete_s >= 0), label
    score = cls.score.cpu().numpy()
    qe = cls.qend.cpu().numpy()
    te = cls.tend.cpu().numpy()
    cig = K.decode_cigars(cls.walk)
    rs = cls.walk.read_start.cpu().numpy()
    ts = cls.walk.ref_start.cpu().numpy()
    for k, pidx in enumerate(cls.picks):
        e = exp[int(pidx)]
        ctx = (label, k, int(pidx))
        assert int(score[k]) == e["score"], ctx
        assert int(qe[k]) == e["read_end"] and int(te[k]) == e["ref_end"], ctx
        assert cig[k] == e["cigar"], ctx
        assert int(rs[k]) == e["read_start"] and int(ts[k]) == e["ref_start"], ctx


@pytest.fixture(scope="module")
def mix():
    from graphmambaformer_amd import serving as S

    ref, pools = _small_mix()
    ref_dev = S.ResidentReference.upload(ref, DEV)
    pools_dev = [S.RequestPool.upload(qs, pos, lens, offs, hb, DEV)
                 for (_, hb, qs, pos, lens, offs) in pools]
    exp = [_expected(ref, p) for p in pools]
    return ref_dev, pools_dev, exp


def test_serving_paced_mixed_stream_matches_oracle(mix):
    from graphmambaformer_amd import serving as S

    ref_dev, pools_dev, exp = mix
    rng = np.random.default_rng(3)
    picks = [rng.integers(0, len(e), size=n) for e, n in zip(exp, (120, 40, 16))]
    arrivals = [np.sort(rng.uniform(0.0, 0.05, size=len(p))) for p in picks]
    classes = _build(ref_dev, pools_dev, picks, arrivals)
    _say("case: serving paced mixed stream", [c.n for c in classes])
    res = S.run(classes, timeout_s=30.0)
    assert res.completed == sum(c.n for c in classes)
    assert not res.timed_out and not res.waves_gave_up
    for c, e in zip(classes, exp):
        _check(c, e, c.name)
        assert np.all(c.complete_s >= c.arrival_s)


def test_serving_all_at_once_and_alone_identical(mix):
    """Every request arrives at t = 0 (saturation); then the noisy class alone.
    Outputs equal the oracle's in both, i.e. independent of the mix."""
    from graphmambaformer_amd import serving as S

    ref_dev, pools_dev, exp = mix
    picks = [np.arange(len(e)) for e in exp]
    arrivals = [np.zeros(len(p)) for p in picks]
    classes = _build(ref_dev, pools_dev, picks, arrivals, slots=(4, 8, 3))
    _say("case: serving saturation", [c.n for c in classes])
    res = S.run(classes, timeout_s=30.0)
    assert res.completed == sum(c.n for c in classes) and not res.timed_out
    for c, e in zip(classes, exp):
        _check(c, e, c.name + "/sat")
    _say("case: serving noisy class alone")
    alone = S.make_class("noisy-alone", ref_dev, pools_dev[2], picks[2][::-1].copy(),
                         np.linspace(0.0, 0.01, len(picks[2])), SC, 2)
    res = S.run([alone], timeout_s=30.0)
    assert res.completed == alone.n and not res.timed_out
    _check(alone, exp[2], "noisy-alone")


def test_serving_rejects_bad_inputs(mix):
    from graphmambaformer_amd import kernels as K
    from graphmambaformer_amd import serving as S

    ref_dev, pools_dev, exp = mix
    _say("case: serving invalid inputs")
    with pytest.raises(ValueError):
        S.make_class("x", ref_dev, pools_dev[0], np.array([0, 1]), np.zeros(3), SC, 2)
    with pytest.raises(ValueError):  # a 16-lane tile holds several pairs per wavefront
        S.make_class("x", ref_dev, pools_dev[0], np.array([0]), np.zeros(1), SC, 2,
                     tile=K.resolve_tile(32, 5, 16))
    c = S.make_class("x", ref_dev, pools_dev[0], np.array([0, 1]), np.array([0.0, 1.0]), SC, 2)
    c.arrival_s = np.array([1.0, 0.0])  # not nondecreasing: rejected before any launch
    with pytest.raises(ValueError):
        S.run([c], timeout_s=5.0)
    c = S.make_class("x", ref_dev, pools_dev[1], np.array([0]), np.zeros(1), SC, 2)
    c.slab_words = 1  # too small for the query
    with pytest.raises(ValueError):
        S.run([c], timeout_s=5.0)
    torch.cuda.synchronize()


def _online_classes(pools, slots, rings, cap=None):
    from graphmambaformer_amd import serving as S

    return [S.OnlineClass(half_band=hb, max_qlen=max(len(q) for q in qs), n_slots=n, ring=r,
                          cigar_cap=cap)
            for (_, hb, qs, _, _, _), n, r in zip(pools, slots, rings)]


def _check_online(results, tickets, exp, label):
    for r in results:
        k = tickets[(r.cls, r.ticket)]
        e = exp[r.cls][k]
        ctx = (label, r.cls, r.ticket, k)
        assert not r.truncated and not r.no_pages, ctx
        assert r.score == e["score"], ctx
        assert r.query_end == e["read_end"] and r.target_end == e["ref_end"], ctx
        assert r.cigar == e["cigar"], ctx
        assert r.read_start == e["read_start"] and r.ref_start == e["ref_start"], ctx


def test_online_session_matches_oracle(mix):
    """Requests submitted one at a time into small rings (ring-full waits and
    wrap-around), a pause longer than the idle timeout (every wave exits; the
    next submit relaunches the class), and every result bit-identical to the
    oracle's."""
    import time

    from graphmambaformer_amd import serving as S

    ref_dev, pools_dev, exp = mix
    _, pools = _small_mix()
    rng = np.random.default_rng(5)
    classes = _online_classes(pools, slots=(8, 8, 4), rings=(16, 8, 4))
    _say("case: online session, 3 classes, rings 16/8/4, idle 0.05 s")
    tickets, results = {}, []
    sess = S.ServeSession(ref_dev, classes, SC, idle_timeout_s=0.05)
    try:
        for phase, n in enumerate((150, 40)):
            order = rng.integers(0, 3, size=n)
            for c in order:
                _, hb, qs, pos, lens, offs = pools[c]
                k = int(rng.integers(0, len(qs)))
                t = sess.submit(int(c), qs[k], int(pos[k]), int(lens[k]), int(offs[k]))
                tickets[(int(c), t)] = k
                results += sess.poll()
            results += sess.wait(len(tickets) - len(results), timeout_s=60.0)
            if phase == 0:
                time.sleep(0.3)  # every wave idles out
        st = sess.stats()
        _say("online stats", st)
        assert st["submitted"] == len(tickets) == st["returned"] and st["in_flight"] == 0
        assert st["launches"] > len(classes)  # relaunched after the pause
    finally:
        rep = sess.close()
    assert rep.completed == len(tickets)
    assert len(results) == len(tickets)
    assert len({(r.cls, r.ticket) for r in results}) == len(tickets)
    _check_online(results, tickets, exp, "online")


def test_online_session_truncation_and_close_drain(mix):
    """A CIGAR longer than cigar_cap comes back flagged, without runs, and the
    rest of the record intact; close() drains requests never polled."""
    from graphmambaformer_amd import serving as S

    ref_dev, pools_dev, exp = mix
    _, pools = _small_mix()
    _say("case: online session, cigar_cap 3, close with unpolled requests")
    classes = _online_classes(pools[:1], slots=(4,), rings=(64,), cap=3)
    sess = S.ServeSession(ref_dev, classes, SC, idle_timeout_s=1.0)
    _, hb, qs, pos, lens, offs = pools[0]
    tickets = {}
    try:
        for k in range(len(qs)):
            tickets[(0, sess.submit(0, qs[k], int(pos[k]), int(lens[k]), int(offs[k])))] = k
        got = sess.wait(len(qs), timeout_s=60.0)
        for r in got:
            e = exp[0][tickets[(0, r.ticket)]]
            assert r.truncated == (len(e["cigar"]) > 3), (r.ticket, e["cigar"])
            assert r.score == e["score"] and r.read_start == e["read_start"]
            if not r.truncated:
                assert r.cigar == e["cigar"]
            else:
                assert r.cigar == []
        for k in range(8):
            sess.submit(0, qs[k], int(pos[k]), int(lens[k]), int(offs[k]))
    finally:
        rep = sess.close()
    assert rep.completed == len(qs) + 8


def test_online_session_rejects_bad_requests(mix):
    from graphmambaformer_amd import serving as S

    ref_dev, pools_dev, exp = mix
    _, pools = _small_mix()
    _say("case: online session invalid requests")
    classes = _online_classes(pools[:1], slots=(2,), rings=(4,))
    with S.ServeSession(ref_dev, classes, SC, idle_timeout_s=1.0) as sess:
        _, hb, qs, pos, lens, offs = pools[0]
        q = qs[0]
        with pytest.raises(ValueError):  # longer than the class's max_qlen
            sess.submit(0, np.zeros(classes[0].max_qlen + 1, np.uint8), 0, 10, 0)
        with pytest.raises(ValueError):  # window leaves the reference
            sess.submit(0, q, ref_dev.n_bases - 5, 10, 0)
        with pytest.raises(ValueError):  # code outside 0..3
            sess.submit(0, np.full(10, 7, np.uint8), 0, 10, 0)
        with pytest.raises(ValueError):  # an N without N masks
            sess.submit(0, np.full(10, 5, np.uint8), 0, 10, 0, two_bit=False)
        with pytest.raises(ValueError):
            sess.submit(3, q, 0, 10, 0)  # no such class
        t = sess.submit(0, q, int(pos[0]), int(lens[0]), int(offs[0]))
        r = sess.wait(1)
        assert r[0].ticket == t and r[0].score == exp[0][0]["score"]
    with pytest.raises(RuntimeError):
        sess.submit(0, q, 0, 10, 0)  # closed


def test_online_submit_many_bursts_match_oracle(mix):
    """Bursts through gmfa_serve_submit_many (partial at a full ring, resumed
    after polls) and results read back as arrays; every request bit-identical
    to the oracle's."""
    from graphmambaformer_amd import _lib
    from graphmambaformer_amd import serving as S

    ref_dev, pools_dev, exp = mix
    _, pools = _small_mix()
    _say("case: online submit_many bursts, rings 8/4/4")
    classes = _online_classes(pools, slots=(4, 4, 2), rings=(8, 4, 4))
    flat, base = [], 0
    reqs = []  # (class, pool index, code offset)
    for c, (_, hb, qs, pos, lens, offs) in enumerate(pools):
        for k, q in enumerate(qs):
            reqs.append((c, k, base))
            flat.append(np.asarray(q, np.uint8))
            base += len(q)
    codes = np.concatenate(flat)
    rng = np.random.default_rng(9)
    order = rng.permutation(len(reqs))
    cl = np.array([reqs[i][0] for i in order], np.int32)
    pk = np.array([reqs[i][1] for i in order])
    co = np.array([reqs[i][2] for i in order], np.int64)
    ql = np.array([len(pools[c][2][k]) for c, k in zip(cl, pk)], np.int32)
    to = np.array([pools[c][3][k] for c, k in zip(cl, pk)], np.int64)
    tl = np.array([pools[c][4][k] for c, k in zip(cl, pk)], np.int32)
    bo = np.array([pools[c][5][k] for c, k in zip(cl, pk)], np.int32)
    tickets = {}
    recs, cigs = [], []
    with S.ServeSession(ref_dev, classes, SC, idle_timeout_s=1.0) as sess:
        i, partial = 0, 0
        while i < len(order):
            tk = sess.submit_many(cl[i:], codes, co[i:], ql[i:], to[i:], tl[i:], bo[i:])
            for m, t in enumerate(tk):
                tickets[(int(cl[i + m]), int(t))] = int(pk[i + m])
            partial += len(tk) < len(order) - i
            i += len(tk)
            r, cg = sess.poll_arrays()
            recs.append(r)
            cigs.append(cg)
        while sum(len(r) for r in recs) < len(order):
            r, cg = sess.poll_arrays()
            recs.append(r)
            cigs.append(cg)
    assert partial > 0  # the small rings cut some bursts short
    n = 0
    for r, cg in zip(recs, cigs):
        for x in r:
            e = exp[int(x["cls"])][tickets[(int(x["cls"]), int(x["ticket"]))]]
            assert x["status"] == 0
            runs = cg[x["cigar_off"]:x["cigar_off"] + x["n_cigar"]]
            cig = [(_lib.OP_CHAR[int(v) & 15], int(v) >> 4) for v in runs]
            ctx = (int(x["cls"]), int(x["ticket"]))
            assert x["score"] == e["score"] and cig == e["cigar"], ctx
            assert x["query_end"] == e["read_end"] and x["target_end"] == e["ref_end"], ctx
            assert x["read_start"] == e["read_start"] and x["ref_start"] == e["ref_start"], ctx
            n += 1
    assert n == len(order)
    # an invalid request in a burst stops it there; the ones before are published
    with S.ServeSession(ref_dev, classes, SC, idle_timeout_s=1.0) as sess:
        bad_ql = ql[:3].copy()
        bad_ql[2] = classes[int(cl[2])].max_qlen + 1
        with pytest.raises(ValueError):
            sess.submit_many(cl[:3], codes, co[:3], bad_ql, to[:3], tl[:3], bo[:3])
        assert sess.stats()["submitted"] == 2
        assert len(sess.wait(2)) == 2


def _tight_budgets(pools, factor):
    """Plane budgets holding about `factor` of each class's longest request."""
    from graphmambaformer_amd import serving as S

    out = []
    for (_, hb, qs, _, _, _) in pools:
        tile = S.serve_tile(hb)
        mq = max(len(q) for q in qs)
        blocks = (mq + 7) // 8
        words = blocks * tile.lanes_stored * tile.cells_per_lane
        out.append(int(words * 4 * factor) + 1)
    return out


# 8-row blocks per page: the default (pages of about 1 MB, a few per request)
# and pages small enough that the long and noisy classes' longest requests
# hold more pages than their tiles have stored band lanes (57 and 63), so
# page numbers past every stored lane are looked up during the plane store
PAGE_BLOCKS = {"default": (None, None, None), "small": (1, 2, 2)}


def _check_many_pages(pages_by_class, tag):
    from graphmambaformer_amd import serving as S

    if tag != "small":
        return
    _, pools = _small_mix()
    for c in (1, 2):
        lanes = S.serve_tile(pools[c][1]).lanes_stored
        assert pages_by_class[c] > lanes, (c, pages_by_class[c], lanes)


@pytest.mark.parametrize("tag", sorted(PAGE_BLOCKS))
def test_serving_paged_planes_tight_budget_matches_oracle(mix, tag):
    """Staged serving with the traceback planes in a page pool that holds
    about two of a class's longest requests: requests wait for pages
    (admission by footprint) and every result equals the oracle's."""
    from graphmambaformer_amd import serving as S

    ref_dev, pools_dev, exp = mix
    _, pools = _small_mix()
    budgets = _tight_budgets(pools, 2.5)
    picks = [np.arange(len(e)) for e in exp]
    arrivals = [np.zeros(len(p)) for p in picks]
    classes = [S.make_class(f"paged{c}", ref_dev, pools_dev[c], picks[c], arrivals[c], SC, n,
                            plane_budget_bytes=budgets[c], page_blocks=PAGE_BLOCKS[tag][c])
               for c, n in enumerate((16, 16, 8))]
    _check_many_pages([c.pages.max_pages for c in classes], tag)
    _say(f"case: staged serving, paged planes ({tag} pages)",
         [(c.pages.page_blocks, c.pages.n_pages, c.pages.max_pages) for c in classes])
    res = S.run(classes, timeout_s=60.0)
    assert res.completed == sum(c.n for c in classes) and not res.timed_out
    assert not res.waves_gave_up
    for c, e in zip(classes, exp):
        _check(c, e, c.name)


@pytest.mark.parametrize("tag", sorted(PAGE_BLOCKS))
def test_online_paged_planes_tight_budget_matches_oracle(mix, tag):
    """Online sessions with paged planes and tight budgets: every request
    completes and equals the oracle's."""
    from graphmambaformer_amd import serving as S

    ref_dev, pools_dev, exp = mix
    _, pools = _small_mix()
    budgets = _tight_budgets(pools, 2.5)
    classes = _online_classes(pools, slots=(8, 8, 4), rings=(64, 32, 16))
    for oc, b, pb in zip(classes, budgets, PAGE_BLOCKS[tag]):
        oc.plane_budget_bytes, oc.page_blocks = b, pb
    geo = [S.page_geometry(S.serve_tile(oc.half_band), oc.max_qlen, oc.plane_budget_bytes,
                           oc.page_blocks) for oc in classes]
    _check_many_pages([g.max_pages for g in geo], tag)
    rng = np.random.default_rng(17)
    _say(f"case: online session, paged planes ({tag} pages)",
         [(g.page_blocks, g.n_pages, g.max_pages) for g in geo])
    tickets, results = {}, []
    with S.ServeSession(ref_dev, classes, SC, idle_timeout_s=1.0) as sess:
        for c in rng.integers(0, 3, size=160):
            _, hb, qs, pos, lens, offs = pools[c]
            k = int(rng.integers(0, len(qs)))
            t = sess.submit(int(c), qs[k], int(pos[k]), int(lens[k]), int(offs[k]))
            tickets[(int(c), t)] = k
            results += sess.poll()
        results += sess.wait(len(tickets) - len(results), timeout_s=60.0)
    assert len(results) == len(tickets)
    _check_online(results, tickets, exp, "online-paged")


def _contended_noisy(pools):
    """The noisy class's requests that each need more than half of a pool
    holding exactly the longest one's pages, with that pool's budget: at
    most one of them holds pages at a time."""
    from graphmambaformer_amd import serving as S

    _, hb, qs, _, _, _ = pools[2]
    tile = S.serve_tile(hb)
    mq = max(len(q) for q in qs)
    geo = S.page_geometry(tile, mq, 1 << 40, page_blocks=2)
    budget = geo.max_pages * geo.page_words * 4
    need = np.array([((len(q) + 7) // 8 + 1) // 2 for q in qs])
    big = np.flatnonzero(2 * need > geo.max_pages)
    assert len(big) >= 3
    return big, budget


def test_serving_no_pages_within_bound_completes_without_outputs(mix):
    """Staged serving with a page pool that holds one of the contending
    requests at a time and a 20 us wait bound: the requests that find no
    pages complete with n_cigar = -1 (and the run reports waves that gave
    up); every other request equals the oracle."""
    from graphmambaformer_amd import kernels as K
    from graphmambaformer_amd import serving as S

    ref_dev, pools_dev, exp = mix
    _, pools = _small_mix()
    big, budget = _contended_noisy(pools)
    cls = S.make_class("nopages", ref_dev, pools_dev[2], big, np.zeros(len(big)), SC, len(big),
                       plane_budget_bytes=budget, page_blocks=2, page_wait_s=20e-6)
    _say("case: staged serving, no pages within 20 us", (len(big), cls.pages.n_pages))
    res = S.run([cls], timeout_s=60.0)
    assert res.completed == len(big) and not res.timed_out
    ncig = cls.walk.n_cigar.cpu().numpy()
    failed = np.flatnonzero(ncig < 0)
    assert len(failed) >= 1 and res.waves_gave_up
    score = cls.score.cpu().numpy()
    cig = K.decode_cigars(cls.walk)
    for k, pidx in enumerate(cls.picks):
        if ncig[k] < 0:
            continue
        e = exp[2][int(pidx)]
        assert int(score[k]) == e["score"] and cig[k] == e["cigar"], (k, int(pidx))


def test_online_no_pages_within_bound_is_reported(mix):
    """Online: requests that find no pages within page_wait_s come back with
    no_pages set and no CIGAR; the others equal the oracle."""
    from graphmambaformer_amd import serving as S

    ref_dev, pools_dev, exp = mix
    _, pools = _small_mix()
    big, budget = _contended_noisy(pools)
    _, hb, qs, pos, lens, offs = pools[2]
    oc = S.OnlineClass(half_band=hb, max_qlen=max(len(q) for q in qs), n_slots=len(big),
                       ring=16, plane_budget_bytes=budget, page_blocks=2, page_wait_s=20e-6)
    _say("case: online session, no pages within 20 us", len(big))
    tickets, results = {}, []
    with S.ServeSession(ref_dev, [oc], SC, idle_timeout_s=1.0) as sess:
        for k in big:
            tickets[(0, sess.submit(0, qs[k], int(pos[k]), int(lens[k]), int(offs[k])))] = int(k)
        results += sess.wait(len(big), timeout_s=60.0)
    assert len(results) == len(big)
    failed = [r for r in results if r.no_pages]
    assert len(failed) >= 1
    assert all(r.cigar == [] for r in failed)
    _check_online([r for r in results if not r.no_pages], tickets, [exp[2]], "online-nopages")


def test_paged_budget_below_one_request_is_rejected(mix):
    from graphmambaformer_amd import serving as S

    ref_dev, pools_dev, exp = mix
    _, pools = _small_mix()
    budgets = _tight_budgets(pools, 0.5)
    with pytest.raises(ValueError):
        S.make_class("x", ref_dev, pools_dev[1], np.arange(3), np.zeros(3), SC, 4,
                     plane_budget_bytes=budgets[1])


@pytest.mark.parametrize("paged", [False, True])
@pytest.mark.parametrize("slices", [2, 3, 4])
def test_split_class_matches_oracle_alone_and_in_mix(mix, slices, paged):
    """A staged class whose every request runs as band slices on a workgroup
    (the noisy band, half_band 751): bit-identical to the oracle alone, and in
    a mixed run beside the short and long classes and a one-wavefront noisy
    class holding the same requests."""
    from graphmambaformer_amd import serving as S

    ref_dev, pools_dev, exp = mix
    n = len(exp[2])
    budget = dict(plane_budget_bytes=64 << 20, page_blocks=4) if paged else {}
    _say("case: split serving class alone", slices, "paged", paged)
    alone = S.make_class("noisy-split", ref_dev, pools_dev[2], np.arange(n),
                         np.linspace(0.0, 0.005, n), SC, 3, split_slices=slices, **budget)
    res = S.run([alone], timeout_s=30.0)
    assert res.completed == n and not res.timed_out and not res.waves_gave_up
    _check(alone, exp[2], f"split{slices}-alone")
    _say("case: split serving class in a mix", slices, "paged", paged)
    rng = np.random.default_rng(slices)
    picks = [rng.integers(0, len(e), size=m) for e, m in zip(exp, (60, 20, 8))]
    arrivals = [np.sort(rng.uniform(0.0, 0.02, size=len(p))) for p in picks]
    classes = _build(ref_dev, pools_dev, picks, arrivals)
    sp = rng.permutation(n)
    classes.append(S.make_class("noisy-split", ref_dev, pools_dev[2], sp,
                                np.sort(rng.uniform(0.0, 0.02, size=n)), SC, 4,
                                split_slices=slices, **budget))
    res = S.run(classes, timeout_s=30.0)
    assert res.completed == sum(c.n for c in classes)
    assert not res.timed_out and not res.waves_gave_up
    for c, e in zip(classes, exp + [exp[2]]):
        _check(c, e, c.name + f"/mix{slices}")

"""The served mix at its full size (pipeline.py:355-363 serves each window of a
read through the extension; here the requests of three band classes are
served side by side by persistent kernels over one resident reference):

* a 46 Mbp resident 2-bit reference and the serving pools of 2,048 / 1,024 /
  512 distinct windows (150 bp reads at half_band 32; lognormal(12 kb, 0.4)
  clipped 5-25 kb at half_band 256; lognormal(8 kb, 1.0) clipped 0.5-100 kb
  with 6% errors at half_band 751);
* 4,000 requests drawn 60 / 30 / 10 % from those pools, plus every noisy
  window of 64 kb and more and the 8 longest long-read windows, arriving as a
  Poisson stream;
* traceback planes paged from per-class budgets (256 MB / 8 GB / 32 GB), so
  requests are admitted by their planes' footprint, and 512 / 1,024 / 2,048
  persistent wavefronts per class;

served staged (gmfa_serve_run) and online (gmfa_serve_submit_many / poll).
Every request is checked through the size-independent properties of
tests/alignment_checks.py, the online results equal the staged ones request by
request, a sample (including every window of 64 kb and more) equals the
oracle's, and each sampled request served alone -- staged, and online with
nothing else in flight -- equals its result in the mix."""

import time

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.alignment_checks import check_alignments

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
N_REQ = 4000
SLOTS = (512, 1024, 2048)
PLANE_BUDGET = (256 << 20, 8 << 30, 32 << 30)
BIG_NOISY = 64_000


def _say(*a):
    print(*a, flush=True)


def _runs(arr):
    """BAM runs -> [(op char, length)]."""
    from graphmambaformer_amd import _lib

    return [(_lib.OP_CHAR[int(v) & 15], int(v) >> 4) for v in np.asarray(arr, np.uint32)]


@pytest.fixture(scope="module")
def mix():
    return build_mix()


def build_mix():
    from graphmambaformer_amd import serving as S
    from graphmambaformer_amd import workloads as WL

    _say("case: full-size serving mix: 46 Mbp reference, pools 2048/1024/512")
    m = WL.serving_mix()
    ref = S.ResidentReference.upload(m.reference, DEV)
    pools = [S.RequestPool.upload(p.queries, p.t_start, p.t_len, p.band_offset, p.half_band, DEV)
             for p in m.pools]
    cls, idx = WL.serving_requests(m, N_REQ, seed=21)
    nq = np.array([len(q) for q in m.pools[2].queries])
    big = np.flatnonzero(nq >= BIG_NOISY)
    lq = np.array([len(q) for q in m.pools[1].queries])
    longest = np.argsort(-lq, kind="stable")[:8]
    assert len(big) >= 4, "the noisy pool holds windows of 64 kb and more"
    cls = np.concatenate([cls, np.full(len(big), 2, np.int32), np.full(8, 1, np.int32)])
    idx = np.concatenate([idx, big, longest]).astype(np.int64)
    perm = np.random.default_rng(22).permutation(len(cls))
    cls, idx = cls[perm], idx[perm]
    arrivals = WL.poisson_arrivals(len(cls), 100_000.0, seed=23)
    sc = (m.match, m.mismatch, m.gap_open, m.gap_extend)
    # per request: its query's place in one concatenation of the three pools
    flat, pool_base, base = [], [], 0
    for p in m.pools:
        o = np.zeros(p.n, np.int64)
        for k, q in enumerate(p.queries):
            o[k] = base
            base += len(q)
        pool_base.append(o)
        flat += [np.asarray(q, np.uint8) for q in p.queries]
    codes = np.concatenate(flat)
    q_base = np.array([pool_base[c][k] for c, k in zip(cls, idx)], np.int64)
    t_start = np.array([m.pools[c].t_start[k] for c, k in zip(cls, idx)], np.int64)
    _say(f"  {len(cls)} requests: {np.bincount(cls, minlength=3).tolist()} by class, "
         f"{len(big)} noisy windows >= 64 kb (longest {nq.max()})")
    return dict(m=m, ref=ref, pools=pools, cls=cls, idx=idx, arrivals=arrivals, sc=sc,
                codes=codes, q_base=q_base, t_start=t_start, big=big)


def _staged(mx, picks_by_class, arrivals_by_class, slots, label):
    from graphmambaformer_amd import serving as S

    classes = [S.make_class(f"{label}{c}", mx["ref"], mx["pools"][c], picks_by_class[c],
                            arrivals_by_class[c], mx["sc"], slots[c],
                            plane_budget_bytes=PLANE_BUDGET[c])
               for c in range(3)]
    res = S.run(classes, timeout_s=300.0)
    assert res.completed == sum(c.n for c in classes), label
    assert not res.timed_out and not res.waves_gave_up, label
    return classes


def _gather(mx, classes):
    """Staged outputs in request order, on the device."""
    cls = mx["cls"]
    n = len(cls)
    out = {k: torch.zeros(n, dtype=torch.long, device=DEV)
           for k in ("score", "qend", "tend", "rs", "ts", "n_cigar", "cigar_off")}
    out["counts"] = torch.zeros((n, 4), dtype=torch.long, device=DEV)
    cig, base = [], 0
    for c, k in enumerate(classes):
        sel = torch.from_numpy(np.flatnonzero(cls == c)).to(DEV)
        w = k.walk
        for key, t in (("score", k.score), ("qend", k.qend), ("tend", k.tend),
                       ("rs", w.read_start), ("ts", w.ref_start), ("n_cigar", w.n_cigar),
                       ("counts", w.counts)):
            out[key][sel] = t.long()
        out["cigar_off"][sel] = w.cigar_off.long() + base
        cig.append(w.cigar)
        base += w.cigar.numel()
    out["cigar"] = torch.cat(cig)
    return out


def _host(out):
    h = {k: v.cpu().numpy() for k, v in out.items()}
    h["cigar"] = h["cigar"].view(np.uint32) if h["cigar"].dtype == np.int32 else \
        h["cigar"].astype(np.uint32)
    return h


def _request_runs(h, r):
    o, nc = int(h["cigar_off"][r]), int(h["n_cigar"][r])
    return h["cigar"][o:o + nc]


def _same_request(a, b, r, label):
    for key in ("score", "qend", "tend", "rs", "ts", "n_cigar"):
        assert int(a[key][r]) == int(b[key][r]), (label, key, r)
    assert np.array_equal(a["counts"][r], b["counts"][r]), (label, "counts", r)
    assert np.array_equal(_request_runs(a, r), _request_runs(b, r)), (label, "cigar", r)


def _online(mx, caps, label):
    """Every request through an online session in arrival order (bursts cut
    short by full rings, polled as they go); outputs in request order."""
    from graphmambaformer_amd import serving as S

    m, cls, idx = mx["m"], mx["cls"], mx["idx"]
    n = len(cls)
    ocls = [S.OnlineClass(half_band=p.half_band, max_qlen=max(len(q) for q in p.queries),
                          n_slots=SLOTS[c], cigar_cap=caps[c],
                          plane_budget_bytes=PLANE_BUDGET[c])
            for c, p in enumerate(m.pools)]
    co = mx["q_base"]
    ql = np.array([len(m.pools[c].queries[k]) for c, k in zip(cls, idx)], np.int32)
    tl = np.array([m.pools[c].t_len[k] for c, k in zip(cls, idx)], np.int32)
    bo = np.array([m.pools[c].band_offset[k] for c, k in zip(cls, idx)], np.int32)
    by_ticket = [[] for _ in range(3)]
    h = {k: np.zeros(n, np.int64) for k in ("score", "qend", "tend", "rs", "ts", "n_cigar",
                                             "cigar_off")}
    h["counts"] = np.zeros((n, 4), np.int64)
    runs, base, seen = [], 0, 0

    def take(rec, cg):
        nonlocal base, seen
        for x in rec:
            r = by_ticket[int(x["cls"])][int(x["ticket"])]
            assert x["status"] == 0, (label, "status", r, int(x["status"]))
            h["score"][r], h["qend"][r], h["tend"][r] = x["score"], x["query_end"], x["target_end"]
            h["rs"][r], h["ts"][r], h["n_cigar"][r] = x["read_start"], x["ref_start"], x["n_cigar"]
            h["counts"][r] = x["counts"]
            h["cigar_off"][r] = base + int(x["cigar_off"])
        runs.append(cg)
        base += len(cg)
        seen += len(rec)

    end = time.monotonic() + 300.0
    with S.ServeSession(mx["ref"], ocls, mx["sc"], idle_timeout_s=5.0,
                        poll_results=8192) as sess:
        i = 0
        while i < n:
            assert time.monotonic() < end, (label, f"{i} of {n} online requests submitted")
            tk = sess.submit_many(cls[i:], mx["codes"], co[i:], ql[i:], mx["t_start"][i:],
                                  tl[i:], bo[i:])
            for j, t in enumerate(tk):
                c = int(cls[i + j])
                assert int(t) == len(by_ticket[c]), (label, "tickets in order")
                by_ticket[c].append(i + j)
            i += len(tk)
            take(*sess.poll_arrays())
        while seen < n:
            take(*sess.poll_arrays())
            assert time.monotonic() < end, (label, f"{seen} of {n} online results")
    h["cigar"] = np.concatenate(runs) if runs else np.zeros(0, np.uint32)
    return h


def test_served_mix_at_full_size(mix):
    from graphmambaformer_amd import serving as S

    mx = mix
    cls, idx, arr = mx["cls"], mx["idx"], mx["arrivals"]
    m = mx["m"]
    sel = [np.flatnonzero(cls == c) for c in range(3)]

    # staged, Poisson arrivals
    _say("case: staged serving of the full-size mix, paged planes")
    classes = _staged(mx, [idx[s] for s in sel], [arr[s] for s in sel], SLOTS, "mix")
    out = _gather(mx, classes)
    torch.cuda.synchronize()
    _say("  properties of every request")
    cols = check_alignments(
        "staged mix", q_codes=torch.from_numpy(mx["codes"]).to(DEV),
        q_base=torch.from_numpy(mx["q_base"]).to(DEV),
        t_codes=torch.from_numpy(np.ascontiguousarray(m.reference)).to(DEV),
        t_base=torch.from_numpy(mx["t_start"]).to(DEV), score=out["score"], qend=out["qend"],
        tend=out["tend"], rs=out["rs"], ts=out["ts"], n_cigar=out["n_cigar"],
        cigar_off=out["cigar_off"], cigar=out["cigar"], counts=out["counts"], scoring=mx["sc"])
    staged = _host(out)
    del out, classes
    torch.cuda.empty_cache()
    _say(f"  {len(cls)} requests, {cols} aligned columns checked against the bases")
    assert (staged["score"] > 0).all()

    # online, every request, against the staged outputs
    caps = [max(64, int(staged["n_cigar"][s].max())) for s in sel]
    _say("case: online serving of the full-size mix, paged planes, CIGAR caps", caps)
    online = _online(mx, caps, "online mix")
    for r in range(len(cls)):
        _same_request(staged, online, r, "online vs staged")

    # an oracle sample: 8 requests per class and every noisy window >= 64 kb
    rng = np.random.default_rng(24)
    sample = [rng.choice(s, size=8, replace=False) for s in sel]
    big_req = sel[2][np.isin(idx[sel[2]], mx["big"])]
    sample[2] = np.unique(np.concatenate([sample[2], big_req]))
    for c in range(3):
        p = m.pools[c]
        req = sample[c]
        _say(f"case: oracle sample, class {p.name}: {len(req)} requests, longest query "
             f"{max(len(p.queries[int(idx[r])]) for r in req)}")
        qs = [p.queries[int(idx[r])].astype(np.int16) + 1 for r in req]
        ts = [p.pair(int(idx[r]), m.reference)[1].astype(np.int16) + 1 for r in req]
        exp = O.sw_align_pairs(qs, ts, p.band_offset[idx[req]].astype(np.int64), p.half_band,
                               *mx["sc"], threads=8)
        for r, e in zip(req, exp):
            ctx = (p.name, int(r), int(idx[r]))
            assert int(staged["score"][r]) == e["score"], ctx
            assert int(staged["qend"][r]) == e["read_end"], ctx
            assert int(staged["tend"][r]) == e["ref_end"], ctx
            assert int(staged["rs"][r]) == e["read_start"], ctx
            assert int(staged["ts"][r]) == e["ref_start"], ctx
            assert _runs(_request_runs(staged, r)) == e["cigar"], ctx

    # the sample served alone equals its in-mix result
    for c in range(3):
        p = m.pools[c]
        _say(f"case: class {p.name}: {len(sample[c])} sampled requests each served alone")
        for r in sample[c]:
            one = S.make_class("alone", mx["ref"], mx["pools"][c], idx[[r]], np.zeros(1),
                               mx["sc"], 1, plane_budget_bytes=PLANE_BUDGET[c])
            res = S.run([one], timeout_s=60.0)
            assert res.completed == 1 and not res.timed_out, (p.name, int(r))
            w = one.walk
            got = {"score": one.score, "qend": one.qend, "tend": one.tend, "rs": w.read_start,
                   "ts": w.ref_start, "n_cigar": w.n_cigar}
            for key, t in got.items():
                assert int(t[0]) == int(staged[key][r]), (p.name, "alone", key, int(r))
            assert np.array_equal(w.counts[0].cpu().numpy(), staged["counts"][r])
            nc = int(w.n_cigar[0])
            mine = w.cigar[:nc].cpu().numpy().view(np.uint32)
            assert np.array_equal(mine, _request_runs(staged, r)), (p.name, "alone", int(r))
            del one
        oc = S.OnlineClass(half_band=p.half_band, max_qlen=max(len(q) for q in p.queries),
                           n_slots=4, cigar_cap=caps[c], plane_budget_bytes=PLANE_BUDGET[c])
        with S.ServeSession(mx["ref"], [oc], mx["sc"], idle_timeout_s=5.0) as sess:
            for r in sample[c]:
                k = int(idx[r])
                t = sess.submit(0, p.queries[k], int(p.t_start[k]), int(p.t_len[k]),
                                int(p.band_offset[k]))
                got = sess.wait(1, timeout_s=60.0)[0]
                ctx = (p.name, "online alone", int(r))
                assert got.ticket == t and not got.truncated and not got.no_pages, ctx
                assert got.score == int(staged["score"][r]), ctx
                assert got.query_end == int(staged["qend"][r]), ctx
                assert got.target_end == int(staged["tend"][r]), ctx
                assert got.read_start == int(staged["rs"][r]), ctx
                assert got.ref_start == int(staged["ts"][r]), ctx
                assert got.cigar == _runs(_request_runs(staged, r)), ctx
    torch.cuda.synchronize()

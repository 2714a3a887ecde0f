#!/usr/bin/env python3
"""Benchmark of the banded affine SW + traceback path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--skip-extras]

A step is one pass of the hot path over the whole noisy batch (8,192 pairs,
lengths lognormal median 8 kb clipped 0.5-100 kb, 6% errors with
homopolymer-weighted indels, band half-width 751, scoring 2/-4/-4/-2): one
launch of the forward kernel writes scores, ends and 4-bit traceback planes
and walks every pair into a CIGAR, longest pair first, on the tile each
length bin's autotune picks, with the library's default launch for a batch of
uneven lengths (kernels.default_launch: the pairs above one resident
wavefront's share of the cells split over several wavefronts each); the
whole-batch autotune of the alternatives (plain, longest pairs at raised
priority, other split sizes) is reported beside it.
Inputs are resident in HBM before the timed region.  ``value`` is GCUPS over
in-band cells; ``roofline`` is the int32 VALU roof of the forward kernel (the
binding one), ``roofline_hbm`` the HBM view.  Beside it the line carries the
per-pair-launch GPU form, the CPU oracle on every host core (bounded
stratified sample, one C call), the HIP-graph replay, the long-read and
short-read batches, serving latency, chaining, seeding and WFA.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

# The serving legs run one persistent kernel per band class side by side.
# HIP maps streams onto a per-process pool of hardware queues (4 by default)
# and a kernel waits behind a running one on a shared queue, so give the
# process more queues than streams before the HIP runtime starts.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

METRIC = "banded affine-SW GCUPS batched vs per-pair; p50/p99 latency at fixed QPS"
HBM_PEAK_GBS = 8000.0
INT32_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12  # 39.3 T int32 VALU lane-ops/s
OPS_PER_CELL_TB = 20  # int32 ops per cell with traceback (score/end 14 + traceback 6)
CHAIN_OPS_PER_EVAL = 20  # int32 ops per chaining pair evaluation
CHAIN_BYTES_PER_ANCHOR = 24  # ref_end 8 + read_end 4 + weight 4 read, f 4 + parent 4 written
TB_BYTES_PER_CELL = 0.5


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(name: str, cells: int, tile=None, launch=None):
    """HBM bytes per forward launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs (profiles/sw_forward_<name>_traffic.json, calibrated
    against a plain 1 GiB copy), if they were taken on this exact batch, tile
    and launch; else None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"sw_forward_{name}_traffic.json")
    try:
        with open(path) as fh:
            rec = json.load(fh)
    except (OSError, ValueError):
        return None
    if rec.get("cells") != cells:
        return None
    if tile is not None and rec.get("tile") != [tile.cells_per_lane, tile.lanes_per_pair]:
        return None
    if launch is not None and rec.get("launch", launch) != launch:
        return None
    return rec.get("traffic_bytes_per_launch")


def lpt_order(cells: np.ndarray):
    return np.argsort(-cells, kind="stable").astype(np.int32)


def time_events(torch, fn, reps, stat="mean"):
    """Mean (or median: stat="median") ms of fn() over reps, HIP events on
    torch's current stream (the stream the library launches on)."""
    s = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    for a, b in ev:
        a.record(s)
        fn()
        b.record(s)
    torch.cuda.synchronize()
    t = [a.elapsed_time(b) for a, b in ev]
    return float(np.median(t) if stat == "median" else np.mean(t))


def scoring(w):
    return (w.match, w.mismatch, w.gap_open, w.gap_extend)


def tile_timings(t: dict) -> dict:
    """autotune_tile's timings as JSON keys: "CxG" on the tuning input,
    "full:CxG" re-timed on the whole batch."""
    return {(f"full:{k[1]}x{k[2]}" if k[0] == "full" else f"{k[0]}x{k[1]}"): v
            for k, v in t.items()}


class BatchRun:
    """One workload resident on the device: packed pairs, longest-first
    order, tile and every buffer preallocated; ``step()`` is one forward
    launch (+ the walk, fused or as a second launch)."""

    def __init__(self, torch, K, w, dev, fused=True, tile=None, launch=None):
        self.w, self.K, self.fused = w, K, fused
        self.launch = dict(launch or {})
        self.cells = w.cells()
        self.total_cells = int(self.cells.sum())
        self.inp = K.upload_pairs(w.queries, w.targets, w.band_offset, dev, alphabet="2bit")
        self.order = torch.from_numpy(lpt_order(self.cells)).to(dev)
        self.tile = tile or K.resolve_tile(w.half_band)
        self.ws = K.make_workspace(self.inp, self.tile, cigar=True)
        torch.cuda.synchronize()

    def forward(self):
        K, w = self.K, self.w
        return K.sw_forward(self.inp, *scoring(w), w.half_band, traceback=True, order=self.order,
                            tile=self.tile, workspace=self.ws, cigar=self.fused,
                            **(self.launch if self.fused else {}))

    def walk(self, fwd):
        return fwd.walk if self.fused else self.K.sw_traceback(self.inp, fwd, workspace=self.ws)

    def step(self):
        fwd = self.forward()
        return fwd, self.walk(fwd)


def per_pair_gcups(torch, K, w, idx, dev):
    """The plain per-pair form: one forward launch + one walk launch per pair."""
    from graphmambaformer_amd import workloads as WL

    sub = WL.subsample(w, idx)
    tile = K.batch_tile(w.half_band, 1, dev)  # the tile a one-pair launch gets
    ws1 = {"counter": torch.zeros(1, dtype=torch.int32, device=dev)}
    inps = [K.upload_pairs([sub.queries[k]], [sub.targets[k]], sub.band_offset[k:k + 1], dev,
                           alphabet="2bit") for k in range(sub.n_pairs)]

    def run():
        for one in inps:
            f1 = K.sw_forward(one, *scoring(w), w.half_band, traceback=True, tile=tile,
                              workspace=ws1)
            K.sw_traceback(one, f1)

    run()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    run()
    torch.cuda.synchronize()
    s = time.perf_counter() - t1
    return int(sub.cells().sum()) / s / 1e9


def cpu_gcups(w, idx, threads):
    """The C oracle (forward + traceback of independent pairs, one C call:
    POSIX threads take pairs from a shared counter, as the reference's
    parallel_map fans reads over host cores)."""
    from oracle import oracle as O

    if not os.path.exists(O._LIB_PATH):
        O.build()
    qs = [w.as_ref_codes(int(k))[0] for k in idx]
    ts = [w.as_ref_codes(int(k))[1] for k in idx]
    t1 = time.perf_counter()
    O.sw_align_batch(qs, ts, w.band_offset[idx], w.half_band, *scoring(w), threads=threads)
    s = time.perf_counter() - t1
    return float(w.cells()[idx].sum()) / s / 1e9, s


SERVE_SLOTS = (512, 1024, 2048)  # persistent wavefronts of the short / long / noisy classes
# traceback-plane budgets of the serving classes (paged planes: a request is
# admitted when its planes' pages are free); per-slot slabs for each class's
# longest request would take 0.02 / 6.6 / 156 GB
SERVE_PLANE_BUDGET = (256 << 20, 8 << 30, 32 << 30)
# staged serving, --serve-prio-rows R: requests of the noisy class with at
# least R query rows run at raised wave priority.  Off by default: priority
# moves the tail between classes rather than shortening it.  R = 32,768 (the
# longest ~8% of the noisy lengths), two runs each: noisy p99 146-148 -> 120
# ms and capacity 301-311 k -> 338 k req/s, but long-class p99 18 -> 39 ms,
# short-class p99 0.17 -> 0.34 ms, overall p99 51.9 -> 52.5 ms; online form
# (not applied there): noisy p99 161 -> 173 ms.
SERVE_PRIO_ROWS = 0


# staged serving, --serve-split R,K,G,N: noisy-class requests with at least R
# query rows go to a class of their own whose G workgroups run each request as
# K band slices (K wavefronts); the noisy class keeps the shorter ones on N
# wavefronts (R = 0: no split class).  Every class's persistent kernel must be
# resident at once: per SIMD, the long class's wavefront (~104 VGPRs), half a
# short one and the noisy ones (~224 each) leave room for split wavefronts
# (~152 each) only if the noisy class gives some up; the split class launches
# first so its workgroups are placed before the others fill the CUs.
# Measured at 65% load, two runs: (32768, 3, 256, 1024) against none: noisy
# p99 137.5 -> 88.5-90.7 ms, overall p99 48.8 -> 45.1-45.6 ms, long-class p99
# 16.8 -> 22.2 ms, saturation 314 k -> 304-305 k req/s; (16384, 3, 256, 1024):
# saturation 212 k; (32768, 3, 128, 1280): 198 k, noisy p99 114 ms.
SERVE_SPLIT = (32768, 3, 256, 1024)


def serving_extra(torch, WL, dev, n_req, load, online=True, prio_rows=SERVE_PRIO_ROWS,
                  split=SERVE_SPLIT):
    """p50 / p99 latency of a Poisson request stream at a fixed rate.

    A resident 46 Mbp reference; requests 60% short (150 bp, half_band 32), 30%
    long (lognormal 12 kb, half_band 256), 10% noisy (lognormal 8 kb clipped
    0.5-100 kb, 6% errors, half_band 751), drawn from pools of distinct
    windows.  First a saturation run (every request at t = 0) gives the
    capacity; then the paced run offers `load` x capacity requests/s.
    Latency = completion seen by the host - arrival; the first 10% of arrivals
    are warm-up."""
    from graphmambaformer_amd import serving as S

    t0 = time.time()
    mix = WL.serving_mix()
    ref = S.ResidentReference.upload(mix.reference, dev)
    pools = [S.RequestPool.upload(p.queries, p.t_start, p.t_len, p.band_offset, p.half_band, dev)
             for p in mix.pools]
    cls, idx = WL.serving_requests(mix, n_req)
    pool_cells = [p.cells() for p in mix.pools]
    req_cells = np.array([pool_cells[c][k] for c, k in zip(cls, idx)], np.int64)
    total_cells = int(req_cells.sum())
    sc = (mix.match, mix.mismatch, mix.gap_open, mix.gap_extend)
    log(f"[bench] serving stream: {n_req} requests, {total_cells:.3e} cells, "
        f"{time.time() - t0:.1f}s")

    # the staged run's classes: the three bands, and the long noisy requests
    # in a split class when asked
    split_rows, split_k, split_groups, split_noisy = (tuple(split) + (0,))[:4]
    qlen_req = np.array([len(mix.pools[c].queries[k]) for c, k in zip(cls, idx)], np.int64)
    scls = cls.copy()
    noisy = [p.name for p in mix.pools].index("noisy")
    if split_rows > 0:
        scls[(cls == noisy) & (qlen_req >= split_rows)] = len(mix.pools)
    slots = list(SERVE_SLOTS)
    if split_rows > 0:
        slots[noisy] = split_noisy or max(1, slots[noisy] - split_groups * split_k)
    # plane budget of the split class: the part of the noisy class's budget
    # its requests' planes take
    split_frac = (float(req_cells[scls == len(mix.pools)].sum()) /
                  max(float(req_cells[cls == noisy].sum()), 1.0)) if split_rows > 0 else 0.0
    budgets = list(SERVE_PLANE_BUDGET)

    def build(arrivals):
        """The classes in launch order, each with its class id (the index into
        scls): a split class first, so its multi-wavefront workgroups are
        placed before the one-wavefront classes take the remaining room."""
        out = []
        if split_rows > 0:
            sel = np.flatnonzero(scls == len(mix.pools))
            out.append((len(mix.pools), S.make_class(
                "noisy_split", ref, pools[noisy], idx[sel], arrivals[sel], sc, split_groups,
                split_slices=split_k,
                plane_budget_bytes=int(budgets[noisy] * split_frac) + (4 << 30))))
        for c, p in enumerate(mix.pools):
            sel = np.flatnonzero(scls == c)
            budget = budgets[c]
            if c == noisy and split_rows > 0:
                budget = int(budget * (1.0 - split_frac)) + (4 << 30)
            out.append((c, S.make_class(p.name, ref, pools[c], idx[sel], arrivals[sel], sc,
                                        slots[c], plane_budget_bytes=budget,
                                        prio_rows=prio_rows if p.name == "noisy" else 0)))
        return out

    def gather(classes):
        comp = np.empty(n_req)
        for c, k in classes:
            comp[np.flatnonzero(scls == c)] = k.complete_s
        return comp

    sat = build(np.zeros(n_req))
    slab_gb = lambda k: round(k.n_slots * k.slab_words * 4 / 2**30, 2)  # noqa: E731
    pages = {k.name: ({"page_blocks": k.pages.page_blocks, "pages": k.pages.n_pages,
                       "pool_gb": round(k.pages.pool_bytes / 2**30, 2),
                       "slab_equivalent_gb": slab_gb(k)} if k.pages else {"slab_gb": slab_gb(k)})
             for _, k in sat}
    r_sat = S.run([k for _, k in sat], timeout_s=300.0)
    capacity = n_req / r_sat.wall_s
    del sat
    torch.cuda.empty_cache()
    log(f"[bench] serving saturation: {r_sat.wall_s * 1e3:.1f} ms -> {capacity:.0f} req/s, "
        f"{total_cells / r_sat.wall_s / 1e9:.1f} GCUPS")

    rate = load * capacity
    arr = WL.poisson_arrivals(n_req, rate)
    paced = build(arr)
    r = S.run([k for _, k in paced], timeout_s=300.0)
    comp = gather(paced)
    skip_t = arr[n_req // 10]
    keep = arr >= skip_t
    lat_all = S.latency_summary(arr[keep], comp[keep])
    per_class = {}
    for c, p in enumerate(mix.pools):
        m = keep & (cls == c)
        per_class[p.name] = S.latency_summary(arr[m], comp[m])
    if split_rows > 0:
        for label, m in (("noisy_split", scls == len(mix.pools)),
                         ("noisy_unsplit", (cls == noisy) & (scls == noisy))):
            per_class[label] = S.latency_summary(arr[keep & m], comp[keep & m])
    del paced
    res = {
        "qps": round(rate, 1),
        "offered_load": load,
        "capacity_qps": round(capacity, 1),
        "saturation_gcups": round(total_cells / r_sat.wall_s / 1e9, 2),
        "paced_wall_s": round(r.wall_s, 4),
        "paced_gcups": round(total_cells / r.wall_s / 1e9, 2),
        "latency": lat_all,
        "latency_by_class": per_class,
        "completed": r.completed,
        "timed_out": r.timed_out,
        "slots": dict(zip([p.name for p in mix.pools], slots)),
        "noisy_prio_rows": prio_rows,  # staged form only
        "noisy_split": ({"rows": split_rows, "slices": split_k, "workgroups": split_groups,
                         "noisy_waves": slots[noisy],
                         "requests": int((scls == len(mix.pools)).sum()),
                         "cell_frac_of_noisy": round(split_frac, 4)} if split_rows > 0 else None),
        "planes": pages,
        "workload": f"{n_req} Poisson requests, 60% short (150 bp, half_band 32) / 30% long "
                    "(lognormal 12 kb, half_band 256) / 10% noisy (lognormal 8 kb, 6% errors, "
                    "half_band 751) against a resident 46 Mbp 2-bit reference; pools of "
                    "2048/1024/512 distinct windows, seed 4; first 10% of arrivals are warm-up; "
                    "queries pre-staged in HBM; traceback planes paged from per-class budgets "
                    "(admission by footprint)",
    }
    log(f"[bench] serving paced at {rate:.0f} req/s: {res['latency']} by class "
        f"{per_class}")
    del pools
    torch.cuda.empty_cache()
    if online:
        res["online"] = online_serving_extra(torch, S, mix, ref, cls, idx, total_cells, load, sc,
                                             prio_rows=0)
    return res


ONLINE_CIGAR_CAP = (512, 8192, 16384)  # CIGAR runs kept per request, short / long / noisy


def online_serving_extra(torch, S, mix, ref, cls, idx, total_cells, load, sc, prio_rows):
    """The same request stream through the online session API
    (gmfa_serve_submit_many / gmfa_serve_poll): queries packed by the host
    into the class rings as they arrive, results polled back to host memory.
    Saturation (everything offered at t = 0, ring-limited) gives the online
    capacity; the paced run offers `load` x that capacity, Poisson.  Latency =
    when the polling host thread has the result (score, ends, CIGAR in host
    memory) - arrival."""
    from graphmambaformer_amd import workloads as WL

    n_req = len(cls)
    flat, offs = [], []
    base = 0
    for p in mix.pools:
        o = np.zeros(p.n, np.int64)
        for k, q in enumerate(p.queries):
            o[k] = base
            base += len(q)
        offs.append(o)
        flat += [np.asarray(q, np.uint8) for q in p.queries]
    codes = np.concatenate(flat)
    co = np.array([offs[c][k] for c, k in zip(cls, idx)], np.int64)
    ql = np.array([len(mix.pools[c].queries[k]) for c, k in zip(cls, idx)], np.int32)
    to = np.array([mix.pools[c].t_start[k] for c, k in zip(cls, idx)], np.int64)
    tl = np.array([mix.pools[c].t_len[k] for c, k in zip(cls, idx)], np.int32)
    bo = np.array([mix.pools[c].band_offset[k] for c, k in zip(cls, idx)], np.int32)
    ocls = [S.OnlineClass(half_band=p.half_band, max_qlen=max(len(q) for q in p.queries),
                          n_slots=SERVE_SLOTS[c], cigar_cap=ONLINE_CIGAR_CAP[c],
                          plane_budget_bytes=SERVE_PLANE_BUDGET[c],
                          prio_rows=prio_rows if p.name == "noisy" else 0)
            for c, p in enumerate(mix.pools)]

    def drive(arrivals):
        comp = np.full(n_req, -1.0)
        kern = np.zeros(n_req, np.int64)
        sess = S.ServeSession(ref, ocls, sc, idle_timeout_s=10.0, poll_results=8192)
        base_t = [0] * len(ocls)
        by_ticket = [np.full(int((cls == c).sum()), -1, np.int64) for c in range(len(ocls))]
        truncated = 0
        host = {"iterations": 0, "submit_s": 0.0, "poll_s": 0.0}
        try:
            i = done = 0
            t0 = time.perf_counter()
            last_done, last_t, last_log = 0, 0.0, 0.0
            while done < n_req:
                now = time.perf_counter() - t0
                if done != last_done:
                    last_done, last_t = done, now
                elif now - last_t > 60.0:
                    raise RuntimeError(f"online serving made no progress for 60 s "
                                       f"({done} of {n_req} done, {i} submitted)")
                if now - last_log > 5.0:
                    last_log = now
                    log(f"[bench] online: {done} of {n_req} done, {i} submitted, {now:.1f} s")
                j = int(np.searchsorted(arrivals, now, side="right"))
                host["iterations"] += 1
                if j > i:
                    ts = time.perf_counter()
                    tk = sess.submit_many(cls[i:j], codes, co[i:j], ql[i:j], to[i:j], tl[i:j],
                                          bo[i:j])
                    m = len(tk)
                    if m:
                        cl = cls[i:i + m]
                        for c in range(len(ocls)):
                            sel = cl == c
                            by_ticket[c][tk[sel] - base_t[c]] = np.arange(i, i + m)[sel]
                        i += m
                    host["submit_s"] += time.perf_counter() - ts
                ts = time.perf_counter()
                rec, _ = sess.poll_arrays()
                host["poll_s"] += time.perf_counter() - ts
                if len(rec):
                    t = time.perf_counter() - t0
                    for c in range(len(ocls)):
                        r = rec[rec["cls"] == c]
                        q = by_ticket[c][r["ticket"] - base_t[c]]
                        comp[q] = t
                        kern[q] = r["done_tick"] - r["claim_tick"]
                    truncated += int((rec["status"] != 0).sum())
                    done += len(rec)
            wall = time.perf_counter() - t0
            st = sess.stats()
        finally:
            sess.close()
        st["host_loop"] = {k: round(v, 4) for k, v in host.items()}
        return comp, wall, truncated, st, kern / st["ticks_per_s"]

    zero = np.zeros(n_req)
    comp, wall_sat, trunc_sat, st_sat, _ = drive(zero)
    capacity = n_req / wall_sat
    log(f"[bench] online saturation: {wall_sat * 1e3:.1f} ms -> {capacity:.0f} req/s, "
        f"{total_cells / wall_sat / 1e9:.1f} GCUPS, truncated {trunc_sat}")
    rate = load * capacity
    arr = WL.poisson_arrivals(n_req, rate)
    comp, wall, trunc, st, kern_s = drive(arr)
    keep = arr >= arr[n_req // 10]
    per_class = {p.name: S.latency_summary(arr[keep & (cls == c)], comp[keep & (cls == c)])
                 for c, p in enumerate(mix.pools)}
    # claim -> outputs written, on the device clock: the kernel's share of the latency
    kernel_by_class = {p.name: S.latency_summary(np.zeros(int((keep & (cls == c)).sum())),
                                                 kern_s[keep & (cls == c)])
                       for c, p in enumerate(mix.pools)}
    res = {
        "qps": round(rate, 1),
        "offered_load": load,
        "capacity_qps": round(capacity, 1),
        "saturation_gcups": round(total_cells / wall_sat / 1e9, 2),
        "paced_wall_s": round(wall, 4),
        "latency": S.latency_summary(arr[keep], comp[keep]),
        "latency_by_class": per_class,
        "kernel_time_by_class": kernel_by_class,
        "truncated_cigars": trunc + trunc_sat,
        "launches": int(st["launches"]),
        "host_loop": {"saturation": st_sat["host_loop"], "paced": st["host_loop"]},
        "cigar_cap": dict(zip([p.name for p in mix.pools], ONLINE_CIGAR_CAP)),
        "note": "one Python thread submits bursts of arrived requests (host 2-bit packing into "
                "the pinned rings) and polls results into host memory; the reference stays "
                "resident in HBM",
    }
    log(f"[bench] online paced at {rate:.0f} req/s: {res['latency']} by class {per_class}; "
        f"kernel claim->done {kernel_by_class}")
    return res


def seeding_extra(torch, WL, dev):
    """Stage-1 k-mer seeding of 4,096 10 kb reads against a 46 Mbp minimizer
    index (k 15, w 10): device select + lookup + expand vs the oracle's
    per-read query on one host core."""
    from oracle import seeding as OS

    from graphmambaformer_amd.alignment import GPUKmerIndex

    t0 = time.time()
    ref, reads = WL.seeding_reads()
    idx = GPUKmerIndex(ref, k=15, window=10, device=dev)
    log(f"[bench] seeding index: {int(idx.keys.numel())} keys, {time.time() - t0:.1f}s")
    t1 = time.perf_counter()
    prep = idx.prepare(reads)
    torch.cuda.synchronize()
    prep_s = time.perf_counter() - t1
    idx.seed(prep)
    torch.cuda.synchronize()
    reps = 5
    t1 = time.perf_counter()
    for _ in range(reps):
        rp, fp, offs = idx.seed(prep)
    torch.cuda.synchronize()
    gpu_s = (time.perf_counter() - t1) / reps
    n_anchors = int(offs[-1])
    sample = reads[::32]
    tab = OS.KmerTable(idx._table.keys, idx._table.offsets, idx._table.positions, 15)
    t1 = time.perf_counter()
    n_cpu = 0
    for r in sample:
        n_cpu += len(OS.kmer_query(tab, r, 10)[0])
    cpu_s = time.perf_counter() - t1
    res = {
        "reads_per_s": round(len(reads) / gpu_s, 1),
        "anchors_per_s": round(n_anchors / gpu_s, 1),
        "ms": round(gpu_s * 1e3, 3),
        "host_pack_upload_ms": round(prep_s * 1e3, 1),
        "anchors": n_anchors,
        "cpu_reads_per_s": round(len(sample) / cpu_s, 1),
        "cpu_threads": 1,
        "workload": "4096 reads x 10 kb (1% errors) against a 46 Mbp random reference, "
                    "minimizer index k 15 w 10, seed 7; reads resident on the device, GPU time = "
                    "select + lookup + scans + anchor-count readback + expand; CPU: the "
                    "oracle's per-read query on every 32nd read",
    }
    log(f"[bench] seeding: {res}")
    return res


def wfa_extra(torch, K, WL, dev):
    """Batched wavefront alignment (distance + CIGAR) of 4,096 10 kb pairs
    with 1% errors vs the oracle's WFA on a sample."""
    from oracle import wfa as OW

    rng = np.random.default_rng(8)
    qs, ts = [], []
    for _ in range(4096):
        t = rng.integers(0, 4, size=10_000).astype(np.uint8)
        qs.append(WL.mutate(rng, t, 0.01, 0.5, 3).astype(np.int8) + 1)
        ts.append(t.astype(np.int8) + 1)
    inp = K.upload_pairs(qs, ts, np.zeros(len(qs), np.int32), dev, alphabet="ref")
    K.wfa_align_device(inp, 4096)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        r = K.wfa_align_device(inp, 4096)
    torch.cuda.synchronize()
    gpu_s = (time.perf_counter() - t1) / reps
    dist = r.distance
    t1 = time.perf_counter()
    K.wfa_align(inp, 4096)
    host_s = time.perf_counter() - t1
    sub = range(0, len(qs), 256)
    t1 = time.perf_counter()
    for p in sub:
        OW.wfa_align(qs[p], ts[p], 4096)
    cpu_s = (time.perf_counter() - t1) / len(sub)
    res = {
        "pairs_per_s": round(len(qs) / gpu_s, 1),
        "ms": round(gpu_s * 1e3, 3),
        "with_host_cigars_ms": round(host_s * 1e3, 3),
        "mean_distance": round(float(np.mean(dist)), 2),
        "cpu_pairs_per_s": round(1.0 / cpu_s, 2),
        "cpu_threads": 1,
        "workload": "4096 pairs x 10 kb, 1% errors (half 1-3 bp indels), max_distance 4096, "
                    "seed 8; pairs resident on the device; GPU: distance pass + distance readback + "
                    "history/traceback pass, CIGARs left on the device (with_host_cigars_ms adds "
                    "the CIGAR readback and decoding into Python lists); CPU: the oracle's (pure "
                    "Python) WFA on every 256th pair",
    }
    log(f"[bench] wfa: {res}")
    return res


def long_read_extra(torch, K, WL, dev, args, threads):
    """The long-read batch (16,384 pairs, 5-25 kb, band 513): batched GCUPS on
    the autotuned tile, its HIP-graph replay, the x-drop extension path, the
    rate with host buffers in and out, the per-pair form and the CPU oracle."""
    t0 = time.time()
    w = WL.sw_long()
    log(f"[bench] sw_long generated: {w.n_pairs} pairs, {w.cells().sum():.3e} cells, "
        f"{time.time() - t0:.1f}s")
    # tile autotuning on every 4th pair (by cell count), the two fastest tiles
    # and the default then timed again on the whole batch, where the choice is
    # made (the sample can rank near-equal tiles the other way round)
    sub_t = WL.subsample(w, np.argsort(w.cells(), kind="stable")[::4])
    inp_t = K.upload_pairs(sub_t.queries, sub_t.targets, sub_t.band_offset, dev, alphabet="2bit")
    inp_f = K.upload_pairs(w.queries, w.targets, w.band_offset, dev, alphabet="2bit")
    tuned, tune_ms = K.autotune_tile(inp_t, w.half_band, scoring(w),
                                     order=torch.from_numpy(lpt_order(sub_t.cells())).to(dev),
                                     confirm=(inp_f, torch.from_numpy(lpt_order(w.cells())).to(dev)))
    del inp_t, inp_f
    torch.cuda.empty_cache()
    run = BatchRun(torch, K, w, dev, True, tile=tuned)
    total_cells, cells, tile = run.total_cells, run.cells, run.tile
    run.step()
    ms = time_events(torch, run.step, 5)
    gcups = total_cells / (ms / 1e3) / 1e9
    gr = K.SwGraph(run.inp, scoring(w), w.half_band, order=run.order, tile=tile,
                   workspace=run.ws, cigar=True)
    gms = time_events(torch, gr.replay, 5)
    del gr
    valu_tops = total_cells * OPS_PER_CELL_TB / (ms / 1e3) / 1e12
    out = {
        "batched_gcups": round(gcups, 2),
        "batched_ms": round(ms, 3),
        "graph_gcups": round(total_cells / (gms / 1e3) / 1e9, 2),
        "graph_ms": round(gms, 3),
        "valu_frac": round(valu_tops / INT32_PEAK_TOPS, 4),
        "traffic": measured_traffic("sw_long", total_cells, tile),
        "tile": [tile.cells_per_lane, tile.lanes_per_pair],
        "autotune_ms": tile_timings(tune_ms),
        "workload": "16384 pairs, lognormal(12 kb, 0.4) clipped 5-25 kb, 0.5% errors, "
                    "half_band 256, seed 2, fused walk",
    }
    log(f"[bench] sw_long: {ms:.1f} ms -> {gcups:.1f} GCUPS, graph {gms:.1f} ms")

    # host buffers in and out: upload + device packing, the step, the CIGAR
    # runs back to the host
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    inp_h = K.upload_pairs(w.queries, w.targets, w.band_offset, dev, alphabet="2bit")
    fwd_h = K.sw_forward(inp_h, *scoring(w), w.half_band, traceback=True, order=run.order,
                         tile=tile, workspace=run.ws, cigar=True)
    tb_h = fwd_h.walk
    nc_h = tb_h.n_cigar.to(torch.int64)
    tot = int(nc_h.sum().item())
    owner = torch.repeat_interleave(torch.arange(w.n_pairs, device=dev), nc_h)
    within = torch.arange(tot, device=dev) - (torch.cumsum(nc_h, 0) - nc_h)[owner]
    runs_h = tb_h.cigar[tb_h.cigar_off[owner] + within].cpu()
    _ = (fwd_h.score.cpu(), fwd_h.query_end.cpu(), fwd_h.target_end.cpu(), nc_h.cpu())
    host_s = time.perf_counter() - t1
    out["host_inclusive"] = {
        "gcups": round(total_cells / host_s / 1e9, 2),
        "s": round(host_s, 3),
        "cigar_runs": int(runs_h.numel()),
        "note": "host arrays in: concatenation, one upload per stream, 2-bit packing on the "
                "device, forward with fused walk, scores/ends/CIGAR runs copied back",
    }
    del inp_h, fwd_h, tb_h, runs_h

    # the x-drop extension path (drop-in form): forward with planes and per-row
    # keys, the batch-wide x-drop of each group of 8 pairs, the walk from the
    # x-drop ends
    gofs = np.arange(0, w.n_pairs + 1, 8)
    xd = {}

    def xdrop_step():
        f = K.sw_forward(run.inp, *scoring(w), w.half_band, traceback=True, rowkeys=True,
                         order=run.order, tile=tile, workspace=run.ws)
        sc_, qe_, te_ = K.sw_xdrop(f, run.inp, gofs, 600.0, w.half_band)
        xd["tb"] = K.sw_traceback(run.inp, f, sc_, qe_, te_, workspace=run.ws)
        return f

    xdrop_step()
    xms = time_events(torch, xdrop_step, 3)
    out["xdrop"] = {
        "gcups": round(total_cells / (xms * 1e-3) / 1e9, 2),
        "ms": round(xms, 3),
        "workload": "forward with traceback planes + per-row keys (longest first), batch "
                    "x-drop 600 over groups of 8 pairs, standalone walk from the x-drop ends",
    }
    del xd, run
    torch.cuda.empty_cache()
    sub = np.arange(0, w.n_pairs, 256)[:64]
    out["per_pair_gcups"] = round(per_pair_gcups(torch, K, w, sub, dev), 4)
    srt = np.argsort(cells, kind="stable")
    subc = srt[::args.cpu_sample_stride]
    cg, cs = cpu_gcups(w, subc, threads)
    out["cpu_gcups"] = round(cg, 4)
    out["cpu_threads"] = threads
    out["cpu_sample"] = (f"{len(subc)} pairs (every {args.cpu_sample_stride}th by cell count, "
                         f"{int(cells[subc].sum()):.3e} cells), one C call, {cs:.1f} s")
    log(f"[bench] sw_long extras: {out}")
    return out


def single_batch(torch, K, WL, dev, args):
    """--workload long / short: W untimed + K timed steps (forward with the
    fused walk) of that batch alone, on the --tile tile or the library's
    default for the batch (profiling runs of that kernel)."""
    w = WL.sw_long() if args.workload == "long" else WL.sw_short()
    if args.tile:
        tile = K.resolve_tile(w.half_band, *(int(x) for x in args.tile.split(",")))
    else:
        tile = K.batch_tile(w.half_band, w.n_pairs, dev)
    run = BatchRun(torch, K, w, dev, True, tile=tile)
    for _ in range(args.warmup):
        run.step()
    torch.cuda.synchronize()
    ms = time_events(torch, run.step, max(args.steps, 1))
    gcups = run.total_cells / (ms / 1e3) / 1e9
    log(f"[bench] {args.workload}: {ms:.4f} ms -> {gcups:.1f} GCUPS on "
        f"{tile.cells_per_lane}x{tile.lanes_per_pair}")
    return {"workload": args.workload, "pairs": w.n_pairs, "cells": run.total_cells,
            "tile": [tile.cells_per_lane, tile.lanes_per_pair], "steps": args.steps,
            "ms_per_step": round(ms, 4), "gcups": round(gcups, 2),
            "valu_frac": round(run.total_cells * OPS_PER_CELL_TB / (ms / 1e3) / 1e12 /
                               INT32_PEAK_TOPS, 4)}


def chains_extra(torch, K, WL, dev, threads):
    """The chaining configuration (4,096 reads, Poisson(400) anchors): the DP
    in the integer and float64 profiles with their VALU / HBM fractions, the
    oracle's DP on every host core, and anchors -> chains on the device (sort,
    DP, backtrack, select) against the oracle's per-read chain_read."""
    from oracle import oracle as O

    wc = WL.chains()
    re_ = torch.from_numpy(wc.read_end).to(dev)
    fe_ = torch.from_numpy(wc.ref_end).to(dev)
    wt = torch.from_numpy(wc.weight).to(dev)
    so = torch.from_numpy(wc.seg_off).to(dev)

    def chain_step():
        K.chain_dp(re_, fe_, wt, so, wc.max_lookback, wc.max_gap, gap_open=wc.gap_open,
                   gap_extend=wc.gap_extend)

    chain_step()
    cms = time_events(torch, chain_step, 5)
    from graphmambaformer_amd.alignment.chaining import penalty_table
    from graphmambaformer_amd.config import ChainingConfig

    lut = penalty_table(ChainingConfig(max_lookback=wc.max_lookback, max_gap=wc.max_gap), dev)
    wt64 = wt.to(torch.float64)

    def chain_step64():
        K.chain_dp(re_, fe_, wt64, so, wc.max_lookback, wc.max_gap, pen_lut=lut)

    chain_step64()
    cms64 = time_events(torch, chain_step64, 5)
    # CPU: the oracle's C DP per read, reads fanned over every host core
    # (chaining.py:657-666 runs reads in a thread pool; the C call releases
    # the interpreter lock)
    from concurrent.futures import ThreadPoolExecutor

    def cpu_read(s_):
        a, b = wc.seg_off[s_], wc.seg_off[s_ + 1]
        O.chain_dp_i32(wc.read_end[a:b], wc.ref_end[a:b], wc.weight[a:b], wc.max_lookback,
                       wc.max_gap, wc.gap_open, wc.gap_extend)

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(cpu_read, range(64)))
        t1 = time.perf_counter()
        list(ex.map(cpu_read, range(wc.n_reads)))
        ccpu = time.perf_counter() - t1
    # anchors -> chains on the device: sort into DP order, DP (float64 reference
    # defaults), backtrack, select; host arrays in, chain records out
    ccfg = ChainingConfig(max_lookback=wc.max_lookback, max_gap=wc.max_gap)
    nrd = wc.n_reads
    seg_of = np.repeat(np.arange(nrd), np.diff(wc.seg_off))
    shuf = np.lexsort((np.random.default_rng(0).random(len(seg_of)), seg_of))  # unsorted input
    rp_ = wc.read_end[shuf] - 15
    fp_ = wc.ref_end[shuf] - 15
    ids = (np.arange(len(seg_of)) - wc.seg_off[seg_of])[shuf]
    read_seg = np.stack([np.arange(nrd), np.full(nrd, -1)], axis=1).astype(np.int32)
    lut64 = penalty_table(ccfg, dev)

    def pipeline():
        return K.chain_pipeline(rp_ + 15, fp_ + 15, rp_, fp_, np.full(len(rp_), 15.0), ids,
                                wc.seg_off, np.ones(nrd, np.int32), read_seg, wc.seg_off[:-1],
                                ccfg.max_lookback, ccfg.max_gap, lut64, ccfg.min_chain_anchors,
                                ccfg.min_chain_score, ccfg.secondary_score_ratio,
                                ccfg.secondary_overlap, ccfg.max_chains, dev)

    pipeline()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(3):
        pres = pipeline()
    pms_host = (time.perf_counter() - t1) / 3 * 1e3
    up_ = K.chain_upload(rp_ + 15, fp_ + 15, rp_, fp_, np.full(len(rp_), 15.0), ids, wc.seg_off,
                         read_seg, wc.seg_off[:-1], dev)

    def device_run():
        return K.chain_run(up_, ccfg.max_lookback, ccfg.max_gap, lut64, ccfg.min_chain_anchors,
                           ccfg.min_chain_score, ccfg.secondary_score_ratio,
                           ccfg.secondary_overlap, ccfg.max_chains)

    device_run()
    pms = time_events(torch, device_run, 5)
    sub = np.arange(0, nrd, 16)
    t1 = time.perf_counter()
    for r in sub:
        m = seg_of[shuf] == r
        O.chain_read(rp_[m], fp_[m], np.full(int(m.sum()), 15), np.ones(int(m.sum()), np.int64),
                     max_lookback=ccfg.max_lookback, max_gap=ccfg.max_gap)
    cpu_pipe = (time.perf_counter() - t1) / len(sub)
    evals = wc.pair_evaluations()
    n_anchor = int(wc.seg_off[-1])
    res = {
        "reads_per_s": round(wc.n_reads / (cms / 1e3), 1),
        "pair_evals_per_s": round(evals / (cms / 1e3), 1),
        "roofline": {
            "bound": "valu_int32",
            "frac": round(evals * CHAIN_OPS_PER_EVAL / (cms / 1e3) / 1e12 / INT32_PEAK_TOPS, 4),
            "ops_per_eval": CHAIN_OPS_PER_EVAL,
            "hbm_frac": round(n_anchor * CHAIN_BYTES_PER_ANCHOR / (cms / 1e3) / 1e9 /
                              HBM_PEAK_GBS, 5),
            "bytes_per_anchor": CHAIN_BYTES_PER_ANCHOR,
            "note": "integer profile DP; the DP is sequential in the anchor index, so one "
                    "wavefront per read runs at its latency, not at the VALU or HBM roof",
        },
        "ms": round(cms, 4),
        "f64_ref_default_ms": round(cms64, 4),
        "f64_reads_per_s": round(wc.n_reads / (cms64 / 1e3), 1),
        "cpu_reads_per_s": round(wc.n_reads / ccpu, 1),
        "cpu_threads": threads,
        "anchors_to_chains_ms": round(pms, 3),
        "anchors_to_chains_reads_per_s": round(nrd / (pms / 1e3), 1),
        "anchors_to_chains_with_host_ms": round(pms_host, 3),
        "anchors_to_chains_kept": int(pres.n_kept.sum()),
        "cpu_anchors_to_chains_reads_per_s": round(1.0 / cpu_pipe, 1),
        "workload": "4096 reads, Poisson(400) anchors, lookback 50, max_gap 5000, "
                    "int profile 6/1 (and float64 6/0.05/0.5), seed 1; CPU: whole batch, "
                    "int profile, the oracle's C DP per read on every host core; "
                    "anchors_to_chains: unsorted anchors resident on "
                    "the device, sort + float64 DP + backtrack + select (with_host adds the "
                    "uploads and the readback of the chain records); CPU: the oracle's "
                    "chain_read on every 16th read",
    }
    log(f"[bench] chains: {res}")
    return res


def host_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--skip-extras", action="store_true")
    ap.add_argument("--tile", default=None,
                    help="C,G: run the headline batch on this tile instead of the autotuner's "
                         "pick (profiling runs pin the tile they measure)")
    ap.add_argument("--launch", default=None,
                    help="split_rows=R,split_slices=S,prio_rows=P: run the headline batch with "
                         "this launch instead of the library's default launch (profiling runs; "
                         "no launch autotune); 'default': the library's default launch "
                         "without the autotune; 'auto': the launch autotuner's pick")
    ap.add_argument("--cpu-sample-stride", type=int, default=8)
    ap.add_argument("--serve-requests", type=int, default=200_000,
                    help="requests of the serving stream (0: skip the serving measurement)")
    ap.add_argument("--no-online", action="store_true",
                    help="skip the online-session serving leg (keeps a profiled run short: "
                         "its Python front end makes ~10^5 library calls)")
    ap.add_argument("--serve-prio-rows", type=int, default=SERVE_PRIO_ROWS,
                    help="staged serving: noisy-class requests with at least this many query "
                         "rows run at raised wave priority (0: none)")
    ap.add_argument("--serve-split", default=",".join(str(x) for x in SERVE_SPLIT),
                    help="R,K,G,N: staged serving routes noisy requests of at least R query "
                         "rows to a class of G workgroups that run each as K band slices, the "
                         "noisy class keeping N wavefronts (R = 0: none)")
    ap.add_argument("--serve-load", type=float, default=0.65,
                    help="offered load of the paced serving run, as a fraction of the "
                         "saturation capacity")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (HBM counter calibration)")
    ap.add_argument("--workload", choices=("noisy", "long", "short", "serving", "chains"),
                    default="noisy",
                    help="long / short: only time the long-read or the short-read batch's step "
                         "(profiling runs; --tile pins its tile, else the library's default); "
                         "serving / chains: only that measurement")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(pairs and reads are independent and shard across GPUs with no collective).",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from graphmambaformer_amd import kernels as K
    from graphmambaformer_amd import workloads as WL

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        x = torch.ones(1 << 28, dtype=torch.int32, device=dev)  # 1 GiB
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": x.numel() * 4, "copies": 3}), flush=True)
        return

    threads = os.cpu_count() or 1
    if args.workload == "serving":
        print(json.dumps(serving_extra(torch, WL, dev, args.serve_requests, args.serve_load,
                                       online=not args.no_online,
                                       prio_rows=args.serve_prio_rows,
                                       split=tuple(int(x) for x in args.serve_split.split(",")))),
              flush=True)
        return
    if args.workload == "chains":
        print(json.dumps(chains_extra(torch, K, WL, dev, threads)), flush=True)
        return
    if args.workload != "noisy":
        print(json.dumps(single_batch(torch, K, WL, dev, args)), flush=True)
        return
    # ------------------------------------------------------------ headline
    # the noisy uneven-length batch (8,192 pairs, 0.5-100 kb, band 1503): the
    # largest single-GPU configuration, and the one whose tail the launch
    # choices address
    t0 = time.time()
    wn = WL.sw_noisy()
    ncells = wn.cells()
    log(f"[bench] sw_noisy generated: {wn.n_pairs} pairs, {ncells.sum():.3e} cells, "
        f"{time.time() - t0:.1f}s")
    # tile per length bin: 8 log-spaced bins, each autotuned on itself
    lens = np.array([len(q) for q in wn.queries])
    edges = np.unique(np.round(np.geomspace(lens.min(), lens.max() + 1, 9)).astype(np.int64))
    bin_idx = [np.flatnonzero((lens >= lo) & (lens < hi)) for lo, hi in zip(edges[:-1], edges[1:])]
    bins = []
    for (lo, hi), idx in zip(zip(edges[:-1], edges[1:]), bin_idx):
        if not len(idx):
            continue
        if args.tile:  # pinned tile (profiling runs): no per-bin autotune launches
            bins.append({"len": [int(lo), int(hi)], "pairs": int(len(idx)),
                         "cells": int(ncells[idx].sum()),
                         "tile": [int(x) for x in args.tile.split(",")]})
            continue
        sb = WL.subsample(wn, idx)
        inp_b = K.upload_pairs(sb.queries, sb.targets, sb.band_offset, dev, alphabet="2bit")
        tb_, tms = K.autotune_tile(inp_b, wn.half_band, scoring(wn), reps=2,
                                   order=torch.from_numpy(lpt_order(sb.cells())).to(dev))
        del inp_b
        bins.append({"len": [int(lo), int(hi)], "pairs": int(len(idx)),
                     "cells": int(ncells[idx].sum()),
                     "tile": [tb_.cells_per_lane, tb_.lanes_per_pair],
                     "autotune_ms": tile_timings(tms)})
    torch.cuda.empty_cache()
    picks = {tuple(b["tile"]) for b in bins}
    tile_n = K.resolve_tile(wn.half_band, *next(iter(picks))) if len(picks) == 1 else \
        K.resolve_tile(wn.half_band)
    if args.tile:
        c_, g_ = (int(x) for x in args.tile.split(","))
        tile_n = K.resolve_tile(wn.half_band, c_, g_)
    rn = BatchRun(torch, K, wn, dev, True, tile=tile_n)
    total_cells = rn.total_cells
    # launch choice on the whole batch: plain, the longest pairs at raised
    # priority, or split over several wavefronts
    if args.launch == "default":  # the library's rule alone (profiling runs): no autotune
        launch, launch_ms = dict(K.default_launch(rn.inp, wn.half_band, rn.tile), prio_rows=0), {}
    elif args.launch and args.launch != "auto":  # pinned launch (profiling runs): no autotune
        launch, launch_ms = {"split_rows": 0, "split_slices": 0, "prio_rows": 0}, {}
        for kv in args.launch.split(","):
            k_, v_ = kv.split("=")
            launch[k_] = int(v_)
    else:
        # the library's rule (default_launch) for the headline, so every run
        # times the same launch; the autotuner's table of the alternatives
        # (timed on the whole batch) is reported beside it, and --launch auto
        # takes its pick instead
        tuned, launch_ms = K.autotune_launch(rn.inp, wn.half_band, scoring(wn), rn.order,
                                             tile=rn.tile, workspace=rn.ws)
        launch = dict(K.default_launch(rn.inp, wn.half_band, rn.tile), prio_rows=0)
        if args.launch == "auto":
            launch = tuned
    rn.launch = launch
    log(f"[bench] launch autotune (ms): {launch_ms} -> {launch}")
    log(f"[bench] resident: tb planes {rn.ws['tb_bytes'] / 1e9:.1f} GB, tile "
        f"{rn.tile.cells_per_lane}x{rn.tile.lanes_per_pair}")

    stream = torch.cuda.current_stream()
    fwd_ev = []

    def step(record):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        rn.forward()
        b.record(stream)
        if record:
            fwd_ev.append((a, b))

    for k in range(args.warmup):
        step(False)
        torch.cuda.synchronize()
        log(f"[bench] warmup step {k + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        step(True)
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    ms_step = t_total * 1e3 / max(args.steps, 1)
    fwd_ms = float(np.mean([a.elapsed_time(b) for a, b in fwd_ev]))
    gcups = total_cells / (ms_step / 1e3) / 1e9
    log(f"[bench] timed {args.steps} steps: {ms_step:.1f} ms/step (launch {fwd_ms:.1f} ms, "
        f"walk fused) -> {gcups:.1f} GCUPS")

    inp = rn.inp
    seq_bytes = float((inp.qlen_host.sum() + inp.tlen_host.sum()) / 4)
    alg_bytes = total_cells * TB_BYTES_PER_CELL + seq_bytes
    hbm_gbs = alg_bytes / (fwd_ms / 1e3) / 1e9
    valu_tops = total_cells * OPS_PER_CELL_TB / (fwd_ms / 1e3) / 1e12
    launch_label = ("plain" if not launch["split_rows"] and not launch["prio_rows"] else
                    ",".join(f"{k}={v}" for k, v in launch.items() if v))
    result = {
        "metric": METRIC,
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": "sw_noisy: 8192 pairs, lognormal(8 kb, 1.0) clipped 0.5-100 kb, 6% "
                        "errors (homopolymer-weighted indels), half_band 751, scoring "
                        "2/-4/-4/-2, traceback + CIGAR, seed 3",
            "pairs": wn.n_pairs,
            "cells": total_cells,
            "half_band": wn.half_band,
            "tile": [rn.tile.cells_per_lane, rn.tile.lanes_per_pair],
            "length_bins": bins,
            "launch": launch,
            "launch_autotune_ms": launch_ms,
            "step": "one launch: sw_forward (scores, ends, 4-bit planes) with the CIGAR walk of "
                    "every pair fused, longest first",
        },
        "kernels_ms": {"sw_forward": round(fwd_ms, 3), "walk_mode": "fused"},
        "roofline": {
            "bound": "valu_int32",
            "achieved": round(valu_tops, 3),
            "peak": round(INT32_PEAK_TOPS, 2),
            "unit": "TOP/s",
            "frac": round(valu_tops / INT32_PEAK_TOPS, 4),
            "ops_per_cell": OPS_PER_CELL_TB,
            "kernel": "sw_forward_kernel",
            "traffic": measured_traffic("sw_noisy", total_cells, rn.tile, launch_label),
        },
        "roofline_hbm": {
            "bound": "hbm",
            "achieved": round(hbm_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(hbm_gbs / HBM_PEAK_GBS, 5),
            "note": "algorithmic bytes: 0.5 B of planes per cell + the 2-bit sequences",
        },
    }
    if args.skip_extras:
        print(json.dumps(result), flush=True)
        return
    result["host"] = host_info()

    # the same step captured in a HIP graph
    gn = K.SwGraph(rn.inp, scoring(wn), wn.half_band, order=rn.order, tile=rn.tile,
                   workspace=rn.ws, cigar=True, **launch)
    gnms = time_events(torch, gn.replay, 3)
    del gn
    # the longest pair alone, plain and with the chosen launch
    longest = int(np.argmax(ncells))
    one = BatchRun(torch, K, WL.subsample(wn, [longest]), dev, True, tile=rn.tile)
    one.step()
    ms1 = time_events(torch, one.step, 2)
    one.launch = {"split_rows": 1, "split_slices": launch["split_slices"]} \
        if launch["split_rows"] else dict(launch)
    one.step()
    ms1l = time_events(torch, one.step, 2)
    # what its plain step is made of, and its time over each compiled split
    one_parts = {}
    for label, kw in (("scores_only", dict(traceback=False)), ("with_planes", dict(traceback=True))):
        fn = lambda kw=kw: K.sw_forward(one.inp, *scoring(wn), wn.half_band, order=one.order,  # noqa: E731
                                        tile=one.tile, workspace=one.ws, **kw)
        fn()
        one_parts[label] = round(time_events(torch, fn, 2), 3)
    one_parts["with_fused_walk"] = round(ms1, 3)
    one_split = {}
    for _, sl in K.split_variants(wn.half_band, rn.tile.cells_per_lane):
        one.launch = {"split_rows": 1, "split_slices": sl}
        one.step()
        one_split[f"{sl}_slices"] = round(time_events(torch, one.step, 2), 3)
    del one
    # each length bin as a batch of its own (its autotuned tile): plain launch,
    # and the library's default launch for that batch
    for b, idx in zip(bins, [i for i in bin_idx if len(i)]):
        rb = BatchRun(torch, K, WL.subsample(wn, idx), dev, True,
                      tile=K.resolve_tile(wn.half_band, *b["tile"]))
        rb.step()
        bms = time_events(torch, rb.step, 2)
        b["ms"] = round(bms, 3)
        b["gcups"] = round(b["cells"] / (bms / 1e3) / 1e9, 2)
        dl = K.default_launch(rb.inp, wn.half_band, rb.tile)
        if dl["split_rows"]:
            rb.launch = dl
            rb.step()
            bms = time_events(torch, rb.step, 2)
            b["default_launch"] = dl
            b["default_launch_ms"] = round(bms, 3)
            b["default_launch_gcups"] = round(b["cells"] / (bms / 1e3) / 1e9, 2)
        del rb
    del rn
    torch.cuda.empty_cache()
    srt = np.argsort(ncells, kind="stable")
    ppn = per_pair_gcups(torch, K, wn, srt[::256], dev)
    subn = srt[::args.cpu_sample_stride]
    cgn, csn = cpu_gcups(wn, subn, threads)
    result["per_pair_gpu"] = {"value": round(ppn, 4), "unit": "GCUPS",
                              "sample": f"{len(srt[::256])} sw_noisy pairs (every 256th by cell "
                                        "count), one forward launch + one walk launch per pair"}
    result["batched_over_per_pair"] = round(gcups / ppn, 1)
    result["cpu_baseline"] = {
        "value": round(cgn, 4),
        "unit": "GCUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"{len(subn)} sw_noisy pairs (every {args.cpu_sample_stride}th by cell count, "
                  f"{int(ncells[subn].sum()):.3e} cells), forward + traceback in one C call "
                  f"on {threads} threads, {csn:.1f} s",
    }
    result["gpu_over_cpu"] = round(gcups / cgn, 1)
    result["sw_noisy"] = {
        "graph_ms": round(gnms, 3),
        "graph_gcups": round(total_cells / (gnms / 1e3) / 1e9, 2),
        "longest_pair_ms": round(ms1, 3),
        "longest_pair_launch_ms": round(ms1l, 3),
        "longest_pair_cells": int(ncells[longest]),
        "longest_pair_frac_of_batch": round(ms1l / ms_step, 3),
        "longest_pair_parts_ms": one_parts,
        "longest_pair_split_ms": one_split,
        "length_bins_note": "each bin also run as its own batch (its tile): plain launch, and "
                            "the library's default launch where it splits",
    }
    log(f"[bench] sw_noisy extras: {result['sw_noisy']}, per-pair {ppn:.3f}, CPU {cgn:.3f} "
        f"GCUPS on {threads} threads")

    # ------------------------------------------------------- long-read batch
    result["sw_long"] = long_read_extra(torch, K, WL, dev, args, threads)

    # ------------------------------------------------------- short-read batch
    ws_ = WL.sw_short()
    inp_s = K.upload_pairs(ws_.queries, ws_.targets, ws_.band_offset, dev, alphabet="2bit")
    tuned_s, tune_s = K.autotune_tile(inp_s, ws_.half_band, scoring(ws_), reps=5,
                                      order=torch.from_numpy(lpt_order(ws_.cells())).to(dev))
    del inp_s
    rs = BatchRun(torch, K, ws_, dev, True, tile=tuned_s)
    rs.step()
    # a launch of this batch lasts ~36 us, so one slow launch among a few
    # moves the mean by several percent: the median of 20 launches (the
    # mean beside it)
    ms = time_events(torch, rs.step, 20, stat="median")
    ms_mean = time_events(torch, rs.step, 20)
    gr = K.SwGraph(rs.inp, scoring(ws_), ws_.half_band, order=rs.order, tile=rs.tile,
                   workspace=rs.ws, cigar=True)
    gms = time_events(torch, gr.replay, 20, stat="median")
    # what the step is made of: the forward pass without planes, with planes,
    # with the fused walk (medians of 20 launches each)
    def part_ms(**kw):
        fn = lambda: K.sw_forward(rs.inp, *scoring(ws_), ws_.half_band, order=rs.order,  # noqa: E731
                                  tile=rs.tile, workspace=rs.ws, **kw)
        fn()
        s_ = torch.cuda.current_stream()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(20)]
        for a_, b_ in ev:
            a_.record(s_)
            fn()
            b_.record(s_)
        torch.cuda.synchronize()
        return round(float(np.median([a_.elapsed_time(b_) for a_, b_ in ev])), 4)

    parts = {"scores_only": part_ms(traceback=False), "with_planes": part_ms(traceback=True),
             "with_fused_walk": part_ms(traceback=True, cigar=True)}
    pps = per_pair_gcups(torch, K, ws_, np.arange(ws_.n_pairs), dev)
    cgs, css = cpu_gcups(ws_, np.arange(ws_.n_pairs), threads)
    result["sw_short"] = {
        "batched_gcups": round(rs.total_cells / (ms / 1e3) / 1e9, 3),
        "batched_ms": round(ms, 4),
        "batched_mean_ms": round(ms_mean, 4),
        "timing": "median of 20 launches (batched_mean_ms: their mean); graph: median of 20 "
                  "replays",
        "graph_gcups": round(rs.total_cells / (gms / 1e3) / 1e9, 3),
        "graph_ms": round(gms, 4),
        "per_pair_gcups": round(pps, 4),
        "cpu_gcups": round(cgs, 4),
        "cpu_threads": threads,
        "tile": [rs.tile.cells_per_lane, rs.tile.lanes_per_pair],
        "autotune_ms": tile_timings(tune_s),
        "parts_ms": parts,
        "workload": "1024 x 150 bp, half_band 32 at offset 32, seed 0 (CPU: whole batch, one C "
                    "call)",
    }
    log(f"[bench] sw_short: {result['sw_short']}")
    del rs, gr
    torch.cuda.empty_cache()

    # ----------------------------------------------------------------- serving
    if args.serve_requests > 0:
        result["serving"] = serving_extra(torch, WL, dev, args.serve_requests, args.serve_load,
                                          online=not args.no_online,
                                          prio_rows=args.serve_prio_rows,
                                          split=tuple(int(x) for x in args.serve_split.split(",")))
        torch.cuda.empty_cache()

    # ------------------------------------------------------------------ chains
    result["chains"] = chains_extra(torch, K, WL, dev, threads)

    # ----------------------------------------------------------------- seeding
    result["seeding"] = seeding_extra(torch, WL, dev)
    result["wfa"] = wfa_extra(torch, K, WL, dev)

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

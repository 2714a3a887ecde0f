#!/usr/bin/env python3
"""Benchmark of the banded affine SW + traceback path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--skip-extras]

A step is one pass of the hot path over the whole long-read batch (16,384
pairs, lengths lognormal median 12 kb clipped 5-25 kb, 0.5% errors, band
half-width 256, scoring 2/-4/-4/-2): the forward kernel writes scores, ends
and 4-bit traceback planes, then the walk turns every pair into a CIGAR.
Inputs are resident in HBM before the timed region.  ``value`` is GCUPS over
in-band cells.  Beside it the line carries the per-pair-launch GPU form, the
CPU oracle on the host cores (bounded stratified sample), the short-read
batch, the chaining DP and the roofline of the dominant kernel.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "banded affine-SW GCUPS batched vs per-pair; p50/p99 latency at fixed QPS"
HBM_PEAK_GBS = 8000.0
INT32_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12  # 39.3 T int32 VALU lane-ops/s
OPS_PER_CELL_TB = 20  # int32 ops per cell with traceback (score/end 14 + traceback 6)
TB_BYTES_PER_CELL = 0.5


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(cells: int, fused: bool):
    """HBM bytes per forward launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs (calibrated against a plain 1 GiB copy), if they were
    taken on this exact batch and walk mode; else None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "sw_forward_sw_long_traffic.json")
    try:
        with open(path) as fh:
            rec = json.load(fh)
    except (OSError, ValueError):
        return None
    if rec.get("cells") != cells or not fused:
        return None
    return rec.get("traffic_bytes_per_launch")


def lpt_order(cells: np.ndarray):
    return np.argsort(-cells, kind="stable").astype(np.int32)


def time_events(torch, fn, reps):
    """Mean ms of fn() over reps, HIP events on torch's current stream (the
    stream the library launches on)."""
    s = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    for a, b in ev:
        a.record(s)
        fn()
        b.record(s)
    torch.cuda.synchronize()
    return float(np.mean([a.elapsed_time(b) for a, b in ev]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--skip-extras", action="store_true")
    ap.add_argument("--walk", choices=["separate", "fused"], default="fused",
                    help="CIGAR walk as its own launch (one wavefront per pair) or inside "
                         "the forward launch")
    ap.add_argument("--cpu-sample-stride", type=int, default=32)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(pairs and reads shard with no collective; see DESIGN.md).", file=sys.stderr)
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

    # ------------------------------------------------------------------ batch
    t0 = time.time()
    w = WL.sw_long()
    cells = w.cells()
    total_cells = int(cells.sum())
    log(f"[bench] sw_long generated: {w.n_pairs} pairs, {total_cells:.3e} cells, "
        f"{time.time() - t0:.1f}s")
    inp = K.upload_pairs(w.queries, w.targets, w.band_offset, dev, alphabet="2bit")
    order = torch.from_numpy(lpt_order(cells)).to(dev)
    tile = K.resolve_tile(w.half_band)
    sc = (w.match, w.mismatch, w.gap_open, w.gap_extend)
    ws = {"counter": torch.zeros(1, dtype=torch.int32, device=dev)}
    nb = K.tb_bytes(inp.qlen_host, tile)
    offs = np.zeros(w.n_pairs + 1, np.int64)
    np.cumsum(nb // 4, out=offs[1:])
    ws["tb"] = torch.empty(int(offs[-1]), dtype=torch.int32, device=dev)
    ws["tb_off"] = torch.from_numpy(offs[:-1]).to(dev)
    cap = inp.qlen_host.astype(np.int64) + inp.tlen_host.astype(np.int64) + 1
    coffs = np.zeros(w.n_pairs + 1, np.int64)
    np.cumsum(cap, out=coffs[1:])
    ws["cigar"] = torch.empty(int(coffs[-1]), dtype=torch.int32, device=dev)
    ws["cigar_off"] = torch.from_numpy(coffs[:-1]).to(dev)
    torch.cuda.synchronize()
    log(f"[bench] resident: tb planes {offs[-1] * 4 / 1e9:.1f} GB, tile "
        f"{tile.cells_per_lane}x{tile.lanes_per_pair}")

    stream = torch.cuda.current_stream()
    fwd_ev = []

    fused = args.walk == "fused"
    tb_ev = []

    def step(record):
        # forward rows + 4-bit planes, then every pair's CIGAR walk: either a
        # second launch with one wavefront per pair ("separate") or inside the
        # forward launch right after each pair's last row ("fused")
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        c = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        fwd = K.sw_forward(inp, *sc, w.half_band, traceback=True, order=order, tile=tile,
                           workspace=ws, cigar=fused)
        b.record(stream)
        tbk = fwd.walk if fused else K.sw_traceback(inp, fwd, workspace=ws)
        c.record(stream)
        if record:
            fwd_ev.append((a, b))
            tb_ev.append((b, c))
        return fwd, tbk

    for k in range(args.warmup):
        step(False)
        torch.cuda.synchronize()
        log(f"[bench] warmup step {k + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        fwd, tbk = step(True)
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    ms_step = t_total * 1e3 / max(args.steps, 1)
    fwd_ms = float(np.mean([a.elapsed_time(b) for a, b in fwd_ev]))
    tb_ms = float(np.mean([a.elapsed_time(b) for a, b in tb_ev]))
    gcups = total_cells / (ms_step / 1e3) / 1e9
    log(f"[bench] timed {args.steps} steps: {ms_step:.1f} ms/step (forward {fwd_ms:.1f} ms, "
        f"walk {tb_ms:.1f} ms, {args.walk}) -> {gcups:.1f} GCUPS")

    # algorithmic traffic of the forward kernel: 0.5 B/cell of traceback planes
    # + the 2-bit sequences read once
    seq_bytes = float((inp.qlen_host.sum() + inp.tlen_host.sum()) / 4)
    alg_bytes = total_cells * TB_BYTES_PER_CELL + seq_bytes
    hbm_gbs = alg_bytes / (fwd_ms / 1e3) / 1e9
    valu_tops = total_cells * OPS_PER_CELL_TB / (fwd_ms / 1e3) / 1e12
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
            "workload": "sw_long: 16384 pairs, lognormal(12 kb, 0.4) clipped 5-25 kb, 0.5% errors, "
                        "half_band 256, scoring 2/-4/-4/-2, traceback + CIGAR, seed 2",
            "pairs": w.n_pairs,
            "cells": total_cells,
            "half_band": w.half_band,
            "tile": [tile.cells_per_lane, tile.lanes_per_pair],
            "step": "sw_forward (scores, ends, 4-bit planes) + CIGAR walk of every pair "
                    f"({args.walk})",
        },
        "kernels_ms": {"sw_forward": round(fwd_ms, 3), "sw_walk": round(tb_ms, 3),
                       "walk_mode": args.walk},
        "roofline": {
            "bound": "hbm",
            "achieved": round(hbm_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(hbm_gbs / HBM_PEAK_GBS, 5),
            "traffic": measured_traffic(total_cells, fused),
            "kernel": "sw_forward_kernel",
            "note": "the kernel is int32-VALU bound; see roofline_valu",
        },
        "roofline_valu": {
            "bound": "valu_int32",
            "achieved": round(valu_tops, 3),
            "peak": round(INT32_PEAK_TOPS, 2),
            "unit": "TOP/s",
            "frac": round(valu_tops / INT32_PEAK_TOPS, 4),
            "ops_per_cell": OPS_PER_CELL_TB,
        },
    }

    # ------------------------------------------------------------- per-pair form
    if not args.skip_extras:
        sub = np.arange(0, w.n_pairs, 256)[:64]
        ws1 = {"counter": torch.zeros(1, dtype=torch.int32, device=dev)}
        sw1 = WL.subsample(w, sub)
        inps = [K.upload_pairs([sw1.queries[k]], [sw1.targets[k]], sw1.band_offset[k:k + 1], dev,
                               alphabet="2bit") for k in range(sw1.n_pairs)]
        sub_cells = int(cells[sub].sum())

        def per_pair():
            for one in inps:
                f1 = K.sw_forward(one, *sc, w.half_band, traceback=True, tile=tile,
                                  workspace=ws1)
                K.sw_traceback(one, f1)

        per_pair()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        per_pair()
        torch.cuda.synchronize()
        pp_s = time.perf_counter() - t1
        result["per_pair_gpu"] = {"value": round(sub_cells / pp_s / 1e9, 3), "unit": "GCUPS",
                                  "sample": f"{len(sub)} pairs of sw_long (every 256th), one "
                                            "forward launch + one walk launch per pair"}
        result["batched_over_per_pair"] = round(gcups / (sub_cells / pp_s / 1e9), 1)
        log(f"[bench] per-pair GPU form: {result['per_pair_gpu']['value']} GCUPS")

    # --------------------------------------------------------------- CPU oracle
    if not args.skip_extras:
        from oracle import oracle as O

        O.build() if not os.path.exists(O._LIB_PATH) else None
        srt = np.argsort(cells, kind="stable")
        sub = srt[::args.cpu_sample_stride]
        threads = min(os.cpu_count() or 1, 16)
        qs = [w.as_ref_codes(int(k))[0] for k in sub]
        ts = [w.as_ref_codes(int(k))[1] for k in sub]
        t1 = time.perf_counter()
        O.sw_align_pairs(qs, ts, w.band_offset[sub], w.half_band, *sc, threads=threads)
        cpu_s = time.perf_counter() - t1
        cpu_gcups = float(cells[sub].sum()) / cpu_s / 1e9
        result["cpu_baseline"] = {
            "value": round(cpu_gcups, 4),
            "unit": "GCUPS",
            "cores": threads,
            "kind": "port",
            "sample": f"{len(sub)} sw_long pairs (every {args.cpu_sample_stride}th by cell count, "
                      f"{int(cells[sub].sum()):.3e} cells), forward + traceback, one pair per "
                      f"thread task, {cpu_s:.1f} s",
        }
        result["gpu_over_cpu"] = round(gcups / cpu_gcups, 1)
        log(f"[bench] CPU oracle: {cpu_gcups:.3f} GCUPS on {threads} threads")

    # ------------------------------------------------------- short reads, chains
    if not args.skip_extras:
        ws_ = WL.sw_short()
        inp_s = K.upload_pairs(ws_.queries, ws_.targets, ws_.band_offset, dev, alphabet="2bit")
        cs = ws_.cells()
        tile_s = K.resolve_tile(ws_.half_band)
        ords = torch.from_numpy(lpt_order(cs)).to(dev)
        wss = {"counter": torch.zeros(1, dtype=torch.int32, device=dev)}

        def short_step():
            f = K.sw_forward(inp_s, *sc, ws_.half_band, traceback=True, order=ords, tile=tile_s,
                             workspace=wss)
            K.sw_traceback(inp_s, f)

        short_step()
        ms = time_events(torch, short_step, 10)
        inps1 = [K.upload_pairs([ws_.queries[k]], [ws_.targets[k]], ws_.band_offset[k:k + 1], dev,
                                alphabet="2bit") for k in range(ws_.n_pairs)]

        def short_pp():
            for one in inps1:
                f1 = K.sw_forward(one, *sc, ws_.half_band, traceback=True, tile=tile_s,
                                  workspace=wss)
                K.sw_traceback(one, f1)

        short_pp()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        short_pp()
        torch.cuda.synchronize()
        pps = time.perf_counter() - t1
        result["sw_short"] = {
            "batched_gcups": round(int(cs.sum()) / (ms / 1e3) / 1e9, 3),
            "per_pair_gcups": round(int(cs.sum()) / pps / 1e9, 4),
            "batched_ms": round(ms, 4),
            "tile": [tile_s.cells_per_lane, tile_s.lanes_per_pair],
            "workload": "1024 x 150 bp, half_band 32 at offset 32, seed 0",
        }
        log(f"[bench] sw_short: {result['sw_short']}")

        wc = WL.chains()
        d = dev
        re_ = torch.from_numpy(wc.read_end).to(d)
        fe_ = torch.from_numpy(wc.ref_end).to(d)
        wt = torch.from_numpy(wc.weight).to(d)
        so = torch.from_numpy(wc.seg_off).to(d)

        def chain_step():
            K.chain_dp(re_, fe_, wt, so, wc.max_lookback, wc.max_gap, gap_open=wc.gap_open,
                       gap_extend=wc.gap_extend)

        chain_step()
        cms = time_events(torch, chain_step, 5)
        result["chains"] = {
            "reads_per_s": round(wc.n_reads / (cms / 1e3), 1),
            "pair_evals_per_s": round(wc.pair_evaluations() / (cms / 1e3), 1),
            "ms": round(cms, 4),
            "workload": "4096 reads, Poisson(400) anchors, lookback 50, max_gap 5000, "
                        "int profile 6/1, seed 1",
        }
        log(f"[bench] chains: {result['chains']}")

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

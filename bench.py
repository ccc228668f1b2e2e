#!/usr/bin/env python3
"""Benchmark of the banded Smith-Waterman extension engine on one MI355X.

Default workload (2): 2,000,000 extension pairs, query length U[30,250],
reference = query + U[0,100] flank, 2% mismatches and 0.5% indels, h0
U[0,120], w = 100, scoring of bench-bsw (match 1, mismatch 4, gap open 6 /
extend 1 both ways, z-drop 100, end bonus 5).  A step is one pass of the
8/16-bit dispatcher over the whole batch with the pairs and sequences already
resident in HBM, replayed from a hipGraph-captured plan after the autotuner
has picked the launch configuration; outputs are written in place.
--workload 1/3/5 measure the other batch workloads the same way (workload 3
as one plan per band class, w in {50,100,200,500}); --serve measures streamed
serving latency.

value = band cells / second (GCUPS) of the whole step, where a pair's band
cells are sum_i max(0, min(len2, i+b+1) - max(0, i-b)) with b its band
(bandedSWA.cpp:2882-2904).  Also reported: GCUPS over rectangle cells
len1*len2 (bench-smith-waterman.cpp:150's count), pairs/s, the eager
(ungraphed) dispatcher, the naive form (one launch per 20,000-pair bucket),
the plain form (per-chunk getScores16 host calls, as bench-bsw drives the
reference), the end-to-end rate with the whole batch moved from pinned host
buffers and back every step, the roofline of the dominant kernel against
integer VALU peak, and the CPU baseline on a bounded sample.  The default run
then adds the rest of the metric: workloads 1, 3 and 5 (plan GCUPS, naive
ratio, parity prefix) under "workloads", and streamed-serving p50/p99 latency
at 50/80/95% of peak under "serving".
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "trans-omics-acceleration-library_amd")
sys.path.insert(0, PKG)

import synth  # noqa: E402
import tal_bsw  # noqa: E402

# MI355X: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz = 39.3e12 int32 lane-ops/s.
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
OPS_PER_CELL = 14  # S select 2, ambiguity 1, M 2, H 1, E 3, F 3, row max 2
PARAMS = dict(o_del=6, e_del=1, o_ins=6, e_ins=1, zdrop=100, end_bonus=5, w_match=1, w_mismatch=4)
# parity is checked against the oracle in the CPU leg only; without it the
# GPU parity tests (tests/) are the check
NO_CHECK = "not checked by this run (--no-cpu); tests/ -m gpu check it"


WORKLOADS = {
    1: {"pairs": 100_000, "name": "plain 16-bit extension, 151-bp queries",
        "shape": {"query_len": 151, "ref_len": "U[151,250]", "h0": "U[10,60]", "w": 100}},
    2: {"pairs": 2_000_000, "name": "uneven batch, 8/16-bit dispatch",
        "shape": {"query_len": "U[30,250]", "ref_flank": "U[0,100]", "h0": "U[0,120]", "w": 100}},
    3: {"pairs": 500_000, "name": "long extensions, one plan per band class",
        "shape": {"query_len": "lognormal median 400, [100,2000]", "ref_flank": "U[0,100]",
                  "w": "{50,100,200,500} per pair"}},
    5: {"pairs": 1_000_000, "name": "edge stress (N, repeats, z-drop, h0 up to 30000)",
        "shape": {"query_len": "as workload 1", "h0": "log-uniform [10,30000]", "w": 100}},
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", type=int, default=2, choices=sorted(WORKLOADS),
                    help="1 plain 16-bit, 2 uneven 8/16-bit batch (the headline), "
                         "3 long extensions per band class, 5 edge stress")
    ap.add_argument("--pairs", type=int, default=0, help="pairs (0 = the workload's size)")
    ap.add_argument("--plain-pairs", type=int, default=200_000)
    ap.add_argument("--cpu-pairs", type=int, default=0,
                    help="pairs of the CPU baseline's sample (0 = the whole batch)")
    ap.add_argument("--check-pairs", type=int, default=20_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-naive", action="store_true")
    ap.add_argument("--no-e2e", action="store_true",
                    help="skip the end-to-end rate from pinned host buffers")
    ap.add_argument("--no-others", action="store_true",
                    help="default run: skip workloads 1, 3 and 5")
    ap.add_argument("--no-serve", action="store_true",
                    help="default run: skip the streamed-serving latencies")
    ap.add_argument("--eager", action="store_true",
                    help="time the eager dispatcher instead of the graph-captured plan")
    ap.add_argument("--tune-pairs", type=int, default=262_144,
                    help="pairs the autotuner times its candidates on (0 = no tuning)")
    ap.add_argument("--bucket", type=int, default=20_000,
                    help="pairs per launch of the naive form (bench-bsw's chunk size)")
    ap.add_argument("--serve", action="store_true",
                    help="streamed serving instead: Poisson arrivals of 1k-64k-pair requests, "
                         "p50/p99 latency at 50/80/95%% of the measured peak")
    ap.add_argument("--serve-requests", type=int, default=400, help="requests per load level")
    ap.add_argument("--serve-mode", default="resident", choices=["resident", "host"],
                    help="serving loop of --serve: the device-resident kernel or host-launched "
                         "batches (the default run measures both)")
    ap.add_argument("--sw", default="", choices=["", *sorted(synth.SW_WORKLOADS)],
                    help="measure the full inter-task SW engine (PairWiseSW drop-in) on one of "
                         "its synthetic workloads instead")
    ap.add_argument("--chain", action="store_true",
                    help="measure the DP chaining engine (dp_chain::mm_dp_vectorized drop-in) "
                         "on its synthetic anchor sets instead")
    ap.add_argument("--extend", action="store_true",
                    help="measure the end-to-end extension pipeline (anchors -> chaining -> "
                         "extension pairs -> banded SW of both sides) on simulated reads instead")
    ap.add_argument("--sw-weights", default="1,-1,1,1", metavar="M,X,O,E",
                    help="full-SW weights (match, mismatch, gap open, gap extend as the reference "
                         "adds them); default: bench-smith-waterman's")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (for scaling the "
                         "FETCH_SIZE/WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        _copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU: the engine has no multi-GPU path. Pairs, anchor sets "
              "and reads are independent, so a batch would shard by contiguous ranges, one "
              "process and handle per GPU, with no collectives on the hot path.", file=sys.stderr)
        sys.exit(2)
    if args.serve:
        _serve(args)
        return
    if args.sw:
        _sw(args)
        return
    if args.chain:
        _chain(args)
        return
    if args.extend:
        _extend(args)
        return

    _batch(args)


# bench-smith-waterman's default weights (bench-smith-waterman.cpp:51-54);
# --sw-weights replaces them
SW_PARAMS = (1, -1, 1, 1)
# VALU lane-ops per cell: packed 16-bit classes xor, neq, mad, add, 3 max for
# H, best, H+open, 2 for E, 2 for F = 13 (two cells per 32-bit lane op, so the
# peak is 2 x VALU_PEAK_TOPS); the 32-bit class has compare and select for the
# score instead of neq and mad, and max3: 11, one cell per lane op
SW_OPS = {8: 13, 16: 13, 32: 11}


def _sw(args):
    """Full inter-task SW, PairWiseSW_{8,16,32}::getScores drop-in: one step
    scores the whole synthetic batch from HBM (bounds kernel, one wait for
    the batch bounds, unit sort, DP kernel); value = sum of len1 * len2 over
    the batch / step time (bench-smith-waterman.cpp:150,246)."""
    import torch
    import tal_sw
    global SW_PARAMS
    SW_PARAMS = tuple(int(x) for x in args.sw_weights.split(","))
    bits, n_def, lo, hi = synth.SW_WORKLOADS[args.sw]
    n = args.pairs or n_def
    t = time.time()
    pairs, buf = synth.sw_generate(bits, n, lo, hi)
    cells = float(tal_sw.cells(pairs).sum())
    log(f"[sw] {args.sw}: {n} pairs len U[{lo},{hi}] class {bits}-bit in {time.time() - t:.1f}s; "
        f"cells {cells:.3e}")
    dev = torch.device("cuda:0")
    d_buf = torch.from_numpy(buf).to(dev)
    d_pairs = torch.from_numpy(pairs.view(np.uint8)).to(dev)
    eng = tal_sw.PairWiseSW(*SW_PARAMS)
    stream = torch.cuda.current_stream(dev).cuda_stream

    def step():
        eng.get_scores_device(bits, d_pairs.data_ptr(), d_buf.data_ptr(), buf.nbytes, n, stream)

    step()
    torch.cuda.synchronize()
    first = d_pairs.cpu().numpy().view(tal_sw.PAIR_DTYPE).copy()  # checked by the CPU leg
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        kms.append(eng.stats()["ms_kernel"])
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    st = eng.stats()
    k_ms = float(np.mean(kms))
    peak = VALU_PEAK_TOPS * (2 if bits != 32 else 1)
    achieved = SW_OPS[bits] * cells / (k_ms / 1e3) / 1e12
    line = {
        "metric": f"GCUPS (len1*len2 cells/s) of bit-exact full SW, PairWiseSW_{bits} drop-in",
        "value": round(cells / dt / 1e9, 2), "unit": "GCUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {8: "int8 (as int16 << 8)", 16: "int16", 32: "int32"}[bits],
        "data": f"synthetic (seeded, {args.sw})",
        "config": {"workload": args.sw, "pairs": n, "len": f"U[{lo},{hi}]", "class": bits,
                   "weights": SW_PARAMS, "cells": cells},
        "pairs_per_s": round(n / dt, 1), "parity": NO_CHECK,
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak, 2),
                     "unit": "TOP/s", "frac": round(achieved / peak, 4), "traffic": None,
                     "kernel": f"sw_main<{bits}>", "kernel_ms": round(k_ms, 3),
                     "kernel_gcups": round(cells / (k_ms / 1e3) / 1e9, 1),
                     "ops_per_cell": SW_OPS[bits],
                     "padded_rows_cols": [st["max_rows"], st["max_cols"]],
                     "regrouped": bool(st["regrouped"])},
    }
    if not args.no_cpu:
        # the CPU leg: the oracle checks the first call, then the reference
        # build (or the oracle) is timed
        g = tal_sw.GROUP[bits]
        m = min(n, max(g, (args.check_pairs // g) * g))
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        want = O.sw_scores(bits, pairs[:m].copy(), buf, SW_PARAMS, nthreads=16)["score"]
        bad = int((first["score"][:m] != want).sum())
        if bad:
            raise SystemExit(f"[sw] parity FAILED on {bad} of the first {m} pairs")
        log(f"[sw] parity: first {m} pairs bit-exact against the oracle")
        line["parity"] = f"bit-exact on the first {m} pairs"
        line["cpu_baseline"] = _sw_cpu(bits, pairs, buf, n if bits != 8 else 200_000)
    print(json.dumps(line), flush=True)


def _sw_cpu(bits, pairs, buf, m, min_s=10.0):
    """The reference's own getScores (AVX-512BW, 16 threads) when oracle/_ref
    holds it and the host runs it, else the oracle (16 threads); whole passes
    over the first m pairs until min_s seconds of CPU work."""
    import oracle as O
    import tal_sw
    g = tal_sw.GROUP[bits]
    m = min(len(pairs), max(g, (m // g) * g))
    sub = pairs[:m].copy()
    kind = "reference"
    if not O.ref_sw_available() or O.ref_sw_scores(bits, sub[:g].copy(), buf, SW_PARAMS) is None:
        kind = "port"

    def one():
        if kind == "reference":
            O.ref_sw_scores(bits, sub, buf, SW_PARAMS, nthreads=16)
        else:
            O.sw_scores(bits, sub, buf, SW_PARAMS, nthreads=16)

    passes = 0
    t = time.perf_counter()
    while True:
        one()
        passes += 1
        dt = time.perf_counter() - t
        if dt >= min_s:
            break
    cells = float(tal_sw.cells(sub).sum()) * passes
    return {"value": round(cells / dt / 1e9, 3), "unit": "GCUPS", "cores": 16, "kind": kind,
            "sample": f"{passes} passes over the first {m} pairs, {dt:.1f}s"}


# Algorithmic lane-ops per candidate predecessor, those of the reference's
# vector loop (parallel_chaining_32_bit.h:262-310): ddr, ddq, ddr - ddq and
# its absolute value 5, the five window tests and their conjunction 9,
# min(ddr, ddq, l) 2, ilog2 2, the float gap 4, the score 3, the running
# maximum and its index 5.  (rocprofv3 counts 32 VALU lane-ops per
# predecessor in the whole kernel, per-anchor work included.)
CHAIN_OPS = 30


def _chain(args):
    """DP chaining, dp_chain::mm_dp_vectorized over many anchor sets: one
    step chains the whole synthetic batch from HBM; value = anchors / s."""
    import torch
    import tal_chain
    nsets, lo, hi = synth.CHAIN_WORKLOADS["chain"]
    t = time.time()
    a, off = synth.chain_generate(nsets, lo, hi)
    n = len(a)
    log(f"[chain] {nsets} sets of U[{lo},{hi}] anchors ({n} anchors) in {time.time() - t:.1f}s")
    dev = torch.device("cuda:0")
    d_a = torch.from_numpy(a.view(np.uint8)).to(dev)
    d_off = torch.from_numpy(off).to(dev)
    d_f = torch.zeros(n, dtype=torch.int32, device=dev)
    d_p = torch.zeros(n, dtype=torch.int32, device=dev)
    d_v = torch.zeros(n, dtype=torch.int32, device=dev)
    eng = tal_chain.DpChain()
    stream = torch.cuda.current_stream(dev).cuda_stream

    def step():
        eng.dp_device(d_a.data_ptr(), d_off.data_ptr(), nsets, d_f.data_ptr(), d_p.data_ptr(),
                      d_v.data_ptr(), stream=stream)

    step()
    torch.cuda.synchronize()
    first = (d_f.cpu().numpy().view(np.uint32).copy(), d_p.cpu().numpy().copy(),
             d_v.cpu().numpy().copy())  # checked by the CPU leg
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        kms.append(eng.stats()["ms_kernel"])
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    st = eng.stats()
    k_ms = float(np.mean(kms))
    pred = float(st["predecessors"])
    achieved = CHAIN_OPS * pred / (k_ms / 1e3) / 1e12
    line = {
        "metric": "anchors/s of bit-exact DP chaining, dp_chain::mm_dp_vectorized drop-in",
        "value": round(n / dt, 1), "unit": "anchors/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32 (+ f32/f64 gap cost)",
        "data": "synthetic (seeded anchor sets)",
        "config": {"workload": "chain", "sets": nsets, "anchors_per_set": f"U[{lo},{hi}]",
                   "anchors": n, "bw": 500, "max_iter": 5000, "dr": 5000, "dq": 5000},
        "predecessors_per_s": round(pred / dt, 1),
        "parity": NO_CHECK,
        "roofline": {"bound": "valu", "achieved": round(achieved, 3),
                     "peak": round(VALU_PEAK_TOPS, 2), "unit": "TOP/s",
                     "frac": round(achieved / VALU_PEAK_TOPS, 4),
                     "traffic": (_traffic("chain_main_chain") or {}).get("traffic_bytes"),
                     "kernel": "chain_main", "kernel_ms": round(k_ms, 3),
                     "ops_per_predecessor": CHAIN_OPS, "lds_ring": st["lds_ring"],
                     "far_anchors": st["far_anchors"]},
    }
    if not args.no_cpu:
        # the CPU leg: the oracle checks the first call, then the reference
        # build is timed
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        m = min(nsets, 200)
        want = O.chain_dp_sets(a[:off[m]], off[:m + 1], nthreads=16)
        bad = sum(int((x[:off[m]] != y).sum()) for x, y in zip(first, want))
        if bad:
            raise SystemExit(f"[chain] parity FAILED: {bad} outputs differ in the first {m} sets")
        log(f"[chain] parity: first {m} sets bit-exact against the oracle")
        line["parity"] = f"bit-exact on the first {m} sets"
        line["cpu_baseline"] = _chain_cpu(a, off)
    print(json.dumps(line), flush=True)


def _chain_cpu(a, off, min_s=10.0):
    """The reference's own mm_dp_vectorized (AVX-512 build in oracle/_ref) on
    16 threads, sets dealt out to the threads, else the oracle on 16 threads;
    whole passes over the batch until min_s seconds."""
    import oracle as O
    from concurrent.futures import ThreadPoolExecutor
    nsets = len(off) - 1
    kind = "reference"
    if not O.ref_chain_available() or O.ref_chain_dp_sets(a[:off[1]], off[:2]) is None:
        kind = "port"
    chunks = [(k, min(nsets, k + 50)) for k in range(0, nsets, 50)]

    def part(c):
        s0, s1 = c
        sub = a[off[s0]:off[s1]]
        so = off[s0:s1 + 1] - off[s0]
        if kind == "reference":
            O.ref_chain_dp_sets(sub, so)
        else:
            O.chain_dp_sets(sub, so, nthreads=1)

    passes = 0
    t = time.perf_counter()
    with ThreadPoolExecutor(16) as ex:
        while True:
            list(ex.map(part, chunks))
            passes += 1
            dt = time.perf_counter() - t
            if dt >= min_s:
                break
    return {"value": round(len(a) * passes / dt, 1), "unit": "anchors/s", "cores": 16,
            "kind": kind, "sample": f"{passes} passes over all {nsets} sets, {dt:.1f}s"}


def _extend(args):
    """The extension pipeline (include/tal_extend.h) end to end on simulated
    reads: one step is one tal_extend_run_device call over every read, inputs
    in HBM; value = reads / s."""
    import torch
    import tal_extend
    nreads, rlen, glen = synth.EXTEND_WORKLOADS["extend"]
    t = time.time()
    d = synth.extend_generate(nreads, rlen, glen)
    na = len(d["anchors"])
    log(f"[extend] {nreads} reads of {rlen} from a {glen}-base genome, {na} anchors, "
        f"in {time.time() - t:.1f}s")
    dev = torch.device("cuda:0")
    d_g = torch.from_numpy(d["genome"]).to(dev)
    d_r = torch.from_numpy(d["reads"]).to(dev)
    d_ro = torch.from_numpy(d["read_off"]).to(dev)
    d_a = torch.from_numpy(d["anchors"].view(np.uint8)).to(dev)
    d_ao = torch.from_numpy(d["anchor_off"]).to(dev)
    d_out = torch.zeros(nreads * tal_extend.READ_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    eng = tal_extend.Extender()
    stream = torch.cuda.current_stream(dev).cuda_stream

    def step():
        eng.run_device(d_g.data_ptr(), glen, d_r.data_ptr(), d_ro.data_ptr(), nreads,
                       d_a.data_ptr(), d_ao.data_ptr(), d_out.data_ptr(), stream=stream)

    # the parity check runs in the CPU leg, on its own batch (below)
    step()
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stages = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        stages.append(eng.stats())
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    st = stages[-1]
    mean = {k: float(np.mean([s[k] for s in stages])) for k in
            ("ms_chain", "ms_produce", "ms_left", "ms_right", "ms_total")}
    left, right, _, _ = eng.pairs()
    cells = float(tal_bsw.band_cells(left["len1"], left["len2"], 100, **PARAMS).sum() +
                  tal_bsw.band_cells(right["len1"], right["len2"], 100, **PARAMS).sum())
    ms_ext = mean["ms_left"] + mean["ms_right"]
    # the extension passes dominate; their roofline as the BSW line's
    achieved = 14 * cells / (ms_ext / 1e3) / 1e12
    line = {
        "metric": "reads/s of the end-to-end extension pipeline (chaining, extension pairs, "
                  "banded SW of both sides), bit-exact",
        "value": round(nreads / dt, 1), "unit": "reads/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int16/int8 (SW), int32 (chaining)",
        "data": "synthetic (reads simulated from a random genome: 1% substitutions, 0.1% indels; "
                "19-mer anchors plus 6 spurious a read)",
        "config": {"workload": "extend", "reads": nreads, "read_len": rlen, "genome": glen,
                   "anchors": na, "left_pairs": st["left_pairs"], "right_pairs": st["right_pairs"],
                   "w": 100, "pad": 100},
        "stages_ms": {k[3:]: round(v, 3) for k, v in mean.items()},
        "parity": NO_CHECK,
        "roofline": {"bound": "valu", "achieved": round(achieved, 3),
                     "peak": round(VALU_PEAK_TOPS, 2), "unit": "TOP/s",
                     "frac": round(achieved / VALU_PEAK_TOPS, 4), "traffic": None,
                     "kernel": "bsw_main (left + right passes)", "kernel_ms": round(ms_ext, 3),
                     "band_cells": cells, "ops_per_cell": 14},
    }
    if not args.no_cpu:
        # the CPU leg: the oracles check the first reads, run as their own
        # batch on a fresh handle (a batch's extension scores depend on its
        # group composition and on the handle's state, as the reference's
        # do, so a slice of the big batch is not compared), then the CPU
        # stages are timed
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import extend_model
        m = 2000
        sub = dict(genome=d["genome"], reads=d["reads"][:d["read_off"][m]],
                   read_off=d["read_off"][:m + 1], anchors=d["anchors"][:d["anchor_off"][m]],
                   anchor_off=d["anchor_off"][:m + 1])
        chk = tal_extend.Extender()
        got = chk.run(**sub)
        chk.close()
        want, _, _ = extend_model.pipeline(**sub)
        bad = sum(int((got[f] != want[f]).sum()) for f in tal_extend.READ_DTYPE.names)
        if bad:
            raise SystemExit(f"[extend] parity FAILED: {bad} fields differ on {m} reads")
        log(f"[extend] parity: a {m}-read batch bit-exact against the oracles")
        line["parity"] = f"bit-exact on a {m}-read batch (the first reads)"
        line["cpu_baseline"] = _extend_cpu(d, left, right, eng)
    print(json.dumps(line), flush=True)


def _extend_cpu(d, left, right, eng, min_s=10.0):
    """The CPU stages on a bounded sample of the reads, 16 threads: the
    reference's own mm_dp_vectorized (oracle/_ref build) over the reads'
    anchor sets, then the BSW oracle (a port of the reference's AVX-512
    path, which cannot be built here) on those reads' left and right pairs.
    The producer's own CPU time is left out."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from concurrent.futures import ThreadPoolExecutor
    nreads = len(d["read_off"]) - 1
    m = 20000
    off = d["anchor_off"][:m + 1]
    a = d["anchors"][:off[-1]]
    chunks = [(k, min(m, k + 500)) for k in range(0, m, 500)]
    ref_ok = O.ref_chain_available() and O.ref_chain_dp_sets(a[:off[1]], off[:2]) is not None

    def chain_part(c):
        s0, s1 = c
        sub, so = a[off[s0]:off[s1]], off[s0:s1 + 1] - off[s0]
        if ref_ok:
            O.ref_chain_dp_sets(sub, so)
        else:
            O.chain_dp_sets(sub, so)

    lsel = left[left["seqid"] < m]
    rsel = right[right["seqid"] < m]
    _, _, ref, qer = eng.pairs()
    passes, t = 0, time.perf_counter()
    while True:
        with ThreadPoolExecutor(16) as ex:
            list(ex.map(chain_part, chunks))
        O.run_mt(lsel.copy(), ref, qer, 100, bits=16, nthreads=16)
        O.run_mt(rsel.copy(), ref, qer, 100, bits=16, nthreads=16)
        passes += 1
        dt = time.perf_counter() - t
        if dt >= min_s:
            break
    return {"value": round(m * passes / dt, 1), "unit": "reads/s", "cores": 16, "kind": "port",
            "sample": f"{passes} passes over the first {m} of {nreads} reads, {dt:.1f}s: "
                      f"chaining by the reference's build ({'yes' if ref_ok else 'no: oracle'}), "
                      "the two extension passes by the BSW oracle (16-bit), producer not timed"}


def _traffic(name="bsw_main16_workload2"):
    """HBM bytes per launch of a kernel from its committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (profiles/<name>.json), if present."""
    f = os.path.join(ROOT, "profiles", name + ".json")
    try:
        with open(f) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return None


def _prepare(cfg, n, torch, dev):
    """Generates workload cfg and uploads it: one band class per w (workload
    3), each a batch on its own fresh handle."""
    t = time.time()
    pairs, ref, qer, wcls = synth.generate(cfg, n)
    ws = sorted(set(int(x) for x in wcls)) if cfg == 3 else [100]
    d_ref = torch.from_numpy(ref).to(dev)
    d_qer = torch.from_numpy(qer).to(dev)
    classes = []
    for w in ws:
        sel = np.where(wcls == w)[0] if cfg == 3 else np.arange(n)
        p = pairs[sel].copy()
        c = {"w": w, "pairs": p, "cells": tal_bsw.band_cells(p["len1"], p["len2"], w, **PARAMS),
             "rect": p["len1"].astype(np.int64) * p["len2"].astype(np.int64),
             "is8": tal_bsw.is_8bit(p, **PARAMS), "eng": tal_bsw.BandedPairWiseSW(**PARAMS)}
        c["d_pairs"] = torch.from_numpy(p.view(np.uint8)).to(dev)
        c["ptrs"] = (c["d_pairs"].data_ptr(), d_ref.data_ptr(), d_qer.data_ptr())
        classes.append(c)
    W = {"cfg": cfg, "n": n, "pairs": pairs, "ref": ref, "qer": qer, "d_ref": d_ref,
         "d_qer": d_qer, "classes": classes, "ws": ws}
    W["cells"] = np.concatenate([c["cells"] for c in classes])
    W["is8"] = np.concatenate([c["is8"] for c in classes])
    W["total_cells"] = float(W["cells"].sum())
    W["total_rect"] = float(sum(c["rect"].sum() for c in classes))
    log(f"[bench] workload {cfg}: {n} pairs in {time.time() - t:.1f}s; band cells "
        f"{W['total_cells']:.3e}; rect cells {W['total_rect']:.3e}; 8-bit share "
        f"{W['is8'].mean():.3f}; classes w={ws}")
    return W


def _release(W, torch):
    for c in W["classes"]:
        if "plan" in c:
            c["plan"].close()
        c["eng"].close()
    W.clear()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _tune(W, sample):
    """Autotunes each band class on its own pairs (a band class is its own
    getScores call, and the best waves per CU depend on the band width);
    returns the largest class's choice, with every class's under per_class."""
    per_class = {}
    big = max(W["classes"], key=lambda c: len(c["pairs"]))
    tuning = None
    for c in W["classes"]:
        t = time.time()
        tc = c["eng"].autotune(*c["ptrs"], len(c["pairs"]), c["w"], sample=sample)
        log(f"[bench] autotune w={c['w']} on {min(len(c['pairs']), sample)} pairs in "
            f"{time.time() - t:.1f}s: waves/CU {tc['waves_per_cu']}, order {tc['order']}; "
            + ", ".join(f"{x['waves_per_cu']}/{x['order']}:{x['ms']:.2f}" for x in tc["trials"]))
        per_class[str(c["w"])] = {"waves_per_cu": tc["waves_per_cu"], "order": tc["order"]}
        if c is big:
            tuning = tc
    if len(per_class) > 1:
        tuning = {**tuning, "per_class": per_class}
    return tuning


def _steps(W, steps, warmup, eager, stream, torch):
    """Builds the plans (unless eager), runs the first call on the fresh
    handles (kept for the parity check), the warmup and `steps` timed
    steps; returns (ms per step, first results per class)."""
    def eager_step(c):
        c["eng"].get_scores_device(0, *c["ptrs"], len(c["pairs"]), c["w"], stream.cuda_stream)

    if not eager:
        t = time.time()
        for c in W["classes"]:
            c["plan"] = c["eng"].plan(*c["ptrs"], len(c["pairs"]), c["w"])
        log(f"[bench] plans (hipGraph of the dispatcher pipeline) built in {time.time() - t:.2f}s")

    def step():
        for c in W["classes"]:
            if eager:
                eager_step(c)
            else:
                c["plan"].launch(stream.cuda_stream)

    step()
    torch.cuda.synchronize()
    firsts = [c["d_pairs"].cpu().numpy().view(tal_bsw.SEQPAIR_DTYPE).copy() for c in W["classes"]]
    for _ in range(max(0, warmup - 1)):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / max(steps, 1), firsts, eager_step


def _parity(W, firsts, per_class):
    bad, checked = 0, 0
    for c, got in zip(W["classes"], firsts):
        b, k = _check_prefix(c["pairs"], W["ref"], W["qer"], got, c["is8"], c["w"], per_class)
        bad += b
        checked += k
    return ("bit-exact" if bad == 0 else f"{bad} mismatching fields"), checked


def _naive_all(W, stream, torch, bucket):
    dt = 0.0
    for c in W["classes"]:
        dt += _naive(c["eng"], c["d_pairs"], W["d_ref"], W["d_qer"], len(c["pairs"]), c["w"],
                     stream, torch, bucket)
    return {"form": f"one dispatcher launch per {bucket}-pair bucket, device-resident, synchronous",
            "pairs": W["n"],
            "buckets": sum((len(c["pairs"]) + bucket - 1) // bucket for c in W["classes"]),
            "gcups": round(W["total_cells"] / dt / 1e9, 2), "ms": round(dt * 1e3, 1)}


def _workload_summary(cfg, args, torch, dev, stream):
    """One of the other batch workloads, measured the same way as the
    headline with fewer steps: plan GCUPS (band and rectangle cells), pairs/s,
    the naive form and a parity prefix against the oracle."""
    W = _prepare(cfg, WORKLOADS[cfg]["pairs"], torch, dev)
    tuning = _tune(W, min(args.tune_pairs, 131_072)) if args.tune_pairs > 0 else None
    ms, firsts, _ = _steps(W, 3, 1, False, stream, torch)
    out = {"workload": WORKLOADS[cfg]["name"], "pairs": W["n"], "band_classes": W["ws"],
           "ms_per_step": round(ms, 3),
           "gcups": round(W["total_cells"] / (ms / 1e3) / 1e9, 2),
           "gcups_rect": round(W["total_rect"] / (ms / 1e3) / 1e9, 2),
           "pairs_per_s": round(W["n"] / (ms / 1e3), 1),
           "pairs_8bit": int(W["is8"].sum()),
           "tuning": {k: tuning[k] for k in ("waves_per_cu", "order", "per_class") if k in tuning}
           if tuning else None}
    if not args.no_naive:
        naive = _naive_all(W, stream, torch, args.bucket)
        out["naive_gcups"] = naive["gcups"]
        out["batched_vs_naive"] = round(out["gcups"] / naive["gcups"], 2)
    if not args.no_cpu and args.check_pairs > 0:
        per = max(1, min(2000, args.check_pairs // 4) // len(W["classes"]))
        out["parity"], out["parity_pairs"] = _parity(W, firsts, per)
    else:
        out["parity"] = NO_CHECK
    log(f"[bench] workload {cfg}: {out}")
    _release(W, torch)
    return out


def _end_to_end(W, stream, torch, reps=3):
    """The whole batch from pinned host buffers: records and sequences in,
    the plan's replay, records out; every step moves all of them over PCIe.
    Measured two ways: one step after another, and pipelined (two device
    buffer sets with a plan each; step k+1's inputs go up and step k-1's
    outputs come down while step k's replay runs; the replays stay in order
    on one stream, so the handle's state evolves as in the serial form)."""
    cs = W["classes"]
    h_pairs = [torch.from_numpy(c["pairs"].view(np.uint8)).pin_memory() for c in cs]
    h_ref = torch.from_numpy(W["ref"]).pin_memory()
    h_qer = torch.from_numpy(W["qer"]).pin_memory()
    h_out = [[torch.empty(h.shape, dtype=h.dtype, pin_memory=True) for h in h_pairs]
             for _ in range(2)]

    # one explicit stream for the copies and the replays, so that each replay
    # runs after its inputs arrive and before its outputs leave (torch's
    # default stream has handle 0, which the library reads as "the handle's
    # own stream")
    s = torch.cuda.Stream()

    def once():
        with torch.cuda.stream(s):
            W["d_ref"].copy_(h_ref, non_blocking=True)
            W["d_qer"].copy_(h_qer, non_blocking=True)
            for c, hp, ho in zip(cs, h_pairs, h_out[0]):
                c["d_pairs"].copy_(hp, non_blocking=True)
                c["plan"].launch(s.cuda_stream)
                ho.copy_(c["d_pairs"], non_blocking=True)

    once()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        once()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    moved = sum(h.numel() for h in h_pairs) * 2 + h_ref.numel() + h_qer.numel()

    # pipelined: a second device buffer set with its own plans (a plan sizes
    # itself from its batch, so the set holds the batch first)
    d_ref2, d_qer2 = W["d_ref"].clone(), W["d_qer"].clone()
    d_pairs2 = [c["d_pairs"].clone() for c in cs]
    torch.cuda.synchronize()
    plans2 = [c["eng"].plan(dp.data_ptr(), d_ref2.data_ptr(), d_qer2.data_ptr(), len(c["pairs"]),
                            c["w"]) for c, dp in zip(cs, d_pairs2)]
    sets = [(W["d_ref"], W["d_qer"], [c["d_pairs"] for c in cs], [c["plan"] for c in cs]),
            (d_ref2, d_qer2, d_pairs2, plans2)]
    up, run, down = torch.cuda.Stream(), torch.cuda.Stream(), torch.cuda.Stream()
    freed = [None, None]  # a set's last download (its buffers may be refilled after it)
    arrived = [None, None]

    def upload(k):
        ref, qer, dps, _ = sets[k % 2]
        with torch.cuda.stream(up):
            if freed[k % 2] is not None:
                up.wait_event(freed[k % 2])
            ref.copy_(h_ref, non_blocking=True)
            qer.copy_(h_qer, non_blocking=True)
            for dp, hp in zip(dps, h_pairs):
                dp.copy_(hp, non_blocking=True)
            arrived[k % 2] = up.record_event()

    def replay_and_download(k):
        _, _, dps, pls = sets[k % 2]
        run.wait_event(arrived[k % 2])
        for pl in pls:  # (each launch waits for its replay: the next upload is queued first)
            pl.launch(run.cuda_stream)
        done = run.record_event()
        with torch.cuda.stream(down):
            down.wait_event(done)
            for ho, dp in zip(h_out[k % 2], dps):
                ho.copy_(dp, non_blocking=True)
            freed[k % 2] = down.record_event()

    def pipeline(n_steps):
        upload(0)
        for k in range(n_steps):
            if k + 1 < n_steps:
                upload(k + 1)
            replay_and_download(k)

    pipeline(2)
    torch.cuda.synchronize()
    n_steps = 2 * reps
    t0 = time.perf_counter()
    pipeline(n_steps)
    torch.cuda.synchronize()
    dtp = (time.perf_counter() - t0) / n_steps
    for pl in plans2:
        pl.close()
    return {"form": "pinned host records and sequences in, plan replay, records out, per step",
            "ms_per_step": round(dt * 1e3, 2), "gcups": round(W["total_cells"] / dt / 1e9, 2),
            "pairs_per_s": round(W["n"] / dt, 1), "pcie_bytes_per_step": int(moved),
            "pipelined": {"form": "the same steps, step k+1's inputs and step k-1's outputs moving "
                                  "while step k's replay runs (two device buffer sets)",
                          "ms_per_step": round(dtp * 1e3, 2),
                          "gcups": round(W["total_cells"] / dtp / 1e9, 2),
                          "pairs_per_s": round(W["n"] / dtp, 1)}}


def _batch(args):
    import torch

    cfg = args.workload
    n = args.pairs or WORKLOADS[cfg]["pairs"]
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    W = _prepare(cfg, n, torch, dev)
    classes, ref, qer = W["classes"], W["ref"], W["qer"]
    cells, is8, total_cells = W["cells"], W["is8"], W["total_cells"]
    tuning = _tune(W, args.tune_pairs) if args.tune_pairs > 0 else None

    ms_step, firsts, eager_step = _steps(W, args.steps, args.warmup, args.eager, stream, torch)
    check = NO_CHECK
    gcups = total_cells / (ms_step / 1e3) / 1e9
    log(f"[bench] {ms_step:.1f} ms/step, {gcups:.1f} GCUPS ({'eager' if args.eager else 'plan'})")

    # the same steps through the eager dispatcher: its HIP events, recorded on
    # the kernels' stream, give the per-kernel times
    ms16, ms8 = [], []
    ke = max(1, args.steps // 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(ke):
        a16 = a8 = 0.0
        for c in classes:
            eager_step(c)
            st = c["eng"].stats()
            a16 += st["ms_main16"]
            a8 += st["ms_main8"]
        ms16.append(a16)
        ms8.append(a8)
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1e3 / ke
    log(f"[bench] eager dispatcher: {eager_ms:.1f} ms/step; main16 {np.mean(ms16):.1f} ms, "
        f"main8 {np.mean(ms8):.1f} ms")

    cells16 = float(cells[~is8].sum())
    k16_ms = float(np.mean(ms16))
    achieved = OPS_PER_CELL * cells16 / (k16_ms / 1e3) / 1e12 if k16_ms > 0 else 0.0
    tr = _traffic() if cfg == 2 and n == WORKLOADS[2]["pairs"] else None
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(VALU_PEAK_TOPS, 2),
                "unit": "TOP/s", "frac": round(achieved / VALU_PEAK_TOPS, 4),
                "traffic": round(tr["traffic_bytes"]) if tr else None,
                "traffic_source": "profiles/bsw_main16_workload2.md (rocprofv3 FETCH_SIZE x2.0 + "
                                  "WRITE_SIZE, per main launch)" if tr else None,
                "traffic_rate_tb_s": round(tr["traffic_bytes"] / (k16_ms / 1e3) / 1e12, 2)
                if tr and k16_ms > 0 else None,
                "traffic_bytes_per_band_cell": round(tr["traffic_bytes"] / cells16, 2)
                if tr else None,
                "kernel": "bsw_main<16> (2 groups of 32 pairs per wave)",
                "kernel_ms": round(k16_ms, 3),
                "kernel_gcups": round(cells16 / (k16_ms / 1e3) / 1e9, 1) if k16_ms > 0 else None,
                "ops_per_cell": OPS_PER_CELL}

    naive = None
    if not args.no_naive:
        naive = _naive_all(W, stream, torch, args.bucket)
        log(f"[bench] naive one launch per bucket: {naive}")

    plain = None
    if not args.no_plain:
        c = classes[0]
        plain = _plain(c["pairs"], ref, qer, c["cells"], c["w"], args.plain_pairs)
        log(f"[bench] plain per-chunk host calls: {plain}")

    e2e = None
    if not args.no_e2e and not args.eager:
        e2e = _end_to_end(W, stream, torch)
        log(f"[bench] end to end from host buffers: {e2e}")

    cpu = None
    if not args.no_cpu:
        # the CPU leg: the oracle checks the first calls' results, then is
        # timed as the CPU baseline
        if args.check_pairs > 0:
            per = max(1, args.check_pairs // len(classes))
            check, checked = _parity(W, firsts, per)
            log(f"[bench] parity vs oracle on {checked} pairs (dispatcher prefixes): {check}")
        c = classes[0]
        cpu = _cpu_baseline(c["pairs"], ref, qer, c["cells"], c["is8"], c["w"], args.cpu_pairs)
        log(f"[bench] cpu baseline: {cpu}")

    line = {
        "metric": "GCUPS (band cells/s) of bit-exact BSW, batched persistent dispatch; pairs/s",
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16+int8",
        "data": f"synthetic (seeded generator, workload {cfg})",
        "config": {"workload": WORKLOADS[cfg]["name"], "pairs": n, "band_classes": W["ws"],
                   **WORKLOADS[cfg]["shape"], "pairs_8bit": int(is8.sum()),
                   "band_cells": total_cells, "rect_cells": W["total_rect"]},
        "gcups_rect": round(W["total_rect"] / (ms_step / 1e3) / 1e9, 2),
        "pairs_per_s": round(n / (ms_step / 1e3), 1),
        "parity": check,
        "pipeline": "eager dispatcher" if args.eager else "hipGraph plan replay",
        "eager_ms_per_step": round(eager_ms, 3),
        "tuning": {k: tuning[k] for k in ("waves_per_cu", "order", "ms", "per_class") if k in tuning}
        if tuning else None,
        "naive_gpu": naive,
        "plain_gpu": plain,
        "end_to_end": e2e,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if naive and naive.get("gcups"):
        line["batched_vs_naive"] = round(gcups / naive["gcups"], 2)
    if plain and plain.get("gcups"):
        line["batched_vs_plain"] = round(gcups / plain["gcups"], 2)
    _release(W, torch)
    # the rest of the metric: the other batch workloads and streamed serving
    if not args.no_others and cfg == 2 and not args.pairs:
        line["workloads"] = {str(k): _workload_summary(k, args, torch, dev, stream)
                             for k in (1, 3, 5)}
    if not args.no_serve and cfg == 2 and not args.pairs:
        line["serving"] = _serve_levels(args, resident=True)
        line["serving_host_batched"] = _serve_levels(args, resident=False)
    print(json.dumps(line), flush=True)


def _slice_request(pairs, ref, qer, a, b):
    """Pairs [a, b) of a generated pool as a self-contained request: its own
    pair records, and views of the pool's sequence bytes (the generator lays
    them out in pair order) with the offsets rebased."""
    p = pairs[a:b].copy()
    r0, r1 = int(p["idr"][0]), int(p["idr"][-1] + p["len1"][-1])
    q0, q1 = int(p["idq"][0]), int(p["idq"][-1] + p["len2"][-1])
    p["idr"] -= r0
    p["idq"] -= q0
    return p, ref[r0:r1], qer[q0:q1]


def _serve_levels(args, resident=True):
    """Streamed serving: requests of log-uniform 1k-64k pairs drawn from the
    workload-1 and workload-2 generators arrive as a Poisson process at 50%,
    80% and 95% of the peak pairs/s measured by submitting a full set at
    once; latency is submit -> outputs in the caller's array.  resident: the
    device-resident loop (one persistent kernel), else host-launched batches
    on two workers.  Returns the per-level p50/p99 and the server's
    counters."""
    import threading
    rng = np.random.default_rng(0xB5A0000000000004)
    pools = {1: synth.generate(1, 400_000), 2: synth.generate(2, 1_000_000)}
    nreq = args.serve_requests
    sizes = np.exp(rng.uniform(np.log(1000), np.log(64000), 3 * nreq + nreq)).astype(int)
    kinds = rng.integers(1, 3, len(sizes))
    reqs = []
    for sz, kd in zip(sizes, kinds):
        pairs, ref, qer, _ = pools[int(kd)]
        a = int(rng.integers(0, len(pairs) - sz))
        reqs.append(_slice_request(pairs, ref, qer, a, a + int(sz)))
    mode = "resident" if resident else "host-batched"
    log(f"[serve] {mode}: {len(reqs)} requests, mean {sizes.mean():.0f} pairs")
    srv = tal_bsw.Server(hbm_budget=1 << 30, workers=2, resident=resident, max_query_len=256,
                         **PARAMS)

    def run_set(sub, rate):
        """Submits sub (open loop at `rate` requests/s, or all at once if None)
        and returns the per-request records and the wall time."""
        live = [r[0].copy() for r in sub]
        tickets = [None] * len(sub)
        infos = [None] * len(sub)

        def waiter():
            for k in range(len(sub)):
                while tickets[k] is None:
                    time.sleep(0.0002)
                infos[k] = srv.wait(tickets[k])

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        nxt = t0
        for k, r in enumerate(sub):
            if rate:
                nxt += rng.exponential(1.0 / rate)
                d = nxt - time.perf_counter()
                if d > 0:
                    time.sleep(d)
            tickets[k] = srv.submit(live[k], r[1], r[2], 100)
        th.join()
        return infos, time.perf_counter() - t0

    warm = reqs[:20]
    run_set(warm, None)
    peak_set = reqs[:nreq]
    infos, dt = run_set(peak_set, None)
    npairs = sum(len(r[0]) for r in peak_set)
    peak = npairs / dt
    log(f"[serve] peak {peak:.3e} pairs/s over {len(peak_set)} requests ({dt:.2f} s)")
    levels = []
    for j, load in enumerate((0.5, 0.8, 0.95)):
        sub = reqs[nreq * (j + 1):nreq * (j + 2)]
        mean_pairs = np.mean([len(r[0]) for r in sub])
        rate = load * peak / mean_pairs
        infos, dt = run_set(sub, rate)
        lat = np.array([i["latency_ms"] for i in infos])
        bad = sum(i["status"] != 0 for i in infos)
        lv = {"load": load, "requests": len(sub), "offered_req_per_s": round(rate, 2),
              "achieved_pairs_per_s": round(sum(len(r[0]) for r in sub) / dt, 1),
              "p50_ms": round(float(np.percentile(lat, 50)), 3),
              "p99_ms": round(float(np.percentile(lat, 99)), 3),
              "mean_batch_requests": round(float(np.mean([i["batch_requests"] for i in infos])), 2),
              "failed": int(bad)}
        log(f"[serve] {lv}")
        levels.append(lv)
    st = srv.stats()
    srv.close()
    cfg = {"mode": mode, "request_pairs": "log-uniform 1k-64k", "arrivals": "Poisson",
           "hbm_budget": 1 << 30, "w": 100}
    cfg.update({"max_query_len": 256} if resident else {"workers": 2})
    return {"metric": "p50/p99 request latency, streamed serving (continuous batching)",
            "p99_ms_at_95pct": levels[-1]["p99_ms"], "peak_pairs_per_s": round(peak, 1),
            "levels": levels, "data": "synthetic (workloads 1 and 2 generators)",
            "config": cfg, "server": st}


def _serve(args):
    r = _serve_levels(args, resident=args.serve_mode != "host")
    line = {"metric": r["metric"], "value": r["p99_ms_at_95pct"], "unit": "ms (p99 at 95% load)",
            "n_gpus": 1, "higher_is_better": False, "peak_pairs_per_s": r["peak_pairs_per_s"],
            "levels": r["levels"], "data": r["data"],
            "config": {"workload": "streamed serving", **r["config"]}, "server": r["server"]}
    print(json.dumps(line), flush=True)


def _copy_calibration(mib):
    """A plain copy kernel of known size: it reads and writes exactly
    mib MiB, so the counters' reading for it gives their scale."""
    import torch
    n = mib * 2**20 // 4
    src = torch.arange(n, dtype=torch.int32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "ok": bool((dst[-1] == n - 1).item())}), flush=True)


def _check_prefix(pairs, ref, qer, got, is8, w, m):
    """Both dispatcher halves' first m pairs against the oracle (fresh
    objects); returns (mismatching fields, pairs checked)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    o = oracle.Oracle(**PARAMS)
    bad, k = 0, 0
    for sel, bits in ((np.where(is8)[0][:m], 8), (np.where(~is8)[0][:m], 16)):
        sub = pairs[sel].copy()
        want = o.scores8(sub, ref, qer, w) if bits == 8 else o.scores16(sub, ref, qer, w)
        for f in tal_bsw.OUT_FIELDS:
            bad += int((got[f][sel] != want[f]).sum())
        k += len(sel)
    o.close()
    return bad, k


def _naive(eng, d_pairs, d_ref, d_qer, n, w, stream, torch, bucket=20_000):
    """The naive GPU form: the same inputs resident in HBM, one dispatcher
    launch per bucket of `bucket` consecutive pairs (bench-bsw's 20,000-pair
    chunks; consecutive, so that the groups stay the reference's), waiting for
    each bucket before the next.  Returns seconds for the whole batch."""
    rec = 56
    base = d_pairs.data_ptr()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for b in range(0, n, bucket):
        m = min(bucket, n - b)
        eng.get_scores_device(0, base + b * rec, d_ref.data_ptr(), d_qer.data_ptr(), m, w,
                              stream.cuda_stream)
        stream.synchronize()
    return time.perf_counter() - t


def _plain(pairs, ref, qer, cells, w, m):
    """bench-bsw's pattern: one blocking getScores16 host call per 20,000
    pairs (copies in and out each call)."""
    eng = tal_bsw.BandedPairWiseSW(**PARAMS)
    sub = pairs[:m].copy()
    eng.get_scores16(sub[:20000].copy(), ref, qer, w)  # warm
    t = time.perf_counter()
    for b in range(0, m, 20000):
        chunk = sub[b:b + 20000]
        eng.get_scores16(chunk, ref, qer, w)
    dt = time.perf_counter() - t
    eng.close()
    return {"form": "getScores16 per 20000-pair chunk, host buffers", "pairs": m,
            "gcups": round(float(cells[:m].sum()) / dt / 1e9, 2), "pairs_per_s": round(m / dt, 1)}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_baseline(pairs, ref, qer, cells, is8, w, m):
    """The CPU baseline: the vectorised restatement of the reference's
    AVX-512BW group loop (oracle/bsw_simd.cpp, bit-exact with the oracle),
    on the host cores, the way bench-bsw drives the reference: one object per
    thread, 20,000-pair chunks, 8/16-bit split as the dispatcher's.  m = 0
    times the whole batch."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    threads = max(1, min(16, avail))
    m = m or len(pairs)
    sub, s8 = pairs[:m].copy(), is8[:m]
    t = time.perf_counter()
    oracle.simd_run_mt(sub[s8].copy(), ref, qer, w, bits=8, nthreads=threads)
    oracle.simd_run_mt(sub[~s8].copy(), ref, qer, w, bits=16, nthreads=threads)
    dt = time.perf_counter() - t
    g = float(cells[:m].sum()) / dt / 1e9
    width = oracle.simd_width()
    return {"value": round(g, 3), "unit": "GCUPS", "cores": threads,
            "gcups_per_core": round(g / threads, 3), "kind": "port",
            "cpu": _cpu_model(),
            "sample": f"{'all' if m == len(pairs) else 'first'} {m} pairs of the workload, "
                      f"dispatcher split, 20000-pair chunks per thread-owned object, {dt:.1f}s",
            "code": f"vectorised restatement of smithWaterman512_16/_8 (oracle/bsw_simd.cpp, "
                    f"{'AVX-512BW' if width == 512 else 'AVX2'} build), bit-exact with the oracle"}


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark of the banded Smith-Waterman extension engine on one MI355X.

Workload: 2,000,000 extension pairs, query length U[30,250], reference =
query + U[0,100] flank, 2% mismatches and 0.5% indels, h0 U[0,120], w = 100,
scoring of bench-bsw (match 1, mismatch 4, gap open 6 / extend 1 both ways,
z-drop 100, end bonus 5).  A step is one pass of the 8/16-bit dispatcher over
the whole batch with the pairs and sequences already resident in HBM
(tal_bsw_get_scores_auto_device); outputs are written in place.

value = band cells / second (GCUPS) of the whole step, where a pair's band
cells are sum_i max(0, min(len2, i+b+1) - max(0, i-b)) with b its band
(bandedSWA.cpp:2882-2904).  Also reported: pairs/s, the plain form (per
20,000-pair host calls of getScores16, as bench-bsw drives the reference), the
roofline of the dominant kernel against integer VALU peak, and the oracle's
CPU rate on a bounded sample.
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=2_000_000)
    ap.add_argument("--plain-pairs", type=int, default=200_000)
    ap.add_argument("--cpu-pairs", type=int, default=60_000)
    ap.add_argument("--check-pairs", type=int, default=20_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-naive", action="store_true")
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
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (for scaling the "
                         "FETCH_SIZE/WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        _copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU: the engine has no multi-GPU path in this build "
              "(pairs are independent; see DESIGN.md for how it would shard).", file=sys.stderr)
        sys.exit(2)
    if args.serve:
        _serve(args)
        return

    import torch

    w = 100
    t = time.time()
    pairs, ref, qer, _ = synth.generate(2, args.pairs)
    cells = tal_bsw.band_cells(pairs["len1"], pairs["len2"], w, **PARAMS)
    is8 = tal_bsw.is_8bit(pairs, **PARAMS)
    log(f"[bench] generated {args.pairs} pairs in {time.time() - t:.1f}s; "
        f"band cells {cells.sum():.3e}; 8-bit share {is8.mean():.3f}")

    dev = torch.device("cuda:0")
    d_pairs = torch.from_numpy(pairs.view(np.uint8)).to(dev)
    d_ref = torch.from_numpy(ref).to(dev)
    d_qer = torch.from_numpy(qer).to(dev)
    stream = torch.cuda.current_stream(dev)
    eng = tal_bsw.BandedPairWiseSW(**PARAMS)
    n = len(pairs)
    ptrs = (d_pairs.data_ptr(), d_ref.data_ptr(), d_qer.data_ptr())
    tuning = None
    if args.tune_pairs > 0:
        t = time.time()
        tuning = eng.autotune(*ptrs, n, w, sample=args.tune_pairs)
        log(f"[bench] autotune on {min(n, args.tune_pairs)} pairs in {time.time() - t:.1f}s: "
            f"waves/CU {tuning['waves_per_cu']}, order {tuning['order']}; "
            + ", ".join(f"{x['waves_per_cu']}/{x['order']}:{x['ms']:.2f}" for x in tuning["trials"]))

    def eager_step():
        eng.get_scores_device(0, *ptrs, n, w, stream.cuda_stream)

    if args.eager:
        step = eager_step
    else:
        t = time.time()
        plan = eng.plan(*ptrs, n, w)
        log(f"[bench] plan (hipGraph of the dispatcher pipeline) built in {time.time() - t:.2f}s")

        def step():
            plan.launch(stream.cuda_stream)

    # first call on a fresh handle doubles as the parity check of a prefix
    step()
    torch.cuda.synchronize()
    got = d_pairs.cpu().numpy().view(tal_bsw.SEQPAIR_DTYPE)
    check = "skipped"
    if args.check_pairs > 0:
        check = _check_prefix(pairs, ref, qer, got, is8, w, args.check_pairs)
        log(f"[bench] parity vs oracle on dispatcher prefixes: {check}")
    for _ in range(max(0, args.warmup - 1)):
        step()
    torch.cuda.synchronize()

    ms16, ms8 = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        st = eng.stats()
        ms16.append(st["ms_main16"])
        ms8.append(st["ms_main8"])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / max(args.steps, 1)
    total_cells = float(cells.sum())
    gcups = total_cells / (ms_step / 1e3) / 1e9
    st = eng.stats()
    log(f"[bench] {ms_step:.1f} ms/step, {gcups:.1f} GCUPS; main16 {np.mean(ms16):.1f} ms, "
        f"main8 {np.mean(ms8):.1f} ms, stats {st}")

    eager_ms = None
    if not args.eager:
        # the same steps through the eager dispatcher, whose per-kernel HIP
        # events (recorded on the kernels' stream) give the 16-bit kernel time
        ms16, ms8 = [], []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(max(1, args.steps // 2)):
            eager_step()
            st = eng.stats()
            ms16.append(st["ms_main16"])
            ms8.append(st["ms_main8"])
        torch.cuda.synchronize()
        eager_ms = (time.perf_counter() - t0) * 1e3 / max(1, args.steps // 2)
        log(f"[bench] eager dispatcher (no graph): {eager_ms:.1f} ms/step; main16 "
            f"{np.mean(ms16):.1f} ms, main8 {np.mean(ms8):.1f} ms")

    cells16 = float(cells[~is8].sum())
    k16_ms = float(np.mean(ms16))
    achieved = OPS_PER_CELL * cells16 / (k16_ms / 1e3) / 1e12 if k16_ms > 0 else 0.0
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(VALU_PEAK_TOPS, 2),
                "unit": "TOP/s", "frac": round(achieved / VALU_PEAK_TOPS, 4), "traffic": None,
                "kernel": "bsw_main<16> (2 groups of 32 pairs per wave)",
                "kernel_ms": round(k16_ms, 3), "kernel_gcups": round(cells16 / (k16_ms / 1e3) / 1e9, 1)
                if k16_ms > 0 else None}

    naive = None
    if not args.no_naive:
        naive = _naive(eng, d_pairs, d_ref, d_qer, cells, n, w, stream, torch, args.bucket)
        log(f"[bench] naive one launch per bucket: {naive}")

    plain = None
    if not args.no_plain:
        plain = _plain(pairs, ref, qer, cells, w, args.plain_pairs)
        log(f"[bench] plain per-chunk host calls: {plain}")

    cpu = None
    if not args.no_cpu:
        cpu = _cpu_baseline(pairs, ref, qer, cells, is8, w, args.cpu_pairs)
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
        "data": "synthetic (seeded generator, workload 2)",
        "config": {"workload": "uneven batch, 8/16-bit dispatch", "pairs": n, "w": w,
                   "query_len": "U[30,250]", "ref_flank": "U[0,100]", "h0": "U[0,120]",
                   "pairs_8bit": int(is8.sum()), "band_cells": total_cells},
        "pairs_per_s": round(n / (ms_step / 1e3), 1),
        "parity": check,
        "pipeline": "eager dispatcher" if args.eager else "hipGraph plan replay",
        "eager_ms_per_step": round(eager_ms, 3) if eager_ms else None,
        "tuning": {k: tuning[k] for k in ("waves_per_cu", "order", "ms")} if tuning else None,
        "naive_gpu": naive,
        "plain_gpu": plain,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if naive and naive.get("gcups"):
        line["batched_vs_naive"] = round(gcups / naive["gcups"], 2)
    if plain and plain.get("gcups"):
        line["batched_vs_plain"] = round(gcups / plain["gcups"], 2)
    print(json.dumps(line), flush=True)
    eng.close()


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


def _serve(args):
    """Streamed serving: requests of log-uniform 1k-64k pairs drawn from the
    workload-1 and workload-2 generators arrive as a Poisson process at 50%,
    80% and 95% of the peak pairs/s measured by submitting a full set at
    once; latency is submit -> outputs in the caller's array."""
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
    log(f"[serve] {len(reqs)} requests, mean {sizes.mean():.0f} pairs")
    srv = tal_bsw.Server(hbm_budget=1 << 30, workers=2, **PARAMS)

    def run_set(sub, rate):
        """Submits sub (open loop at `rate` requests/s, or all at once if None)
        and returns the per-request records and the wall time."""
        live = [r[0].copy() for r in sub]
        tickets = [None] * len(sub)
        done = threading.Event()
        infos = [None] * len(sub)

        def waiter():
            for k in range(len(sub)):
                while tickets[k] is None:
                    time.sleep(0.0002)
                infos[k] = srv.wait(tickets[k])
            done.set()

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
    line = {"metric": "p50/p99 request latency, streamed serving (continuous batching)",
            "value": levels[-1]["p99_ms"], "unit": "ms (p99 at 95% load)", "n_gpus": 1,
            "higher_is_better": False, "peak_pairs_per_s": round(peak, 1),
            "levels": levels, "data": "synthetic (workloads 1 and 2 generators)",
            "config": {"workload": "streamed serving", "request_pairs": "log-uniform 1k-64k",
                       "arrivals": "Poisson", "hbm_budget": 1 << 30, "workers": 2, "w": 100},
            "server": st}
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
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    o = oracle.Oracle(**PARAMS)
    bad = 0
    for sel, bits in ((np.where(is8)[0][:m], 8), (np.where(~is8)[0][:m], 16)):
        sub = pairs[sel].copy()
        want = o.scores8(sub, ref, qer, w) if bits == 8 else o.scores16(sub, ref, qer, w)
        for f in tal_bsw.OUT_FIELDS:
            bad += int((got[f][sel] != want[f]).sum())
    return "bit-exact" if bad == 0 else f"{bad} mismatching fields"


def _naive(eng, d_pairs, d_ref, d_qer, cells, n, w, stream, torch, bucket=20_000):
    """The naive GPU form: the same inputs resident in HBM, one dispatcher
    launch per bucket of `bucket` consecutive pairs (bench-bsw's 20,000-pair
    chunks; consecutive, so that the groups stay the reference's), waiting for
    each bucket before the next, over the whole batch."""
    rec = 56
    base = d_pairs.data_ptr()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for b in range(0, n, bucket):
        m = min(bucket, n - b)
        eng.get_scores_device(0, base + b * rec, d_ref.data_ptr(), d_qer.data_ptr(), m, w,
                              stream.cuda_stream)
        stream.synchronize()
    dt = time.perf_counter() - t
    return {"form": f"one dispatcher launch per {bucket}-pair bucket, device-resident, "
                    f"synchronous", "pairs": n, "buckets": (n + bucket - 1) // bucket,
            "gcups": round(float(cells.sum()) / dt / 1e9, 2), "ms": round(dt * 1e3, 1)}


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


def _cpu_baseline(pairs, ref, qer, cells, is8, w, m):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = min(16, os.cpu_count() or 1)
    sub, s8 = pairs[:m].copy(), is8[:m]
    t = time.perf_counter()
    oracle.run_mt(sub[s8].copy(), ref, qer, w, bits=8, nthreads=threads)
    oracle.run_mt(sub[~s8].copy(), ref, qer, w, bits=16, nthreads=threads)
    dt = time.perf_counter() - t
    return {"value": round(float(cells[:m].sum()) / dt / 1e9, 3), "unit": "GCUPS",
            "cores": threads, "kind": "port",
            "sample": f"first {m} pairs of the workload, dispatcher split, 20000-pair chunks "
                      f"per thread-owned oracle object, {dt:.1f}s"}


if __name__ == "__main__":
    main()

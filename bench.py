#!/usr/bin/env python3
"""Seed-and-filter benchmark: query Mbp/s against a 1 Gbp resident target,
batched path vs the plain per-chunk path.

Workload: a 1,000,000,000 bp uppercase target (iid, 41% GC, seed 0xC0FFEE05)
indexed with 12of19 and held in HBM. The query is a set of chunks sampled from
it (clipped lognormal lengths, median 200 kbp, 10 kbp..2 Mbp, divergence
U(5%, 25%) as substitutions plus indels at a fifteenth of that rate, half
reverse-complemented), each chunk its own query block. HOXD70, X-drop 910,
threshold 3000, both strands, entropy on, transitions on; max_hits is the value
the reference derives from this card's memory (4194304 x GiB), reported.

A step = one seed-and-filter pass over every chunk: raw query blocks in (the
bytes a FASTA record packs to), ordered HSP lists out; the reverse complement
and the encoding of both strands run on the device inside the step. Batched:
one hspz_filter_batch call over all chunks (`value`). Plain: per chunk,
hspz_swap_query_raw + the reference-shaped per-batch pipeline (hspz_run_pass),
reported beside it. Both produce the same lists (checked).
Target generation and the device build of the resident index are setup, outside
the timed region (the index build time is reported).

Before the timed region the batched gate's autotuner (hspz_autotune_batch)
measures its launch grid on the leading ~1 Mbp of chunks and keeps the fastest
cell; every measured cell is reported under `autotune`.

Prints one JSON line (N = 1). `roofline` covers the fused expansion + score
gate kernel of the batched path (HIP events around its launches);
`cpu_baseline` is the C oracle on a bounded sample of the same chunks, one
block per worker thread (all-core), with a 1-thread figure beside it. `io`
times the host paths around the engine outside the timed region: writing the
step's lists as .segments (directory, -D partitions, .tar.gz) and reading a
FASTA / gzipped FASTA slice of the target.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Stored bytes per base the packed gate reads per sequence: a 2-bit code plus
# one exception bit per 64 bases.
W_PACKED = 0.25 + 1.0 / 512


def gate_profile(order_used, coarse_used):
    """The committed rocprofv3 summary of this gate path on the bench workload."""
    if not order_used:
        return "profiles/seed_gate_1gbp_target.md"
    return "profiles/coarse_gate_1gbp_target.md" if coarse_used else "profiles/kmer_gate_1gbp_target.md"


def profiled_traffic(hits_per_launch, order_used, coarse_used):
    """HBM bytes per launch of the batched gate (its kernels for every hit and
    the deferred-hit gates), FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 from the
    committed rocprofv3 counter summary of this same workload and path; None
    when this run's launch is not that one (another hit count)."""
    path = os.path.join(ROOT, gate_profile(order_used, coarse_used))
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic-line: \S+ hits=(\d+) .*bytes_per_launch=(\d+)", open(path).read())
    if not m or int(m.group(1)) != hits_per_launch:
        return None
    return int(m.group(2))


def profiled_valu(order_used, coarse_used):
    """(VALU wave-instructions per hit, VALU issue fraction) of the coarse
    bound from the committed profile of this path, or None."""
    path = os.path.join(ROOT, gate_profile(order_used, coarse_used))
    if not os.path.exists(path):
        return None
    m = re.search(r"valu-line: \S+ valu_per_hit=([0-9.]+) valu_issue=([0-9.]+)", open(path).read())
    return (float(m.group(1)), float(m.group(2))) if m else None


def set_flags(eng, pairs):
    """Engine flags from NAME=VALUE strings."""
    for kv in pairs:
        name, _, v = kv.partition("=")
        eng.set_flag(name.strip(), int(v))
        log("flag %s = %d" % (name.strip(), int(v)))


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--target-mbp", type=float, default=1000.0)
    ap.add_argument("--query-mbp", type=float, default=4.0)
    ap.add_argument("--cpu-blocks", type=int, default=32,
                    help="20 kbp blocks of the query the CPU baseline runs, one block per worker thread")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--plain-only", action="store_true",
                    help="time only the plain per-chunk path (profiling)")
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="requests per arrival-rate arm of the serving measurement (0 = skip)")
    ap.add_argument("--serve-rates", default="50,200,800")
    ap.add_argument("--serve-check", type=int, default=50,
                    help="requests per arm, drawn at random, whose served lists are compared with the "
                         "same request processed alone")
    ap.add_argument("--tune-mbp", type=float, default=1.0,
                    help="query Mbp the gate autotuner measures on before the timed region (0 = skip)")
    ap.add_argument("--tune-reps", type=int, default=1)
    ap.add_argument("--no-io", action="store_true", help="skip the host input/output timing")
    ap.add_argument("--gate", default="",
                    help="step_cells,block_threads,claim_seeds,grid_blocks[,order[,lookahead[,coarse[,cand_planes]]]] "
                         "for the batched gate "
                         "(applied after the autotuner; use with --tune-mbp 0 to fix the configuration)")
    ap.add_argument("--cand-bins", type=int, choices=[0, 1, 2], default=None,
                    help="binning of the coarse bound's candidates by reference position: 0 never, "
                         "1 passes of at most 12 Mi query bases (the engine's default), 2 every pass")
    ap.add_argument("--flag", action="append", default=[], metavar="NAME=VALUE",
                    help="set an engine flag (hspz_amd.engine.FLAG names, e.g. batch_max_positions=11500000) "
                         "before the timed region; repeatable")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device once and exit: the plain copy of known size "
                         "that calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE readings")
    ap.add_argument("--workload", default="bench", choices=["bench", "c4"],
                    help="bench (default): the 1 Gbp line above. c4: the mammal-scale shape on one GPU "
                         "(3 Gbp resident target, 50 Mbp query chunks), one JSON line of its own")
    ap.add_argument("--c4-chunks", type=int, default=1, help="50 Mbp query chunks of the c4 workload")
    args = ap.parse_args()

    if args.workload == "c4":
        return c4_workload(args)

    if args.copy_probe:
        import torch
        a = torch.ones(args.copy_probe << 20, dtype=torch.uint8, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("copied %d bytes" % (args.copy_probe << 20))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the seed-and-filter path has no cross-GPU step "
              "(work units shard by reference/query bin with a host-side ordered merge). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from hspz_amd import engine as E
    from hspz_amd import sequence as S
    from hspz_amd import synth

    if E.device_count() < 1:
        raise SystemExit("no HIP device visible")

    t0 = time.time()
    tlen = int(args.target_mbp * 1e6)
    target = synth.random_bases(0xC0FFEE05, tlen)
    traw = target.tobytes()
    log("target %d bp generated in %.1fs" % (tlen, time.time() - t0))
    from hspz_amd import seqio as X
    ref_enc = X.encode(traw)
    sub = E.build_sub_mat("", 910)

    qwant = int(args.query_mbp * 1e6)
    pieces, total, seed = [], 0, 0xC0FFEE03
    while total < qwant:
        more = synth.sample_chunks(target, seed, 16, 200_000, 0.768, 10_000, 2_000_000, 0.05, 0.25)
        seed += 1
        for c in more:
            if total >= qwant:
                break
            pieces.append(c)
            total += len(c)
    # raw query blocks (one record each); the device encodes both strands
    raws = [S.pack([("q%d" % i, c)])[0][:len(c)] for i, c in enumerate(pieces)]
    chunks = [(r, None) for r in raws]
    query_bp = sum(len(c) for c in pieces)
    log("query: %d chunks, %d bp" % (len(pieces), query_bp))

    _, total_mem = _device_memory(E)
    max_hits = E.default_max_hits(total_mem)
    t1 = time.time()
    # the 12of19 index of the target is built on the device and stays resident
    eng = E.Engine(None, None, ref_enc, sub, seed_size=19, xdrop=910, hspthresh=3000,
                   noentropy=False, max_hits=max_hits, device_index="12of19")
    index_ms = eng.stat("index_build_ns") / 1e6
    log("engine up in %.1fs: device index %.0f ms (%d positions), max_hits %d"
        % (time.time() - t1, index_ms, eng.stat("pos_len"), max_hits))
    index = pos = None
    if not args.no_cpu:
        index, pos = eng.index()  # for the oracle's sample

    def batched():
        return eng.filter_batch(chunks)

    if args.plain_only:
        t_p = time.perf_counter()
        for r in raws:
            eng.swap_query_raw(r)
            eng.run_pass()
        log("plain per-chunk only: %.3f Mbp/s" % (query_bp / 1e6 / (time.perf_counter() - t_p)))
        eng.close()
        return

    tune = None
    if args.tune_mbp > 0:
        # autotuner on the leading chunks (untimed): every cell is reported
        sample, sbp = [], 0
        for r in raws:
            if sbp >= args.tune_mbp * 1e6:
                break
            sample.append((r, None))
            sbp += len(r)
        t1 = time.time()
        cells, best = eng.autotune(sample, reps=args.tune_reps)
        for c in cells:
            log("tune order %d step %2d threads %4d grid %5d claim %3d: gate %.2f ms + after %.2f ms "
                "(%d hits, %d survivors, %d deferred)"
                % (c["order"], c["step_cells"], c["block_threads"], c["grid_blocks"], c["claim_seeds"],
                   c["gate_ms"], c["after_ms"], c["hits"], c["survivors"], c["deferred"]))
        log("autotuner: %d cells on %d bp in %.1fs; chose order %d step %d threads %d grid %d claim %d"
            % (len(cells), sbp, time.time() - t1, best["order"], best["step_cells"],
               best["block_threads"], best["grid_blocks"], best["claim_seeds"]))
        cols = ["order", "step_cells", "block_threads", "grid_blocks", "claim_seeds"]
        tune = {"sample_bp": sbp, "reps": args.tune_reps,
                "order_names": ["seed-major", "k-mer-major"],
                "columns": cols + ["gate_ms", "after_ms"],
                "cells": [[c[k] for k in cols] + [round(c["gate_ms"], 3), round(c["after_ms"], 3)]
                          for c in cells],
                "chosen": [best[k] for k in cols]}

    if args.cand_bins is not None:
        eng.set_flag("batch_cand_bins", args.cand_bins)
    set_flags(eng, args.flag)
    if args.gate:
        vals = [int(x) for x in args.gate.split(",")]
        step, thr, claim, grid = vals[:4]
        order = vals[4] if len(vals) > 4 else 2
        ahead = vals[5] if len(vals) > 5 else eng.get_flag("batch_lookahead")
        coarse = vals[6] if len(vals) > 6 else eng.get_flag("batch_coarse")
        planes = vals[7] if len(vals) > 7 else eng.get_flag("batch_cand_planes")
        for name, v in (("batch_step_cells", step), ("batch_block_threads", thr),
                        ("batch_claim_seeds", claim), ("batch_gate_blocks", grid), ("batch_order", order),
                        ("batch_lookahead", ahead), ("batch_coarse", coarse), ("batch_cand_planes", planes)):
            eng.set_flag(name, v)
        log("gate fixed: step %d threads %d claim %d grid %d order %d lookahead %d coarse %d cand_planes %d"
            % (step, thr, claim, grid, order, ahead, coarse, planes))

    for i in range(args.warmup):
        res_b, st = batched()
        log("warmup %d: %d hits, %d HSPs" % (i, st["hits"], st["hsps"]))

    g0 = {k: eng.stat(k) for k in ("gate_ns", "gate_hits", "gate_launches", "gate_deferred", "batch_prep_ns",
                                   "batch_device_ns", "batch_read_ns", "batch_post_ns")}
    t_start = time.perf_counter()
    st_b = None
    for i in range(args.steps):
        res_b, st_b = batched()
        log("batched step %d done" % i)
    elapsed = time.perf_counter() - t_start
    cand_step = eng.stat("batch_candidates")  # of the last timed step
    g = {k: eng.stat(k) - g0[k] for k in g0}
    order_used = eng.get_flag("batch_order_used")
    coarse_used = eng.get_flag("batch_coarse_used")
    if not order_used:
        gate_kernel = "k_bgate_packed"
    elif coarse_used:
        gate_kernel = "%s (coarse bound over every hit) + %sk_bgate_exact_ahead (its candidates)" % (
            ("k_bgate_coarse", "k_bgate_coarse2", "k_bgate_coarse2p")[coarse_used - 1],
            "k_cand_hist, k_cand_scatter (binned by reference position) + "
            if eng.get_flag("batch_cand_binned") else "")
    elif (eng.get_flag("batch_step_cells") >= 96 and eng.get_flag("batch_block_threads") == 512
          and eng.get_flag("batch_lookahead") and eng.get_flag("ref_records") == 2):
        gate_kernel = "k_bgate_ahead"
    else:
        gate_kernel = "k_bgate_kmer"
    value = args.steps * query_bp / 1e6 / elapsed
    hashes_b = [E.hash_pass(r) for r in res_b]

    plain = None
    if not args.no_plain:
        t_p = time.perf_counter()
        hashes_p = []
        for r in raws:
            eng.swap_query_raw(r)
            res_p, _ = eng.run_pass()
            hashes_p.append(E.hash_pass(res_p))
        plain_s = time.perf_counter() - t_p
        plain = {"value": round(query_bp / 1e6 / plain_s, 4), "unit": "Mbp/s",
                 "ms_per_step": round(plain_s * 1e3, 2),
                 "equals_batched": hashes_p == hashes_b}
        log("plain per-chunk: %.3f Mbp/s; identical lists: %s" % (plain["value"], plain["equals_batched"]))

    cpu, cells_per_hit, parity_sample = None, None, None
    if not args.no_cpu:
        cpu, cells_per_hit, parity_sample = cpu_baseline(E, S, eng, pieces, traw, ref_enc, index, pos, sub,
                                                         max_hits, args.cpu_blocks, args.threads)

    io = None
    if not args.no_io:
        io = io_measure(res_b, pieces, traw)

    serving = None
    if args.serve_requests > 0:
        serving = serve_arms(E, S, synth, eng, index, pos, ref_enc, sub, max_hits, target,
                             [float(x) for x in args.serve_rates.split(",")], args.serve_requests,
                             args.serve_check)

    roof = None
    if g["gate_ns"] > 0 and cells_per_hit is not None:
        # fused expansion + gate: 4 B of pos_table per hit, plus one reference
        # and one query cell per evaluated cell at W_PACKED bytes per base
        per_hit = 4.0 + 2.0 * W_PACKED * cells_per_hit
        secs = g["gate_ns"] * 1e-9
        achieved = g["gate_hits"] * per_hit / secs / 1e9
        traffic = profiled_traffic(st_b["hits"] if st_b else 0, order_used, coarse_used)
        prof = gate_profile(order_used, coarse_used)
        if not order_used:
            limit = "memory fetch: each hit fetches its own reference lines (%s)" % prof
        elif coarse_used:
            pv = profiled_valu(order_used, coarse_used)
            limit = ("VALU issue: the coarse bound over every hit keeps %.0f%% of the SIMDs' issue slots busy "
                     "(%.2f wave-instructions per hit); HBM traffic %s (%s)"
                     % (100 * pv[1], pv[0], "%.1f B/hit" % (traffic / st_b["hits"]) if traffic else "not profiled",
                        prof) if pv else "VALU issue: the coarse bound over every hit (%s)" % prof)
        else:
            limit = "VALU issue: the two-cell maps over every hit saturate the vector ALU (%s)" % prof
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": "%s (FETCH_SIZE x2.0 + WRITE_SIZE, same workload)" % prof if traffic else None,
                "binding_limit": limit,
                "kernel": "%s (+ k_bgate_long, k_bgate_wave for deferred hits)" % gate_kernel,
                "hit_order": "k-mer-major" if order_used else "seed-major",
                "avg_launch_us": round(g["gate_ns"] / max(g["gate_launches"], 1) / 1e3, 1),
                "hits_per_s": round(g["gate_hits"] / secs, 1),
                "deferred_frac": round(g["gate_deferred"] / max(g["gate_hits"], 1), 5),
                "cells_per_hit": round(cells_per_hit, 3), "bytes_per_hit": round(per_hit, 2),
                "coarse_bound": {"form": ("off", "one hit per lane", "two query records per lane",
                                          "two reference positions per lane")[eng.get_flag("batch_coarse_used")],
                                 "candidates_per_step": cand_step,
                                 "candidates_binned": bool(eng.get_flag("batch_cand_binned")),
                                 "candidate_frac": round(cand_step / max(st_b["hits"], 1), 5)}}

    out = {
        "metric": "query Mbp/s vs 1 Gbp target (batched; plain per-chunk beside it)",
        "value": round(value, 4),
        "unit": "Mbp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "seed-and-filter of %d query chunks (%d bp, 10 kbp-2 Mbp, 5-25%% "
                               "divergent) against a %d bp resident target; 12of19+transitions, "
                               "HOXD70, X=910, T=3000, both strands, max_hits=%d"
                               % (len(pieces), query_bp, tlen, max_hits),
                   "target_bp": tlen, "query_bp": query_bp, "chunks": len(pieces),
                   "max_hits": max_hits, "hits_per_step": st_b["hits"] if st_b else 0,
                   "hsps_per_step": st_b["hsps"] if st_b else 0,
                   "raw_hsps_per_step": st_b["raw_hsps"] if st_b else 0,
                   "hash_first_chunk": "%016x" % hashes_b[0] if hashes_b else None,
                   "index_build": {"where": "device", "ms": round(index_ms, 1),
                                   "positions": eng.stat("pos_len")},
                   "parity_sample_vs_oracle": parity_sample},
        "plain": plain,
        "batched_over_plain": round(value / plain["value"], 3) if plain else None,
        "batched_ms_per_step_by_part": {k[6:-3]: round(g[k] / 1e6 / args.steps, 2)
                                        for k in ("batch_prep_ns", "batch_device_ns", "batch_read_ns",
                                                  "batch_post_ns")},
        "roofline": roof,
        "cpu_baseline": cpu,
        "serving": serving,
        "autotune": tune,
        "io": io,
    }
    print(json.dumps(out), flush=True)
    eng.close()


def serve_arms(E, S, synth, eng, index, pos, ref_enc, sub, max_hits, target, rates, n, n_check):
    """Serving on the resident 1 Gbp target: per arrival rate, n requests with
    clipped lognormal sizes (median 100 kbp, sigma 1.535, 1 kbp..10 Mbp),
    divergence U(2%, 30%), half reverse-complemented, arriving as a Poisson
    process; two engines sharing one resident target serve them by continuous
    batching.
    Latency = completion - arrival. Every arm also checks, on n_check requests
    drawn at random, that their served lists equal the same request processed
    alone on the plain path."""
    import threading
    import numpy as np
    # the second engine reads the first's resident target (index, pos table,
    # planes, clean bits, reference records: held once) with its tuned gate
    eng2 = eng.share()
    arms = []
    for k, rate in enumerate(rates):
        pieces = synth.sample_chunks(target, 0xC0FFEE05 + 17 * k, n, 100_000, 1.535, 1_000,
                                     10_000_000, 0.02, 0.30)
        reqs = [S.pack([("r%d" % i, p)])[0][:len(p)] for i, p in enumerate(pieces)]  # raw blocks
        rng = np.random.default_rng(1000 + k)
        gaps = rng.exponential(1.0 / rate, n)
        bases = sum(len(r) for r in reqs)
        srv = E.Server([eng, eng2], max_batch_requests=256)
        ids = [None] * n

        def submitter():
            t0 = time.perf_counter()
            at = 0.0
            for i, r in enumerate(reqs):
                at += gaps[i]
                d = at - (time.perf_counter() - t0)
                if d > 0:
                    time.sleep(d)
                ids[i] = srv.submit(r)

        th = threading.Thread(target=submitter)
        t_start = time.perf_counter()
        th.start()
        th.join()
        check = set(np.random.default_rng(2000 + k).choice(n, size=min(n_check, n), replace=False).tolist())
        lat, last_done, first_sub, results = [], 0.0, None, {}
        for i in range(n):
            res, tm = srv.wait(ids[i])
            lat.append(tm["t_done"] - tm["t_submit"])
            last_done = max(last_done, tm["t_done"])
            first_sub = tm["t_submit"] if first_sub is None else min(first_sub, tm["t_submit"])
            if i in check:
                results[i] = res
        wall = time.perf_counter() - t_start
        st = srv.stats()
        srv.close()
        same = 0
        for i, res in sorted(results.items()):
            eng.swap_query_raw(reqs[i])
            alone, _ = eng.run_pass()
            same += E.hash_pass(alone) == E.hash_pass(res)
        lat = np.array(lat)
        span = max(last_done - first_sub, 1e-9)
        arms.append({"rate_per_s": rate, "requests": n, "query_bp": int(bases),
                     "offered_mbp_s": round(bases / 1e6 / float(gaps.sum()), 3),
                     "achieved_mbp_s": round(bases / 1e6 / span, 3),
                     "p50_s": round(float(np.percentile(lat, 50)), 4),
                     "p99_s": round(float(np.percentile(lat, 99)), 4),
                     "batches": st["batches"], "max_batch_requests": st["max_batch_requests"],
                     "alone_equals_mixed": "%d/%d" % (same, len(results)),
                     "saturated": bool(bases / float(gaps.sum()) > bases / span * 1.05)})
        log("serve rate %g/s: %.2f Mbp/s, p50 %.3fs p99 %.3fs, %d batches (wall %.1fs)" %
            (rate, arms[-1]["achieved_mbp_s"], arms[-1]["p50_s"], arms[-1]["p99_s"],
             st["batches"], wall))
    bytes2 = {"first_engine": eng.stat("device_bytes"), "second_engine": eng2.stat("device_bytes"),
              "note": "the target is held once, counted in the first engine; the second holds its "
                      "work buffers only"}
    eng2.close()
    return {"engines": 2, "shared_target": True, "device_bytes": bytes2, "arms": arms,
            "note": "offered load above the service rate saturates the loop; latency then "
                    "includes queueing"}


def c4_workload(args):
    """The mammal-scale shape on one GPU: a 3 Gbp target (20 x 150 Mbp
    records, 45% repeat-derived, uppercase) packed into one reference block,
    its 12of19 index (about 3e9 positions, 12 GB) built on the device and held
    there; query records of 50 Mbp sampled from it at 12% divergence with
    segmental duplications and reverse-complemented segments, each one
    hspz_filter_batch call (its query bin against the resident reference bin).
    Parity: a 1 Mbp block and a 20 kbp slice of the first chunk, as blocks of
    their own in one call, against the plain path and the oracle."""
    from hspz_amd import engine as E
    from hspz_amd import sequence as S
    from hspz_amd import seqio as X
    from hspz_amd import synth
    if E.device_count() < 1:
        raise SystemExit("no HIP device visible")
    t0 = time.time()
    target = synth.c4_target()
    recs = [("chr%d" % (r + 1), target[r * synth.C4_RECORD_LEN:(r + 1) * synth.C4_RECORD_LEN].tobytes())
            for r in range(synth.C4_RECORDS)]
    buf, chrs, bl = S.pack(recs)
    traw = buf[:bl]
    del recs
    ref_enc = X.encode(traw)
    log("c4 target: %d bp in %d records, generated and packed in %.1fs" % (bl, synth.C4_RECORDS, time.time() - t0))
    sub = E.build_sub_mat("", 910)
    _, total_mem = _device_memory(E)
    max_hits = E.default_max_hits(total_mem)
    t1 = time.time()
    eng = E.Engine(None, None, ref_enc, sub, seed_size=19, xdrop=910, hspthresh=3000, noentropy=False,
                   max_hits=max_hits, device_index="12of19")
    log("engine up in %.1fs: device index %.0f ms, %d positions, %.1f GB held"
        % (time.time() - t1, eng.stat("index_build_ns") / 1e6, eng.stat("pos_len"), eng.stat("device_bytes") / 1e9))
    if args.cand_bins is not None:
        eng.set_flag("batch_cand_bins", args.cand_bins)
    set_flags(eng, args.flag)
    tot_bp, tot_s, hits, hsps, raw_hsps, passes, per, parts, binned = 0, 0.0, 0, 0, 0, 0, [], [], []
    first = None
    for k in range(args.c4_chunks):
        t2 = time.time()
        q = synth.c4_query(target, 0xC0FFEE04 + 1 + k)
        if first is None:
            first = q
        gen = time.time() - t2
        p0 = {n: eng.stat(n) for n in ("batch_prep_ns", "batch_device_ns", "batch_read_ns", "batch_post_ns")}
        t3 = time.perf_counter()
        res, st = eng.filter_batch([(q, None)])
        dt = time.perf_counter() - t3
        parts.append({n[6:-3]: round((eng.stat(n) - v) / 1e9, 3) for n, v in p0.items()})
        tot_bp += len(q)
        tot_s += dt
        hits += st["hits"]
        hsps += st["hsps"]
        raw_hsps += st["raw_hsps"]
        passes += eng.stat("batch_passes")
        per.append(round(dt, 3))
        binned.append(bool(eng.get_flag("batch_cand_binned")))  # the chunk's last pass
        log("c4 chunk %d: %d bp (generated in %.1fs) in %.2fs = %.2f Mbp/s; %d hits, %d HSPs, %d device passes; "
            "seconds by part %s" % (k, len(q), gen, dt, len(q) / 1e6 / dt, st["hits"], st["hsps"],
                                    eng.stat("batch_passes"), parts[-1]))
    # parity sample
    blk, sl = first[10_000_000:11_000_000], first[20_000_000:20_020_000]
    res_b, _ = eng.filter_batch([(blk, None), (sl, None)])
    eng.swap_query_raw(blk)
    res_p, _ = eng.run_pass()
    plain_ok = E.hash_pass(res_p) == E.hash_pass(res_b[0])
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O
    index, pos = eng.index()
    sbuf, schrs, sbl = S.pack([("slice", sl)])
    src, _ = S.reverse_complement(sbuf, schrs, sbl)
    o = O.Oracle(traw, sbuf[:sbl], src, sub, max_hits=max_hits, index=(index, pos), ref_enc=ref_enc)
    _, ost, oh = o.run_pass(threads=args.threads)
    oracle_ok = oh == E.hash_pass(res_b[1])
    log("c4 parity: batched == plain on a 1 Mbp block: %s; batched == oracle on a 20 kbp slice: %s (%d hits)"
        % (plain_ok, oracle_ok, ost["hits"]))
    out = {"metric": "query Mbp/s, mammal-scale shape on one GPU (3 Gbp resident target, 50 Mbp query chunks)",
           "value": round(tot_bp / 1e6 / tot_s, 4), "unit": "Mbp/s", "n_gpus": 1, "higher_is_better": True,
           "dtype": "int32", "data": "synthetic",
           "config": {"workload": "c4: %d query chunk(s) of 50 Mbp (12%% divergent, 10%% segmental duplications "
                                  "1-5%% divergent, 10%% reverse-complemented segments) against a %d bp target "
                                  "(20 x 150 Mbp, 45%% repeat-derived, uppercase); 12of19+transitions, HOXD70, "
                                  "X=910, T=3000, both strands, max_hits=%d" % (args.c4_chunks, bl, max_hits),
                      "target_bp": bl, "query_bp": tot_bp, "chunks": args.c4_chunks, "max_hits": max_hits},
           "positions": eng.stat("pos_len"), "hits": hits, "hsps": hsps, "raw_hsps": raw_hsps,
           "device_passes": passes, "overflow_reruns": eng.stat("batch_reruns"), "seconds_per_chunk": per,
           "candidates_binned": binned,
           "seconds_by_part": parts,
           "peak_device_bytes": eng.stat("peak_device_bytes"), "device_bytes_held": eng.stat("device_bytes"),
           "ref_records": eng.get_flag("ref_records") == 2,
           "index_build_ms": round(eng.stat("index_build_ns") / 1e6, 1),
           "batched_equals_plain_1mbp": plain_ok, "batched_equals_oracle_20kbp": oracle_ok}
    print(json.dumps(out), flush=True)
    eng.close()


def cpu_info():
    """CPU model and logical CPUs of this host (the GPU box's host when run there)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


def cpu_baseline(E, S, eng, pieces, traw, ref_enc, index, pos, sub, max_hits, n_blocks, threads):
    """The C oracle (the reference's semantics restated; the reference itself
    cannot be built here) on a bounded sample of the same workload: the leading
    20 kbp of n_blocks query chunks, each its own block (both strands, against
    the whole resident target). All-core: one block per worker thread, `threads`
    workers busy at once (the oracle call releases the interpreter lock).
    1-thread: the first two blocks run one after the other on one thread. Every
    block's lists are checked equal to the GPU's (one filter_batch over the
    sample)."""
    import concurrent.futures as F
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi as O
    blocks = []
    for c in pieces:
        if len(blocks) >= n_blocks:
            break
        part = c[:20_000]
        if len(part) <= 19:
            continue
        buf, chrs, bl = S.pack([("s", part)])
        rcb, _ = S.reverse_complement(buf, chrs, bl)
        blocks.append((buf[:bl], rcb))
    oracles = [O.Oracle(traw, f, r, sub, max_hits=max_hits, index=(index, pos), ref_enc=ref_enc)
               for f, r in blocks]
    sample_bp = sum(len(f) for f, _ in blocks)
    # 1 thread
    t1 = time.perf_counter()
    one_bp = 0
    for o, (f, _) in list(zip(oracles, blocks))[:2]:
        o.run_pass(threads=1)
        one_bp += len(f)
    one_s = time.perf_counter() - t1
    # all cores: one block per worker
    t1 = time.perf_counter()
    with F.ThreadPoolExecutor(max_workers=threads) as ex:
        outs = list(ex.map(lambda o: o.run_pass(threads=1), oracles))
    all_s = time.perf_counter() - t1
    hits = sum(st["hits"] for _, st, _ in outs)
    cells = sum(st["cells"] for _, st, _ in outs)
    g_res, _ = eng.filter_batch([(f, None) for f, _ in blocks])
    ok = all(E.hash_pass(g) == h for g, (_, _, h) in zip(g_res, outs))
    model, ncpu, usable = cpu_info()
    busy = min(threads, len(blocks))
    cpu = {"value": round(sample_bp / 1e6 / all_s, 5), "unit": "query Mbp/s", "cores": busy, "kind": "port",
           "one_thread_mbp_s": round(one_bp / 1e6 / one_s, 6),
           "threads_busy": busy, "cpu_model": model, "nproc": ncpu, "usable_cpus": usable,
           "sample": "%d bp: the leading 20 kbp of %d query chunks, each its own block (both strands) "
                     "against the full target, one block per worker thread on %d threads (%d hits, %.1f s); "
                     "1-thread figure from %d bp run serially (%.1f s)"
                     % (sample_bp, len(blocks), busy, hits, all_s, one_bp, one_s)}
    log("cpu baseline: %.5f Mbp/s on %d threads, %.6f Mbp/s on 1 thread (%s, nproc %d); sample parity %s"
        % (cpu["value"], busy, cpu["one_thread_mbp_s"], model, ncpu, ok))
    return cpu, cells / max(hits, 1), ok


def io_measure(results, pieces, traw):
    """Host paths around the engine, timed outside the step: the step's lists
    written as .segments files (one work unit per chunk, the target as one
    record) into a directory, with -D partitions, and into one .tar.gz; and the
    native reader over 100 Mbp of the target as 60-column FASTA and 20 Mbp of it
    gzipped."""
    import gzip
    import shutil
    import tempfile
    import numpy as np
    from hspz_amd import seqio as X
    d = tempfile.mkdtemp(prefix="hspz_bench_io_")
    out = {}
    try:
        r_chrs = [X.Chr("target", 0, len(traw))]
        q_chrs = [[X.Chr("q%d" % i, 0, len(c))] for i, c in enumerate(pieces)]
        n_hsps = sum(len(a) for res in results for fw, rc in res for a in (fw, rc))
        for name, tarball, diag in (("directory", None, False), ("directory_D", None, True),
                                    ("tar_gz", "", False)):
            t0 = time.perf_counter()
            em = X.Emitter(os.path.join(d, name), tarball=tarball, diagonal=diag)
            for i, res in enumerate(results):
                em.emit_unit(res, r_chrs, q_chrs[i], q_chrs[i], query_bin=i)
            rep = em.finish()
            ms = (time.perf_counter() - t0) * 1e3
            out["emit_" + name] = {"ms": round(ms, 2), "files": rep["files"], "hsps": n_hsps,
                                   "bytes_in": rep["bytes_in"], "bytes_out": rep["bytes_out"],
                                   "mb_s": round(rep["bytes_in"] / 1e3 / max(ms, 1e-6), 1)}
        for name, mbp, gz in (("read_fasta", 100, False), ("read_fasta_gz", 20, True)):
            body = np.frombuffer(traw[:mbp * 1_000_000 // 60 * 60], dtype=np.uint8).reshape(-1, 60)
            text = b">target\n" + np.hstack([body, np.full((len(body), 1), 10, np.uint8)]).tobytes()
            path = os.path.join(d, name + (".fa.gz" if gz else ".fa"))
            with open(path, "wb") as fh:
                fh.write(gzip.compress(text, compresslevel=1) if gz else text)
            t0 = time.perf_counter()
            s = X.SeqSet.read(path)
            ms = (time.perf_counter() - t0) * 1e3
            ok = s.block_len == len(body) * 60 and s.raw[:1000] == traw[:1000]
            out[name] = {"ms": round(ms, 1), "bases": s.block_len, "file_bytes": s.bytes_read,
                         "mb_s": round(len(text) / 1e3 / ms, 1), "ok": bool(ok)}
            if not gz:
                t0 = time.perf_counter()
                X.encode(s.raw)
                ms = (time.perf_counter() - t0) * 1e3
                out["encode_host"] = {"ms": round(ms, 1), "mb_s": round(s.block_len / 1e3 / ms, 1)}
            s.close()
        log("io: emit %s ms (dir) / %s ms (-D) / %s ms (tar.gz); FASTA read %s MB/s, gz %s MB/s"
            % (out["emit_directory"]["ms"], out["emit_directory_D"]["ms"], out["emit_tar_gz"]["ms"],
               out["read_fasta"]["mb_s"], out["read_fasta_gz"]["mb_s"]))
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return out


def _device_memory(E):
    import ctypes as C
    f, t = C.c_uint64(), C.c_uint64()
    E._check(E.lib().hspz_device_memory(0, C.byref(f), C.byref(t)))
    return int(f.value), int(t.value)


if __name__ == "__main__":
    main()

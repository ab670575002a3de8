#!/usr/bin/env python3
"""Seed-and-filter benchmark: query Mbp/s against a 1 Gbp resident target.

Workload: a 1,000,000,000 bp uppercase target (iid, 41% GC, seed 0xC0FFEE05)
indexed with 12of19 and held in HBM; the query is a set of chunks sampled from
it (clipped lognormal lengths, median 200 kbp, 10 kbp..2 Mbp, divergence
U(5%, 25%), half reverse-complemented), packed as records of one query block.
HOXD70, X-drop 910, threshold 3000, both strands, entropy on, max_hits pinned.

A step = one seed-and-filter pass over the whole query block: encoded query in,
ordered HSP list out (seeding, hit counting, expansion, score gate,
compaction, materialization, dedup). Target generation, index build and index
upload are setup, outside the timed region.

Prints one JSON line (rank 0, N = 1). Extra objects: `roofline` for the score
gate kernel (HIP events around every gate launch in the timed region) and
`cpu_baseline` (the C oracle, multi-threaded, on a bounded sample of the same
query block, timed on this host).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MAX_HITS = 16711680


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--target-mbp", type=float, default=1000.0)
    ap.add_argument("--query-mbp", type=float, default=4.0)
    ap.add_argument("--cpu-sample-kbp", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the seed-and-filter path has no cross-GPU "
              "step here (work units shard by reference/query bin; see the design notes). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from hspz_amd import engine as E
    from hspz_amd import sequence as S
    from hspz_amd import synth

    if E.device_count() < 1:
        raise SystemExit("no HIP device visible")

    t0 = time.time()
    tlen = int(args.target_mbp * 1e6)
    target = synth.random_bases(0xC0FFEE05, tlen)
    log("target %d bp generated in %.1fs" % (tlen, time.time() - t0))
    t1 = time.time()
    traw = target.tobytes()
    index, pos = E.seed_table_build(traw, "12of19", 1, args.threads)
    log("index built in %.1fs (%d positions)" % (time.time() - t1, len(pos)))
    ref_enc = S.encode(traw)
    sub = E.build_sub_mat("", 910)

    # query chunks (C3 model) until query_mbp is reached
    qwant = int(args.query_mbp * 1e6)
    chunks, total, seed = [], 0, 0xC0FFEE03
    while total < qwant:
        more = synth.sample_chunks(target, seed, 16, 200_000, 0.768, 10_000, 2_000_000, 0.05, 0.25)
        seed += 1
        for c in more:
            if total >= qwant:
                break
            chunks.append(c)
            total += len(c)
    records = [("q%d" % i, c) for i, c in enumerate(chunks)]
    qbuf, qchrs, qbl = S.pack(records)
    q_raw = qbuf[:qbl]
    q_rc, _ = S.reverse_complement(qbuf, qchrs, qbl)
    q_enc, q_rc_enc = S.encode(q_raw), S.encode(q_rc)
    query_bp = sum(len(c) for c in chunks)
    log("query: %d chunks, %d bp" % (len(chunks), query_bp))

    t1 = time.time()
    eng = E.Engine(index, pos, ref_enc, sub, seed_size=19, xdrop=910, hspthresh=3000,
                   noentropy=False, max_hits=MAX_HITS)
    log("engine up (index + target resident) in %.1fs" % (time.time() - t1))

    def step():
        eng.swap_query(q_enc, q_rc_enc)
        return eng.run_pass()

    for i in range(args.warmup):
        res, st = step()
        log("warmup %d: %d hits, %d HSPs" % (i, st["hits"], st["hsps"]))

    g_ns0, g_hits0, g_l0 = eng.stat("gate_ns"), eng.stat("gate_hits"), eng.stat("gate_launches")
    g_def0 = eng.stat("gate_deferred")
    t_start = time.perf_counter()
    stats = None
    for i in range(args.steps):
        res, stats = step()
        log("step %d done" % i)
    elapsed = time.perf_counter() - t_start
    g_ns = eng.stat("gate_ns") - g_ns0
    g_hits = eng.stat("gate_hits") - g_hits0
    g_launch = eng.stat("gate_launches") - g_l0
    g_def = eng.stat("gate_deferred") - g_def0
    value = args.steps * query_bp / 1e6 / elapsed
    hits_per_step = stats["hits"] if stats else 0
    hsp_hash = E.hash_pass(res)

    # CPU baseline: the oracle on the first cpu_sample_kbp of the query block
    cpu = None
    cells_per_hit = None
    parity_sample = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_ffi as O
        n = min(qbl, args.cpu_sample_kbp * 1000)
        s_raw = q_raw[:n]
        s_rc = S.reverse_complement(s_raw + b"&", [S.Chr("s", 0, n)], n)[0]
        orc = O.Oracle(traw, s_raw, s_rc, sub, max_hits=MAX_HITS, index=(index, pos))
        t1 = time.perf_counter()
        o_res, o_st, o_hash = orc.run_pass(threads=args.threads)
        cpu_s = time.perf_counter() - t1
        cells_per_hit = o_st["cells"] / max(o_st["hits"], 1)
        cpu = {"value": round(n / 1e6 / cpu_s, 4), "unit": "query Mbp/s", "cores": args.threads,
               "kind": "port",
               "sample": "first %d bp of the query block, both strands, against the full "
                         "target (%d hits, %.1f s)" % (n, o_st["hits"], cpu_s)}
        eng.swap_query(S.encode(s_raw), S.encode(s_rc))
        g_res, g_st = eng.run_pass()
        parity_sample = (E.hash_pass(g_res) == o_hash)
        log("cpu baseline %.3f Mbp/s on %d threads; sample parity %s" %
            (cpu["value"], args.threads, parity_sample))

    roof = None
    if g_ns > 0 and cells_per_hit is not None:
        # score gate algorithmic bytes: 9 per hit (8 B anchor + 1 B flag) plus
        # 2 bytes per evaluated cell (one reference and one query code byte)
        alg_bytes = g_hits * (9.0 + 2.0 * cells_per_hit)
        achieved = alg_bytes / (g_ns * 1e-9) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_gate", "avg_launch_us": round(g_ns / max(g_launch, 1) / 1e3, 2),
                "hits_per_s": round(g_hits / (g_ns * 1e-9), 1),
                "deferred_frac": round(g_def / max(g_hits, 1), 5),
                "cells_per_hit": round(cells_per_hit, 3)}

    out = {
        "metric": "query Mbp/s vs 1 Gbp target (plain per-chunk path)",
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
        "config": {"workload": "plain seed-and-filter, %d bp target resident, %d query chunks "
                               "(%d bp), 12of19+transitions, HOXD70, X=910, T=3000, both "
                               "strands, max_hits=%d" % (tlen, len(chunks), query_bp, MAX_HITS),
                   "target_bp": tlen, "query_bp": query_bp, "hits_per_step": hits_per_step,
                   "hsps_per_step": stats["hsps"] if stats else 0,
                   "hash_pass": "%016x" % hsp_hash, "parity_sample": parity_sample},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()

"""The engine boundary's auxiliary entry points through the C ABI: device
seeds byte for byte against host seeds (Engine::check_seed_bytes,
gpu/mod.rs:1464-1495), asynchronous seed uploads from the pinned pool
(gpu/mod.rs:1165-1192, 1246-1309), the pre-dedup record list
(FilterOutput::raw, gpu/mod.rs:55-67), the launch overhead probe
(gpu/mod.rs:1210-1236), and engines sharing one resident target
(gpu/mod.rs:1130-1148)."""
import sys

import numpy as np
import pytest

from helpers import Pair, same_intervals
from hspz_amd import engine as E
from hspz_amd import synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_generate_seeds_bytes_equal_chunk_seeds(gpu_available):
    """hspz_generate_seeds' u64 array equals hspz_chunk_seeds' byte for byte,
    both strands, with and without transitions, on a query with ambiguous
    bases and soft-masked runs (positions whose window holds one are skipped)."""
    _say("case device seeds vs host seeds, byte for byte")
    t, q = synth.c2_pair(300_000, seed=0xB0DA)
    p = Pair.raw(t, q)
    q = bytearray(p.q_raw)
    rng = np.random.default_rng(5)
    for i in rng.integers(0, len(q), 300):
        q[i] = ord("N")
    p = Pair.raw(t, bytes(q))
    with p.engine() as eng:
        for rev, (lo, hi), trans in [(False, (0, 120_000), True), (True, (7, 200_001), True),
                                     (False, (50_000, 50_500), False), (True, (0, 1), True)]:
            raw = p.q_rc_raw if rev else p.q_raw
            want = E.chunk_seeds(raw, lo, hi, transitions=trans, threads=4)
            n = eng.generate_seeds(1, rev, lo, hi, transitions=trans)
            assert n == len(want)
            eng.check_seed_bytes(1, want)
            if len(want):  # a planted difference is reported, with its index
                bad = want.copy()
                bad[len(bad) // 2] ^= 1
                with pytest.raises(E.HspzError, match="differs at %d" % (len(bad) // 2)):
                    eng.check_seed_bytes(1, bad)
            with pytest.raises(E.HspzError, match="count differs"):
                eng.check_seed_bytes(1, want[:-1] if len(want) else np.zeros(1, np.uint64))


def test_async_upload_from_pinned_pool(gpu_available):
    """Seeds staged from pinned pool buffers, one batch ahead on the other
    slot, give the lists of synchronous uploads; the pool reuses a returned
    buffer; the lifetime rule holds across sync_seeds and a generate."""
    _say("case asynchronous seed uploads, pinned pool")
    t, q = synth.c1_pair(200_000, seed=91)
    p = Pair.raw(t, q)
    batches = [(False, (0, 40_000)), (True, (20_000, 70_000)), (False, (60_000, 110_000)),
               (True, (100_000, 150_000))]
    seeds = [E.chunk_seeds(p.q_rc_raw if rev else p.q_raw, lo, hi, threads=4) for rev, (lo, hi) in batches]
    with p.engine() as eng:
        want = []
        for (rev, _), s in zip(batches, seeds):
            eng.upload_seeds(0, s)
            want.append(eng.seed_and_filter(0, rev))
        assert eng.stat("pinned_active") == 0
        up0 = eng.stat("seed_uploads")
        bufs = []
        for i, s in enumerate(seeds):
            b = eng.pinned_take(len(s))
            b[:] = s
            bufs.append(b)
        assert eng.stat("pinned_active") == 1
        eng.upload_seeds_async(0, bufs[0])
        for i, (rev, _) in enumerate(batches):
            if i + 1 < len(batches):  # the next batch's copy runs while this one computes
                eng.upload_seeds_async((i + 1) % 2, bufs[i + 1])
            got = eng.seed_and_filter(i % 2, rev)
            assert got[1:] == want[i][1:] and np.array_equal(got[0], want[i][0]), "batch %d" % i
        assert eng.stat("seed_uploads") == up0 + len(batches)
        # a returned buffer is handed out again for a request it can hold
        addr = bufs[0].ctypes.data
        for b in bufs:
            eng.pinned_give(b)
        again = eng.pinned_take(min(len(s) for s in seeds))
        assert again.ctypes.data in {b.ctypes.data for b in bufs}
        eng.pinned_give(again)
        with pytest.raises(E.HspzError):  # not from the pool
            eng.pinned_give(np.zeros(4, dtype=np.uint64))
        # sync_seeds ends the lifetime; a generate after an async upload waits for it
        eng.upload_seeds_async(1, seeds[1])
        eng.sync_seeds(1)
        eng.check_seed_bytes(1, seeds[1])
        eng.upload_seeds_async(1, seeds[2])
        n = eng.generate_seeds(1, False, 0, 40_000)
        eng.check_seed_bytes(1, seeds[0])
        assert n == len(seeds[0]) and addr


def test_raw_pre_dedup_list(gpu_available):
    """With dump_raw, hspz_filter_output carries the raw_hsps records before
    dedup, and dedup_and_order of them is the returned list (one max_hits
    chunk); without it the list is empty."""
    _say("case raw pre-dedup records")
    t, q = synth.c1_pair(200_000, seed=77)
    p = Pair.raw(t, q)
    o = p.oracle()
    with p.engine() as eng:
        for rev, (lo, hi) in [(False, (0, 60_000)), (True, (10_000, 90_000))]:
            raw = p.q_rc_raw if rev else p.q_raw
            seeds = E.chunk_seeds(raw, lo, hi, threads=4)
            eng.upload_seeds(0, seeds)
            res, nh, nraw, rows = eng.seed_and_filter(0, rev, with_raw=True)
            assert len(rows) == 0 and nraw > 0
            eng.set_flag("dump_raw", 1)
            res2, nh2, nraw2, rows = eng.seed_and_filter(0, rev, with_raw=True)
            eng.set_flag("dump_raw", 0)
            want, st = o.seed_and_filter(seeds, rev)
            assert np.array_equal(res2, res) and np.array_equal(res, want)
            assert nraw2 == nraw == st["raw_hsps"] == len(rows) and nh2 == nh
            assert len(rows) >= len(res)
            assert np.array_equal(E.dedup_and_order(rows), res)


def test_launch_overhead(gpu_available):
    _say("case launch overhead")
    p = Pair.fixture("identical")
    with p.engine() as eng:
        ms = eng.launch_overhead_ms(2000)
        _say("  %.4f ms per launch" % ms)
        assert 0 < ms < 1.0
        with pytest.raises(E.HspzError):
            eng.launch_overhead_ms(0)


def test_engines_share_one_target(gpu_available):
    """Two engines over one resident target: the second holds no copy (its
    device bytes are its own work buffers), both give the lists of the first
    alone, and the target outlives the engine that built it."""
    _say("case engines sharing one resident target")
    target = synth.random_bases(0x5A4E, 2_000_000)
    pieces = synth.sample_chunks(target, 0x5A4E, 6, 20_000, 0.8, 5_000, 60_000, 0.02, 0.2)
    raws = [bytes(c) for c in pieces]
    from hspz_amd import seqio as X
    sub = E.build_sub_mat("", 910)
    base = E.Engine(None, None, X.encode(target.tobytes()), sub, max_hits=16711680, device_index="12of19")
    held = base.stat("device_bytes")
    want = []
    for r in raws:
        base.swap_query_raw(r)
        want.append(base.run_pass()[0])
    b0, _ = base.filter_batch([(r, None) for r in raws])
    base_bytes = base.stat("device_bytes")  # the target and this batch's work buffers
    other = base.share()
    assert other.stat("reference_uploads") == 0
    assert other.stat("pos_len") == base.stat("pos_len") and other.stat("index_len") == base.stat("index_len")
    got, _ = other.filter_batch([(r, None) for r in raws])
    for i, (g, w) in enumerate(zip(got, b0)):
        assert same_intervals(g, w), "batch of chunk %d differs on the sharing engine" % i
    # the sharing engine's bytes are the same work buffers without the target
    assert other.stat("device_bytes") <= base_bytes - 0.9 * held, (other.stat("device_bytes"), base_bytes, held)
    third = other.share()
    base.close()  # the target stays resident for the engines still holding it
    for i, r in enumerate(raws):
        third.swap_query_raw(r)
        assert same_intervals(third.run_pass()[0], want[i]), "chunk %d differs after the first engine closed" % i
    got, _ = other.filter_batch([(r, None) for r in raws])
    for i, (g, w) in enumerate(zip(got, b0)):
        assert same_intervals(g, w)
    other.close()
    third.close()

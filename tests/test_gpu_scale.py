"""The benchmark configurations at their full shape on one GPU.

C3: 1,024 lognormal query chunks (10 kbp - 2 Mbp, median 200 kbp, 5-25%
divergent, half reverse-complemented) against a 500 Mbp target in ONE
hspz_filter_batch call (run internally as several device passes), plus 8
chunks cut to 20 kbp in the same call. Every chunk's lists are compared with
the plain per-chunk path on 16 chunks drawn at random, and with the oracle on
the 8 cut chunks.

C2: the 100 Mbp x 100 Mbp soft-masked pair, plain per-batch pass vs the
batched path over the whole query (both strands, all 800 batches), and the
oracle on a 20 kbp slice.

C5: serving on the resident 1 Gbp target with Poisson arrivals (400/s) of
small requests (1-20 kbp); every served list is compared with the oracle and
with the same request processed alone.

max_hits is pinned at the reference's smoke value, 16,711,680 (at 500 Mbp a
full 250 kbp batch splits into about six max_hits chunks)."""
import concurrent.futures as F
import sys
import threading
import time

import numpy as np
import pytest

import oracle_ffi as O
from helpers import same_intervals
from hspz_amd import engine as E
from hspz_amd import sequence as S
from hspz_amd import synth

pytestmark = pytest.mark.gpu

MAX_HITS = 16711680


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _raw(records):
    buf, chrs, bl = S.pack(records)
    rc, _ = S.reverse_complement(buf, chrs, bl)
    return buf[:bl], rc


def _oracle_all(traw, ref_enc, index, pos, sub, blocks, **kw):
    """The oracle's run_pass on each (raw fwd, raw rc) block, blocks in parallel."""
    def one(b):
        o = O.Oracle(traw, b[0], b[1], sub, max_hits=MAX_HITS, index=(index, pos), ref_enc=ref_enc)
        return o.run_pass(threads=2, **kw)
    with F.ThreadPoolExecutor(max_workers=8) as ex:
        return list(ex.map(one, blocks))


def test_c3_one_call_1024_chunks(gpu_available):
    _say("case C3: 1024 chunks + 8 cut chunks vs a 500 Mbp target in one filter_batch")
    t0 = time.time()
    target = synth.random_bases(0xC0FFEE03, 500_000_000)
    traw = target.tobytes()
    pieces = synth.sample_chunks(target, 0xC0FFEE03, 1024, 200_000, 0.768, 10_000, 2_000_000, 0.05, 0.25)
    cut = [p[:20_000] for p in pieces[:8]]
    blocks = [_raw([("c%d" % i, p)]) for i, p in enumerate(pieces + cut)]
    ref_enc = S.encode(traw)
    sub = O.sub_mat("", 910)
    qbp = sum(len(p) for p in pieces)
    _say("  generated in %.1fs: %d chunks, %d bp" % (time.time() - t0, len(pieces), qbp))
    assert len(pieces) == 1024 and qbp > 150_000_000
    with E.Engine(None, None, ref_enc, sub, max_hits=MAX_HITS, device_index="12of19") as eng:
        t1 = time.perf_counter()
        got, st = eng.filter_batch([(f, None) for f, _ in blocks])
        secs = time.perf_counter() - t1
        passes = eng.stat("batch_passes")
        _say("  one call: %.2fs, %.1f Mbp/s, %d device passes, %d batches, %d hits, %d HSPs, peak %.1f GB"
             % (secs, qbp / 1e6 / secs, passes, st["batches"], st["hits"], st["hsps"],
                eng.stat("peak_device_bytes") / 1e9))
        assert passes > 1 and st["hsps"] > 0
        rng = np.random.default_rng(0xC3)
        for i in sorted(rng.choice(1024, 16, replace=False).tolist()):
            eng.swap_query_raw(blocks[i][0])
            plain, _ = eng.run_pass()
            assert same_intervals(plain, got[i]), "chunk %d: batched differs from plain" % i
        index, pos = eng.index()
    want = _oracle_all(traw, ref_enc, index, pos, sub, blocks[1024:])
    for k, (w, _, h) in enumerate(want):
        assert same_intervals(got[1024 + k], w) and E.hash_pass(got[1024 + k]) == h, "cut chunk %d" % k


def test_c2_full_size(gpu_available):
    _say("case C2: 100 Mbp x 100 Mbp soft-masked pair, plain vs batched, oracle on a 20 kbp slice")
    t0 = time.time()
    t, q = synth.c2_pair(100_000_000)
    sub = O.sub_mat("", 910)
    ref_enc = S.encode(t)
    fwd, rc = _raw([("query", q)])
    sl_f, sl_r = _raw([("slice", q[37_000_000:37_020_000])])
    _say("  generated in %.1fs" % (time.time() - t0))
    with E.Engine(None, None, ref_enc, sub, max_hits=MAX_HITS, device_index="12of19") as eng:
        t1 = time.perf_counter()
        eng.swap_query_raw(fwd)
        plain, pst = eng.run_pass()
        tp = time.perf_counter() - t1
        t1 = time.perf_counter()
        got, bst = eng.filter_batch([(fwd, None), (sl_f, None)])
        tb = time.perf_counter() - t1
        _say("  plain %.2fs (%.1f Mbp/s), batched %.2fs (%.1f Mbp/s, %d passes): %d batches, %d hits, %d HSPs"
             % (tp, 100 / tp, tb, 100.02 / tb, eng.stat("batch_passes"), pst["batches"], pst["hits"],
                pst["hsps"]))
        # 10 intervals x 40 batches x 2 strands, plus the few the indels' extra length adds
        assert 800 <= pst["batches"] <= 810 and pst["hits"] > 1e10 and bst["hits"] > pst["hits"]
        assert same_intervals(got[0], plain), "C2: batched differs from plain"
        # its passes hold more query bases than candidate binning takes by default; forced, it
        # bins them (multi-tile scatters over large lists) and the lists stay the same
        eng.set_flag("batch_cand_bins", 2)
        t1 = time.perf_counter()
        forced, _ = eng.filter_batch([(fwd, None)])
        _say("  batched with the candidates binned: %.2fs" % (time.perf_counter() - t1))
        assert eng.get_flag("batch_cand_binned") == 1
        assert same_intervals(forced[0], plain), "C2: batched with binned candidates differs from plain"
        eng.set_flag("batch_cand_bins", 1)
        index, pos = eng.index()
    (w, _, h), = _oracle_all(t, ref_enc, index, pos, sub, [(sl_f, sl_r)])
    assert same_intervals(got[1], w) and E.hash_pass(got[1]) == h, "C2 slice differs from the oracle"


def test_c5_serving_1gbp_poisson(gpu_available):
    _say("case C5: serving on a resident 1 Gbp target, Poisson arrivals of 1-20 kbp requests")
    t0 = time.time()
    target = synth.random_bases(0xC0FFEE05, 1_000_000_000)
    traw = target.tobytes()
    ref_enc = S.encode(traw)
    sub = O.sub_mat("", 910)
    pieces = synth.sample_chunks(target, 0xC5, 48, 5_000, 1.0, 1_000, 20_000, 0.02, 0.30)
    reqs = [_raw([("r%d" % i, p)]) for i, p in enumerate(pieces)]
    _say("  generated in %.1fs: %d requests, %d bp" % (time.time() - t0, len(reqs), sum(map(len, pieces))))
    e1 = E.Engine(None, None, ref_enc, sub, max_hits=MAX_HITS, device_index="12of19")
    e2 = e1.share()  # one resident target for both serving engines
    gaps = np.random.default_rng(0xC5).exponential(1.0 / 400, len(reqs))
    ids = [None] * len(reqs)
    with E.Server([e1, e2], max_batch_requests=16) as srv:
        def submitter():
            for i, (f, _) in enumerate(reqs):
                time.sleep(gaps[i])
                ids[i] = srv.submit(f, None)
        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        got = [srv.wait(i) for i in ids]
        sst = srv.stats()
    lat = np.array([tm["t_done"] - tm["t_submit"] for _, tm in got])
    _say("  %d batches (up to %d requests each), p50 %.3fs, p99 %.3fs"
         % (sst["batches"], sst["max_batch_requests"], np.percentile(lat, 50), np.percentile(lat, 99)))
    assert sst["failed_batches"] == 0 and sst["requests"] == len(reqs)
    for k, (f, _) in enumerate(reqs):
        e1.swap_query_raw(f)
        alone, _ = e1.run_pass()
        assert same_intervals(got[k][0], alone), "request %d differs from alone" % k
    index, pos = e1.index()
    e1.close()
    e2.close()
    want = _oracle_all(traw, ref_enc, index, pos, sub, reqs)
    for k, (w, _, h) in enumerate(want):
        assert same_intervals(got[k][0], w) and E.hash_pass(got[k][0]) == h, "request %d differs from the oracle" % k

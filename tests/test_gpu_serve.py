"""The serving loop: requests submitted together are batched (continuous
batching over two engines) and every request's lists equal the same request
processed alone on the plain path and the oracle's run_pass on it alone;
admission refuses a request over the HBM budget; a result is collected once."""
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


def _setup(n_target=400_000, seed=0x5E11):
    target = synth.random_bases(seed, n_target).tobytes()
    sub = O.sub_mat("", 910)
    index, pos = O.seed_table(target, O.shape("12of19"), 1)
    return target, sub, index, pos


def _requests(target, n, seed, raw_every=0):
    """Requests as (fwd, rc, raw_fwd, raw_rc): encoded strands for the engines;
    the raw bytes for the oracle. With raw_every = k, every k-th request is
    submitted raw (the device encodes both strands)."""
    t = np.frombuffer(target, dtype=np.uint8)
    pieces = synth.sample_chunks(t, seed, n, 20_000, 1.0, 1_000, 150_000, 0.02, 0.3)
    out = []
    for i, p in enumerate(pieces):
        buf, chrs, bl = S.pack([("r%d" % i, p)])
        rc, _ = S.reverse_complement(buf, chrs, bl)
        raw = bool(raw_every) and i % raw_every == 0
        out.append((S.encode(buf[:bl]), S.encode(rc), buf[:bl], rc, raw))
    return out


def _oracle_lists(target, sub, index, pos, reqs):
    """The oracle's run_pass on every request alone: (lists, hash_pass)."""
    want = []
    for f, r, raw_f, raw_r, _ in reqs:
        o = O.Oracle(target, raw_f, raw_r, sub, max_hits=16711680, index=(index, pos))
        lists, _, h = o.run_pass(threads=16)
        want.append((lists, h))
    return want


def test_served_results_equal_alone(gpu_available):
    print("case serve: 40 requests (every third raw) through two engines vs each alone and the oracle",
          flush=True)
    target, sub, index, pos = _setup()
    reqs = _requests(target, 40, 3, raw_every=3)
    ref_enc = S.encode(target)
    e1 = E.Engine(index, pos, ref_enc, sub, max_hits=16711680)
    e2 = E.Engine(index, pos, ref_enc, sub, max_hits=16711680)
    alone = E.Engine(index, pos, ref_enc, sub, max_hits=16711680)
    want = []
    for f, r, _, _, _ in reqs:
        alone.swap_query(f, r)
        want.append(alone.run_pass()[0])
    oracle = _oracle_lists(target, sub, index, pos, reqs)
    with E.Server([e1, e2], max_batch_requests=16) as srv:
        ids = []
        # arrivals spread over time so that admission happens while passes run
        for i, (f, r, raw_f, _, raw) in enumerate(reqs):
            ids.append(srv.submit(raw_f, None) if raw else srv.submit(f, r))
            if i % 8 == 7:
                time.sleep(0.01)
        got = [srv.wait(i) for i in ids]
        st = srv.stats()
    for k, ((res, timing), w, (ow, oh)) in enumerate(zip(got, want, oracle)):
        assert same_intervals(res, w), "request %d differs from alone" % k
        assert same_intervals(res, ow) and E.hash_pass(res) == oh, "request %d differs from the oracle" % k
        assert timing["t_done"] >= timing["t_start"] >= timing["t_submit"] - 1e-9
    assert st["requests"] == len(reqs) and st["failed_batches"] == 0
    assert st["max_batch_requests"] > 1, "continuous batching merged no requests"
    for e in (e1, e2, alone):
        e.close()


def test_admission_refuses_request_over_budget(gpu_available):
    print("case serve: admission by HBM footprint", flush=True)
    target, sub, index, pos = _setup(100_000)
    eng = E.Engine(index, pos, S.encode(target), sub, max_hits=16711680)
    reqs = _requests(target, 2, 5)
    with E.Server([eng], hbm_budget=4 << 20) as srv:
        small = np.full(1000, 0, dtype=np.uint8)
        rid = srv.submit(small, small)
        res, _ = srv.wait(rid)
        # a collected result is handed out once: a second wait is an error, not a hang
        with pytest.raises(E.HspzError) as ew:
            srv.wait(rid)
        assert ew.value.status == 2
        with pytest.raises(E.HspzError) as ei:
            big = np.zeros(200_000, dtype=np.uint8)
            srv.submit(big, big)
        assert ei.value.status == 7
        assert srv.stats()["rejected"] == 1
    eng.close()


def test_served_growing_batches_equal_alone(gpu_available):
    print("case serve: growing batches (both hit orders, buffer growth during captures)", flush=True)
    target, sub, index, pos = _setup(600_000, seed=0x5E12)
    t = np.frombuffer(target, dtype=np.uint8)
    pieces = synth.sample_chunks(t, 9, 24, 150_000, 0.8, 20_000, 500_000, 0.02, 0.2)
    reqs, raws = [], []
    for i, p in enumerate(pieces):
        buf, chrs, bl = S.pack([("g%d" % i, p)])
        rc, _ = S.reverse_complement(buf, chrs, bl)
        reqs.append((S.encode(buf[:bl]), S.encode(rc)))
        raws.append((buf[:bl], rc))
    ref_enc = S.encode(target)
    e1 = E.Engine(index, pos, ref_enc, sub, max_hits=16711680)
    e2 = E.Engine(index, pos, ref_enc, sub, max_hits=16711680)
    alone = E.Engine(index, pos, ref_enc, sub, max_hits=16711680)
    want = []
    for f, r in reqs:
        alone.swap_query(f, r)
        want.append(alone.run_pass()[0])
    with E.Server([e1, e2], max_batch_requests=12) as srv:
        ids = []
        for i, (f, r) in enumerate(sorted(reqs, key=lambda x: len(x[0]))):
            ids.append((i, srv.submit(f, r)))
        order = sorted(range(len(reqs)), key=lambda k: len(reqs[k][0]))
        got = {order[i]: srv.wait(rid)[0] for i, rid in ids}
        st = srv.stats()
    for k, w in enumerate(want):
        assert same_intervals(got[k], w), "request %d differs from alone" % k
    for k, (raw_f, raw_r) in enumerate(raws[:8]):  # the oracle on a third of them (CPU time)
        o = O.Oracle(target, raw_f, raw_r, sub, max_hits=16711680, index=(index, pos))
        ow, _, oh = o.run_pass(threads=16)
        assert same_intervals(got[k], ow) and E.hash_pass(got[k]) == oh, "request %d differs from the oracle" % k
    assert st["failed_batches"] == 0
    assert max(e.get_flag("batch_order_used") for e in (e1, e2)) == 1 or st["max_batch_requests"] > 1
    for e in (e1, e2, alone):
        e.close()

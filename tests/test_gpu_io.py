"""Device encoding of raw query bytes: hspz_swap_query_raw, raw chunks of
hspz_filter_batch and raw requests of the serving loop give the same encoded
strands as the host's reverse complement + encode (sequence.rs:212-230 then
:66-79), byte for byte, and the same HSP lists as the encoded path (which the
oracle pins in the other GPU tests)."""
import sys

import numpy as np
import pytest

import oracle_ffi as O
from helpers import same_intervals
from hspz_amd import engine as E
from hspz_amd import seqio as X
from hspz_amd import sequence as S
from hspz_amd import synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _engine(target: bytes):
    sub = O.sub_mat("", 910)
    index, pos = O.seed_table(target, O.shape("12of19"), 1)
    return E.Engine(index, pos, S.encode(target), sub, max_hits=16711680)


def _host_strands(raw: bytes):
    buf, chrs, bl = S.pack([("q", raw)])
    rc, _ = S.reverse_complement(buf, chrs, bl)
    return S.encode(raw), S.encode(rc)


def _raw_cases(target: bytes):
    rng = np.random.default_rng(5)
    t = np.frombuffer(target, dtype=np.uint8)
    homolog = bytearray(t[20_000:60_000].tobytes())
    for i in rng.integers(0, len(homolog), 600):  # substitutions, soft masks, Ns, IUPAC, junk
        homolog[i] = int(rng.choice(list(b"ACGTacgtNnRYKMBDHVWS&-*Jj\xff\x80")))
    every = bytes(range(256)) * 3
    multi = X.SeqSet.pack([("a", target[1_000:17_384]), ("b", bytes(homolog[:5_000])), ("c", b"ACGTN")]).raw
    return [("empty", b""), ("one", b"G"), ("every byte", every), ("segment", target[:16384]),
            ("segment+1", target[5:16390]), ("homolog", bytes(homolog)), ("multi-record", multi)]


def test_swap_query_raw_equals_host_encoding(gpu_available):
    target = synth.random_bases(0x10E1, 200_000).tobytes()
    with _engine(target) as eng:
        for name, raw in _raw_cases(target):
            _say("case raw query: %s (%d bytes)" % (name, len(raw)))
            wf, wr = _host_strands(raw)
            eng.swap_query_raw(raw)
            f, r = eng.query()
            assert np.array_equal(f, wf) and np.array_equal(r, wr), name
            if len(raw) > 19:
                got, _ = eng.run_pass()
                eng.swap_query(wf, wr)
                want, _ = eng.run_pass()
                assert same_intervals(got, want), name


def test_batch_raw_and_mixed_chunks_equal_encoded(gpu_available):
    _say("case raw batch: raw, encoded and mixed chunks of one call")
    target = synth.random_bases(0x10E2, 300_000).tobytes()
    t = np.frombuffer(target, dtype=np.uint8)
    pieces = synth.sample_chunks(t, 21, 12, 20_000, 0.8, 2_000, 70_000, 0.02, 0.2)
    raws = [X.SeqSet.pack([("c%d" % i, p)]).raw for i, p in enumerate(pieces)]
    raws += [c for _, c in _raw_cases(target)]
    enc = [_host_strands(r) for r in raws]
    with _engine(target) as eng:
        want, wst = eng.filter_batch(enc)
        got, st = eng.filter_batch([(r, None) for r in raws])
        mixed, _ = eng.filter_batch([(r, None) if i % 2 else enc[i] for i, r in enumerate(raws)])
    assert st == wst
    for i in range(len(raws)):
        assert same_intervals(got[i], want[i]), "raw chunk %d" % i
        assert same_intervals(mixed[i], want[i]), "mixed chunk %d" % i


def test_served_raw_requests_equal_encoded(gpu_available):
    _say("case raw serve: raw and encoded submissions of the same requests")
    target = synth.random_bases(0x10E3, 300_000).tobytes()
    t = np.frombuffer(target, dtype=np.uint8)
    pieces = synth.sample_chunks(t, 22, 16, 10_000, 1.0, 1_000, 60_000, 0.02, 0.3)
    raws = [X.SeqSet.pack([("r%d" % i, p)]).raw for i, p in enumerate(pieces)]
    with _engine(target) as eng:
        want, _ = eng.filter_batch([_host_strands(r) for r in raws])
        with E.Server([eng]) as srv:
            ids = [srv.submit(r) if i % 2 else srv.submit(*_host_strands(r)) for i, r in enumerate(raws)]
            for i, rid in enumerate(ids):
                got, _ = srv.wait(rid)
                assert same_intervals(got, want[i]), "request %d" % i

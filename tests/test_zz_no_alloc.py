"""hspz_seed_and_filter_noalloc (the allocation-free plain entry point, over a
file-scope device scratch array) gives the same lists, hit counts and raw-HSP
counts as hspz_seed_and_filter on the same slot and strand, bit for bit, and
both equal the oracle's seed_and_filter. Each case runs two calls in a row,
the second on a smaller input, since the scratch array keeps the last call's
contents; the cases also cover many hit tiles per max_hits chunk, several
max_hits chunks, the byte-code gate (X-drop beyond the packed gate) and the
limit errors."""
import sys

import numpy as np
import pytest

from helpers import Pair
from hspz_amd import engine as E
from hspz_amd import synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _both(eng, o, raw_fwd, raw_rc, rev, lo, hi, cap=1 << 20):
    """Seeds of [lo, hi) on one strand, through both entry points and the oracle."""
    ns = eng.generate_seeds(0, rev, lo, hi)
    want, nh, raw = eng.seed_and_filter(0, rev)
    got, gnh, graw = eng.seed_and_filter_noalloc(0, rev, cap)
    assert (gnh, graw) == (nh, raw), "hit / raw counts differ"
    assert np.array_equal(got, want), "lists differ from hspz_seed_and_filter"
    seeds = E.chunk_seeds(raw_rc if rev else raw_fwd, lo, hi, threads=4)
    assert len(seeds) == ns
    ow, st = o.seed_and_filter(seeds, rev)
    assert np.array_equal(got, ow) and st["hits"] == gnh and st["raw_hsps"] == graw, "differs from the oracle"
    return len(got), gnh


def test_noalloc_equals_seed_and_filter(gpu_available):
    _say("case noalloc: C1-shape 600 kbp pair, a 250 kbp batch then a 40 kbp one, both strands")
    t, q = synth.c1_pair(600_000, seed=0xA110C)
    p = Pair.raw(t, q)
    o = p.oracle()
    with p.engine() as eng:
        for rev in (False, True):
            n1, h1 = _both(eng, o, p.q_raw, p.q_rc_raw, rev, 100_000, 350_001)
            n2, h2 = _both(eng, o, p.q_raw, p.q_rc_raw, rev, 20_000, 60_000)
            assert n1 > 0 and h2 < h1


def test_noalloc_tiles_and_chunks(gpu_available):
    _say("case noalloc: 4096-hit tiles, max_hits 20000 chunks, then a smaller call")
    t, q = synth.c1_pair(400_000, seed=0xA110D)
    # a homologous stretch gives long survivor runs across tile edges
    q = q[:150_000] + t[200_000:260_000] + q[210_000:]
    p = Pair.raw(t, q, max_hits=20000)
    o = p.oracle()
    with p.engine() as eng:
        eng.set_flag("noalloc_tile_hits", 4096)
        assert eng.get_flag("noalloc_tile_hits") == 4096
        n1, h1 = _both(eng, o, p.q_raw, p.q_rc_raw, False, 120_000, 300_000)
        assert h1 > 20000 * 3 and n1 > 0  # several max_hits chunks, many tiles each
        n2, h2 = _both(eng, o, p.q_raw, p.q_rc_raw, True, 0, 30_000)
        assert h2 < h1
        eng.set_flag("noalloc_tile_hits", 0)
        _both(eng, o, p.q_raw, p.q_rc_raw, False, 120_000, 300_000)


def test_noalloc_byte_code_gate(gpu_available):
    _say("case noalloc: X-drop 5000 (byte-code gate), noisy soft-masked pair")
    rng = np.random.default_rng(5)
    t0 = synth.random_bases(0xA110E, 300_000)
    q = bytearray(synth.mutate(t0, 6, 0.05, 0.005).tobytes())
    t = bytearray(t0.tobytes())
    for seq in (t, q):  # soft-masked stretches and N runs on both sides
        for a in rng.integers(0, len(seq) - 500, 40):
            n = int(rng.integers(20, 400))
            seq[a:a + n] = bytes(seq[a:a + n]).lower()
        for a in rng.integers(0, len(seq) - 100, 10):
            seq[a:a + 30] = b"N" * 30
    t, q = bytes(t), bytes(q)
    p = Pair.raw(t, q, xdrop=5000)
    o = p.oracle()
    with p.engine() as eng:
        _both(eng, o, p.q_raw, p.q_rc_raw, False, 0, 120_000)
        _both(eng, o, p.q_raw, p.q_rc_raw, True, 50_000, 90_000)


def test_noalloc_limits(gpu_available):
    _say("case noalloc: output capacity below the pre-dedup list, then a call that fits")
    t, q = synth.c1_pair(200_000, seed=0xA110F)
    q = q[:50_000] + t[60_000:90_000] + q[80_000:]
    p = Pair.raw(t, q)
    o = p.oracle()
    with p.engine() as eng:
        eng.generate_seeds(0, False, 40_000, 120_000)
        want, nh, raw = eng.seed_and_filter(0, False)
        assert raw > 2
        with pytest.raises(E.HspzError) as ei:
            eng.seed_and_filter_noalloc(0, False, 1)
        assert ei.value.status == 7 and "before dedup" in str(ei.value)
        got, gnh, graw = eng.seed_and_filter_noalloc(0, False, raw)
        assert np.array_equal(got, want) and (gnh, graw) == (nh, raw)
        _both(eng, o, p.q_raw, p.q_rc_raw, False, 40_000, 60_000)

"""GPU parity: the native engine (HIP kernels through the C ABI) against the
oracle, bit-exact on every HSP list (ref_start, query_start, len, score, in
order, per interval and strand)."""
import sys

import numpy as np
import pytest

import oracle_ffi as O
from helpers import Pair, same_intervals
from hspz_amd import engine as E
from hspz_amd import sequence as S
from hspz_amd import synth
from gen_fixtures import SCORING_ACGT

pytestmark = pytest.mark.gpu

FIXTURES = ["identical", "mismatch", "indel", "revcomp", "repeat", "ambiguous", "xdrop", "multi",
            "softmask", "scoring"]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check(pair, threads=16, **kw):
    want, wst, wh = pair.oracle().run_pass(threads=threads, **kw)
    with pair.engine() as eng:
        got, gst = eng.run_pass(**kw)
    assert gst["hits"] == wst["hits"], (gst, wst)
    assert gst["raw_hsps"] == wst["raw_hsps"], (gst, wst)
    assert same_intervals(got, want)
    assert E.hash_pass(got) == wh
    return got, gst


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_parity(gpu_available, name):
    _say("case fixture %s" % name)
    kw = {"acgt": SCORING_ACGT} if name == "scoring" else {}
    p = Pair.fixture(name, **kw)
    got, _ = _check(p)
    if name == "repeat":
        assert p.segments_digest(got) == (57, 1, "d01edd2118cf5fa5")
    if name == "multi":
        assert p.segments_digest(got) == (3, 2, "0da3bbda656e4f7b")


def test_repeat_with_small_max_hits(gpu_available):
    _say("case repeat max_hits=1000 (chunked)")
    _check(Pair.fixture("repeat", max_hits=1000))


@pytest.mark.parametrize("xdrop", [910, 5000])
def test_c1_shape_small_multi_interval(gpu_available, xdrop):
    # X = 5000 is beyond the packed gate: the plain path runs the byte-code gate
    _say("case c1-shape 400 kbp, 3 intervals, 100 kbp chunks, X=%d" % xdrop)
    t, q = synth.c1_pair(400_000, seed=0xC0FFEE01)
    _check(Pair.raw(t, q, xdrop=xdrop), wga_chunk=100_000, interval_size=150_000)


@pytest.mark.parametrize("ambiguous,noentropy", [("", False), ("n", False), ("iupac,5,7", True)])
def test_noisy_sequences(gpu_available, ambiguous, noentropy):
    _say("case noisy ambiguous=%r noentropy=%s" % (ambiguous, noentropy))
    rng = np.random.default_rng(17)
    t = bytearray(synth.random_bases(5, 200_000).tobytes())
    q = bytearray(synth.mutate(np.frombuffer(bytes(t), np.uint8), 6, 0.08, 0.005).tobytes())
    for arr in (t, q):
        for i in rng.integers(0, len(arr), len(arr) // 100):
            arr[i] = rng.choice(list(b"NnacgtRYKMX"))
    _check(Pair.raw(bytes(t), bytes(q), ambiguous=ambiguous, noentropy=noentropy))


def test_soft_masked_repeats(gpu_available):
    _say("case c2-shape 1 Mbp with soft-masked repeats")
    t, q = synth.c2_pair(1_000_000, seed=0xC0FFEE02)
    _check(Pair.raw(t, q))


def test_upload_seeds_equals_generate_seeds(gpu_available):
    _say("case host seeds vs device seeds")
    t, q = synth.c1_pair(200_000, seed=77)
    p = Pair.raw(t, q)
    o = p.oracle()
    with p.engine() as eng:
        for rev, (lo, hi) in [(False, (0, 60_000)), (True, (10_000, 90_000))]:
            raw = p.q_rc_raw if rev else p.q_raw
            seeds = E.chunk_seeds(raw, lo, hi, threads=4)
            eng.upload_seeds(1, seeds)
            a, nha, rawa = eng.seed_and_filter(1, rev)
            ns = eng.generate_seeds(0, rev, lo, hi)
            assert ns == len(seeds)
            b, nhb, rawb = eng.seed_and_filter(0, rev)
            want, st = o.seed_and_filter(seeds, rev)
            assert nha == nhb == st["hits"] and rawa == rawb == st["raw_hsps"]
            assert np.array_equal(a, want) and np.array_equal(b, want)


def test_error_paths(gpu_available):
    _say("case error paths")
    p = Pair.fixture("identical")
    o = p.oracle()
    eng = E.Engine(o.index, o.pos, o.ref_enc, p.sub, max_hits=16711680)
    with pytest.raises(E.HspzError) as ei:
        eng.seed_and_filter(0, False)
    assert ei.value.status == 1
    with pytest.raises(E.HspzError) as ei:  # a pass before any swap_query
        eng.run_pass()
    assert ei.value.status == 1
    empty = np.zeros(0, dtype=np.uint8)
    eng.swap_query(empty, empty)  # an empty query: seed_and_filter refuses it, a pass is empty
    with pytest.raises(E.HspzError) as ei:
        eng.seed_and_filter(0, False)
    assert ei.value.status == 1
    assert eng.run_pass()[0] == []
    eng.swap_query(o.q_enc, o.q_rc_enc)
    with pytest.raises(E.HspzError) as ei:
        eng.generate_seeds(0, False, 0, len(o.q_enc))
    assert ei.value.status == 4
    with pytest.raises(E.HspzError) as ei:
        eng.upload_seeds(0, np.array([1 << 56], dtype=np.uint64))
    assert ei.value.status == 2
    assert eng.stat("max_hits") == 16711680
    eng.close()


def _big_score_pair(thresh):
    """Six identical segments of 170..333 bp between unrelated flanks, under
    matches of about 100,000 and mismatches of -90,001: HSP totals of 17 M to
    33 M, above 2^24, where float32 spacing is 2 or 4."""
    m = [100003, 99997, 99991, 100007]
    acgt = [m[i] if i == j else -90001 for i in range(4) for j in range(4)]
    ref = bytearray(synth.random_bases(0x2425, 8000).tobytes())
    qry = bytearray(synth.random_bases(0x2426, 8000).tobytes())
    for k, ln in enumerate((170, 171, 190, 230, 260, 333)):
        a, b = 500 + 1200 * k, 700 + 1200 * k
        qry[b:b + ln] = ref[a:a + ln]
    return Pair.raw(bytes(ref), bytes(qry), acgt=acgt, xdrop=1_000_000, hspthresh=thresh, noentropy=True)


def test_scores_above_2_24_keep_the_float32_rounding(gpu_available):
    """kernels.rs:2036-2041: the gate and the materializer compare the total
    through float32. Above 2^24 that rounds: 23,120,075 rounds up to the
    threshold 23,120,076 and is kept; 17,330,177 rounds down below a threshold
    of its own value and is dropped. The matrix is beyond the packed gate's
    limits, so the plain path runs the byte-code gate."""
    _say("case scores above 2^24 (byte-code gate, float32 rounding)")
    for thresh, want_scores in ((23_120_076, [23120075, 26010094, 33330021]),
                                (17_330_177, [19130053, 23120075, 26010094, 33330021])):
        p = _big_score_pair(thresh)
        got, _ = _check(p)
        assert sorted(int(x) for fw, rc in got for a in (fw, rc) for x in a["score"]) == want_scores


@pytest.mark.parametrize("seed,transitions,plus,minus", [
    ("12of19", False, True, True), ("12of19", True, True, False), ("12of19", True, False, True),
    ("1110110111", True, True, True), ("14of22", True, True, True)])
def test_seed_shapes_strands_and_transitions(gpu_available, seed, transitions, plus, minus):
    """The pass options of seed_and_filter_all (run.rs:274-300) and the seed
    shapes of Shape::parse (seed.rs:47-84), on the plain and the batched path."""
    _say("case seed %s transitions=%s plus=%s minus=%s" % (seed, transitions, plus, minus))
    t, q = synth.c1_pair(200_000, seed=0xC0FFEE07)
    p = Pair.raw(t, q, seed=seed)
    kw = dict(transitions=transitions, plus=plus, minus=minus, wga_chunk=60_000, interval_size=90_000)
    want, wst, wh = p.oracle().run_pass(threads=16, **kw)
    assert wst["hsps"] > 0
    with p.engine() as eng:
        got, gst = eng.run_pass(pattern=seed, **kw)
        assert gst["hits"] == wst["hits"] and gst["raw_hsps"] == wst["raw_hsps"], (gst, wst)
        assert same_intervals(got, want)
        assert E.hash_pass(got) == wh
        o = p.oracle()
        for order in (0, 1):  # seed-major and k-mer-major hit order
            eng.set_flag("batch_order", order)
            batched, _ = eng.filter_batch([(o.q_enc, o.q_rc_enc)], pattern=seed, **kw)
            assert eng.get_flag("batch_order_used") == order
            assert same_intervals(batched[0], want), "batched differs in order %d" % order


def test_max_hits_underflow_and_oversize_seed(gpu_available):
    """chunk_limits (gpu/mod.rs:2397-2420): a first seed that alone holds more
    than max_hits hits is the reference's underflow error, on the plain and
    the batched path; a later such seed only makes an oversize chunk, and the
    lists equal the oracle's."""
    _say("case max_hits=60: first seed over the limit (error), later seed over it (oversize chunk)")
    t = bytearray(synth.random_bases(0x51, 50_000).tobytes())
    t[1000:1100] = b"A" * 100
    t = bytes(t)
    # the poly-A k-mer has 82 positions; a batch stays within the 64 chunks the batched path takes
    first = Pair.raw(t, b"A" * 40 + synth.random_bases(0x52, 1000).tobytes(), max_hits=60)
    with pytest.raises(RuntimeError):
        first.oracle().run_pass(threads=4)
    o = first.oracle()
    with first.engine() as eng:
        with pytest.raises(E.HspzError) as ei:
            eng.run_pass()
        assert ei.value.status == 5
        with pytest.raises(E.HspzError) as ei:
            eng.filter_batch([(o.q_enc, o.q_rc_enc)])
        assert ei.value.status == 5 and "single seed exceeds max_hits" in str(ei.value)
    later = Pair.raw(t, synth.random_bases(0x53, 1000).tobytes() + b"A" * 40 + t[20000:21000], max_hits=60)
    want, _ = _check(later)
    o = later.oracle()
    with later.engine() as eng:
        got, _ = eng.filter_batch([(o.q_enc, o.q_rc_enc)])
        assert same_intervals(got[0], want)


@pytest.mark.parametrize("xdrop,thresh", [(0, 100), (1, 300), (4000, 0), (4000, 1)])
def test_extreme_xdrop_and_threshold(gpu_available, xdrop, thresh):
    """X-drop at 0 and 1 (a side stops at its first loss) and thresholds of 0
    and 1 (every hit that scores at all survives the gate), plain and batched."""
    _say("case xdrop=%d thresh=%d" % (xdrop, thresh))
    t = synth.random_bases(0xE5, 60_000).tobytes()
    q = synth.mutate(np.frombuffer(t[10_000:40_000], dtype=np.uint8), 0xE6, 0.12, 0.004).tobytes()
    p = Pair.raw(t, q, xdrop=xdrop, hspthresh=thresh)
    want, st = _check(p)
    assert st["hsps"] > 0
    o = p.oracle()
    with p.engine() as eng:
        for order in (0, 1):
            eng.set_flag("batch_order", order)
            got, _ = eng.filter_batch([(o.q_enc, o.q_rc_enc)])
            assert same_intervals(got[0], want), "batched differs in order %d" % order


def test_long_homolog_extensions(gpu_available):
    """A 20 kbp homolog at 0.5% substitutions: most hits extend thousands of
    cells past the gate's windows, through the deferred-hit gates' stretch
    loop to the sequence ends, on the plain and the batched path."""
    _say("case 20 kbp homolog, 0.5% substitutions (long extensions)")
    t = synth.random_bases(0xE7, 80_000).tobytes()
    q = synth.mutate(np.frombuffer(t[30_000:50_000], dtype=np.uint8), 0xE8, 0.005, 0.0).tobytes()
    p = Pair.raw(t, q)
    want, st = _check(p)
    assert max(int(a["len"].max()) for fw, rc in want for a in (fw, rc) if len(a)) > 5_000
    o = p.oracle()
    with p.engine() as eng:
        for order in (0, 1):
            eng.set_flag("batch_order", order)
            got, _ = eng.filter_batch([(o.q_enc, o.q_rc_enc)])
            assert eng.stat("gate_deferred") > 0
            assert same_intervals(got[0], want), "batched differs in order %d" % order


def test_homologs_at_sequence_edges(gpu_available):
    """Homologs that touch the first and last bases of both sequences: their
    hits sit within the gate's margins of an end and take the exact paths."""
    _say("case homologs at the sequence edges")
    t = synth.random_bases(0xE9, 30_000).tobytes()
    mid = synth.random_bases(0xEA, 2_000).tobytes()
    q = t[-600:] + mid + t[:600]
    for qq in (q, t[:700] + mid + t[-700:]):
        p = Pair.raw(t, qq)
        want, st = _check(p)
        assert st["hsps"] >= 2
        o = p.oracle()
        with p.engine() as eng:
            for order in (0, 1):
                eng.set_flag("batch_order", order)
                got, _ = eng.filter_batch([(o.q_enc, o.q_rc_enc)])
                assert same_intervals(got[0], want), "batched differs in order %d" % order


def test_c1_full_size(gpu_available):
    _say("case c1 full size: 4 Mbp x 4 Mbp, plain and batched")
    t, q = synth.c1_pair(4_000_000, seed=0xC0FFEE01)
    pair = Pair.raw(t, q)
    got, st = _check(pair)
    assert st["hsps"] > 0
    # the batched path on the same query as one block: every hit order and merged step
    with pair.engine() as eng:
        block = [(S.encode(pair.q_raw), S.encode(pair.q_rc_raw))]
        for step, threads, order in [(96, 1024, 2), (128, 1024, 1), (96, 512, 1), (64, 512, 0)]:
            _say("  batched step %d threads %d order %d" % (step, threads, order))
            for name, v in (("batch_step_cells", step), ("batch_block_threads", threads), ("batch_order", order)):
                eng.set_flag(name, v)
            bgot, bst = eng.filter_batch(block)
            assert bst["hits"] == st["hits"] and bst["raw_hsps"] == st["raw_hsps"], (bst, st)
            assert same_intervals(bgot[0], got)


@pytest.mark.parametrize("case", ["random", "softmasked_n", "step3", "tiny", "1110110111", "14of22"])
def test_device_index_equals_host(gpu_available, case):
    """hspz_engine_new_device_index builds the same index as the host
    SeedTable::build (seed.rs:126-373) and the oracle, and the engine it makes
    gives the same HSP lists as one built from the host index."""
    print("case device index: %s" % case, flush=True)
    step = 3 if case == "step3" else 1
    shape = case if case in ("1110110111", "14of22") else "12of19"
    if case == "tiny":
        target = b"AAAAACGTAAAAACGTTTTTTTTTTTTTTTTTTTTTTTGGGGA"
    elif case == "softmasked_n":
        t, _ = synth.c2_pair(300_000, seed=0xD1)
        t = bytearray(t)
        rng = np.random.default_rng(5)
        for i in rng.integers(0, len(t) - 50, 300):
            t[i:i + rng.integers(1, 40)] = b"N" * len(t[i:i + rng.integers(1, 40)])
        target = bytes(t)
    else:
        target = synth.random_bases(0xD2, 500_000 if shape == "12of19" else 200_000).tobytes()
    sub = O.sub_mat("", 910)
    h_index, h_pos = E.seed_table_build(target, shape, step, 8)
    o_index, o_pos = O.seed_table(target, O.shape(shape), step)
    assert np.array_equal(h_index, o_index) and np.array_equal(h_pos, o_pos)
    with E.Engine(None, None, S.encode(target), sub, seed_size=E.shape_parse(shape)[0], max_hits=16711680,
                  device_index=shape, index_step=step) as eng:
        d_index, d_pos = eng.index()
        assert eng.stat("pos_len") == len(h_pos)
        assert np.array_equal(d_index, h_index), "index_table differs"
        assert np.array_equal(d_pos, h_pos), "pos_table differs"
        if case == "random":
            q = synth.mutate(np.frombuffer(target[100_000:200_000], dtype=np.uint8), 7, 0.1, 0.005)
            buf, chrs, bl = S.pack([("q", bytes(q))])
            rc, _ = S.reverse_complement(buf, chrs, bl)
            fw, rv = S.encode(buf[:bl]), S.encode(rc)
            eng.swap_query(fw, rv)
            got, _ = eng.run_pass()
            with E.Engine(h_index, h_pos, S.encode(target), sub, max_hits=16711680) as ref_eng:
                ref_eng.swap_query(fw, rv)
                want, _ = ref_eng.run_pass()
            assert same_intervals(got, want)

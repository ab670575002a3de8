"""GPU parity of the batched path: every chunk's HSP lists from one
hspz_filter_batch call equal the oracle's run_pass on that chunk alone (and
the plain GPU path's), bit for bit."""
import sys

import numpy as np
import pytest

import oracle_ffi as O
from helpers import same_intervals
from hspz_amd import engine as E
from hspz_amd import sequence as S
from hspz_amd import synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _block(records):
    buf, chrs, bl = S.pack(records)
    raw = buf[:bl]
    rc, _ = S.reverse_complement(buf, chrs, bl)
    return raw, rc


def _run(target: bytes, query_blocks, max_hits=16711680, ambiguous="", noentropy=False,
         wga_chunk=250000, interval_size=10000000, hspthresh=3000):
    sub = O.sub_mat(ambiguous, 910)
    index, pos = O.seed_table(target, O.shape("12of19"), 1)
    ref_enc = S.encode(target)
    with E.Engine(index, pos, ref_enc, sub, max_hits=max_hits, noentropy=noentropy, hspthresh=hspthresh) as eng:
        chunks = [(S.encode(f), S.encode(r)) for f, r in query_blocks]
        got, st = eng.filter_batch(chunks, wga_chunk=wga_chunk, interval_size=interval_size)
        for order in (0, 1):  # both hit orders of the batched gate give the same lists
            eng.set_flag("batch_order", order)
            again, _ = eng.filter_batch(chunks, wga_chunk=wga_chunk, interval_size=interval_size)
            assert eng.get_flag("batch_order_used") == order
            for i, (a, g) in enumerate(zip(again, got)):
                assert same_intervals(a, g), "chunk %d differs in hit order %d" % (i, order)
        # k-mer-major order without the coarse bound ahead of the two-cell maps, with
        # it in its paired forms (HOXD70's class scores fit 16 bits) and single-hit form,
        # the candidates' windows read from the planes and from the records
        eng.set_flag("batch_order", 1)
        for coarse, used, planes in ((0, 0, 1), (1, 3, 1), (3, 2, 1), (2, 1, 1), (1, 3, 0), (2, 1, 0)):
            eng.set_flag("batch_coarse", coarse)
            eng.set_flag("batch_cand_planes", planes)
            again, _ = eng.filter_batch(chunks, wga_chunk=wga_chunk, interval_size=interval_size)
            assert eng.get_flag("batch_coarse_used") == used
            # candidates binned by reference position whenever their windows come from the planes
            assert eng.get_flag("batch_cand_binned") == (1 if coarse and planes else 0)
            for i, (a, g) in enumerate(zip(again, got)):
                assert same_intervals(a, g), "chunk %d differs with coarse mode %d, planes %d" % (i, coarse, planes)
        eng.set_flag("batch_coarse", 1)
        eng.set_flag("batch_cand_planes", 1)
        # the planes' candidates in the coarse gate's order, unbinned
        eng.set_flag("batch_cand_bins", 0)
        again, _ = eng.filter_batch(chunks, wga_chunk=wga_chunk, interval_size=interval_size)
        assert eng.get_flag("batch_cand_binned") == 0
        for i, (a, g) in enumerate(zip(again, got)):
            assert same_intervals(a, g), "chunk %d differs with unbinned candidates" % i
        with pytest.raises(E.HspzError):
            eng.set_flag("batch_cand_bins", 3)
        eng.set_flag("batch_cand_bins", 1)
        # every candidate through the two-cell maps, none sent straight to the deferred-hit gate:
        # the same lists, from at least as many candidates
        routed = eng.stat("batch_candidates")
        eng.set_flag("batch_coarse_long", 0)
        again, _ = eng.filter_batch(chunks, wga_chunk=wga_chunk, interval_size=interval_size)
        assert eng.stat("batch_candidates") >= routed
        for i, (a, g) in enumerate(zip(again, got)):
            assert same_intervals(a, g), "chunk %d differs without the long-gate routing" % i
        eng.set_flag("batch_coarse_long", 1)
        # the materializer one wave per survivor throughout, instead of a lane per survivor first
        eng.set_flag("batch_lane_materialize", 0)
        again, _ = eng.filter_batch(chunks, wga_chunk=wga_chunk, interval_size=interval_size)
        for i, (a, g) in enumerate(zip(again, got)):
            assert same_intervals(a, g), "chunk %d differs with the wave materializer" % i
        eng.set_flag("batch_lane_materialize", 1)
        eng.set_flag("batch_order", 2)
        # the call split into device passes between batches gives the same lists:
        # every batch alone (a batch is never split, so 1 position per pass runs each
        # batch in a pass of its own), then runs of batches (a third of the positions)
        npos = sum(2 * max(len(f) - 19, 0) for f, _ in chunks)
        for pmax in (1, max(npos // 3, 1)):
            eng.set_flag("batch_max_positions", pmax)
            again, _ = eng.filter_batch(chunks, wga_chunk=wga_chunk, interval_size=interval_size)
            if pmax == 1:
                assert eng.stat("batch_passes") == st["batches"]
            elif st["batches"] > 3:
                assert eng.stat("batch_passes") > 1
            for i, (a, g) in enumerate(zip(again, got)):
                assert same_intervals(a, g), "chunk %d differs with %d positions per device pass" % (i, pmax)
        eng.set_flag("batch_max_positions", 0)
        assert eng.stat("batch_passes") >= 1
        eng.set_flag("batch_device_dedup", 0)  # host dedup_and_order gives the same lists
        again, _ = eng.filter_batch(chunks, wga_chunk=wga_chunk, interval_size=interval_size)
        for i, (a, g) in enumerate(zip(again, got)):
            assert same_intervals(a, g), "chunk %d differs between device and host dedup" % i
        eng.set_flag("batch_device_dedup", 1)
        plain = []
        for f, r in chunks:
            eng.swap_query(f, r)
            plain.append(eng.run_pass(wga_chunk=wga_chunk, interval_size=interval_size)[0])
    for i, (f, r) in enumerate(query_blocks):
        o = O.Oracle(target, f, r, sub, max_hits=max_hits, noentropy=noentropy, hspthresh=hspthresh,
                     index=(index, pos))
        want, wst, wh = o.run_pass(threads=16, wga_chunk=wga_chunk, interval_size=interval_size)
        assert same_intervals(got[i], want), "chunk %d differs from the oracle" % i
        assert same_intervals(plain[i], want), "plain chunk %d differs from the oracle" % i
        assert E.hash_pass(got[i]) == wh
    return got, st


def _chunks_from(target: bytes, n, seed, lo=2_000, hi=60_000, div=(0.02, 0.2)):
    t = np.frombuffer(target, dtype=np.uint8)
    pieces = synth.sample_chunks(t, seed, n, (lo * hi) ** 0.5, 0.8, lo, hi, div[0], div[1])
    return [_block([("c%d" % i, p)]) for i, p in enumerate(pieces)]


def test_batch_many_uneven_chunks(gpu_available):
    _say("case batch: 24 uneven chunks vs a 600 kbp target")
    target = synth.random_bases(0xB001, 600_000).tobytes()
    blocks = _chunks_from(target, 24, 11)
    # a multi-record block, a tiny block shorter than the seed and an empty-ish one
    blocks.append(_block([("m1", target[1000:3000]), ("m2", target[5000:9000])]))
    blocks.append(_block([("tiny", b"ACGTACGTAC")]))
    got, st = _run(target, blocks)
    assert st["hsps"] > 0 and st["batches"] > 0


def test_batch_multi_interval_and_small_max_hits(gpu_available):
    _say("case batch: multi-interval blocks, max_hits=2000 (chunked scopes)")
    target = synth.random_bases(0xB002, 300_000).tobytes()
    blocks = _chunks_from(target, 6, 12, lo=50_000, hi=120_000, div=(0.01, 0.05))
    _run(target, blocks, max_hits=2000, wga_chunk=20_000, interval_size=45_000)


@pytest.mark.parametrize("ambiguous", ["n", "iupac,5,7"])
def test_batch_noisy_and_soft_masked(gpu_available, ambiguous):
    _say("case batch: soft-masked repeats, N runs, IUPAC codes, ambiguous=%r" % ambiguous)
    t, q = synth.c2_pair(400_000, seed=0xB003)
    rng = np.random.default_rng(3)
    q = bytearray(q)
    for i in rng.integers(0, len(q), len(q) // 200):
        q[i] = rng.choice(list(b"NnRYX"))
    q = bytes(q)
    blocks = [_block([("a", q[:150_000])]), _block([("b", q[150_000:260_000]), ("c", q[260_000:])])]
    _run(t, blocks, ambiguous=ambiguous)


def test_batch_separator_that_does_not_force_a_drop(gpu_available):
    _say("case batch: separator scored -900 (no forced drop): fenced windows go to the exact gate")
    target = synth.random_bases(0xB00A, 300_000).tobytes()
    blocks = _chunks_from(target, 10, 0x1A, lo=5_000, hi=40_000, div=(0.01, 0.1))
    blocks.append(_block([("m1", target[1000:3000]), ("m2", target[2990:9000])]))  # '&' inside a homolog
    sub = O.sub_mat("", 910).copy()
    sub[56:64] = -900  # row of the separator code
    sub[7::8] = -900   # its column
    index, pos = O.seed_table(target, O.shape("12of19"), 1)
    want = []
    for f, r in blocks:
        o = O.Oracle(target, f, r, sub, max_hits=16711680, index=(index, pos))
        want.append(o.run_pass(threads=16)[0])
    with E.Engine(index, pos, S.encode(target), sub, max_hits=16711680) as eng:
        chunks = [(S.encode(f), S.encode(r)) for f, r in blocks]
        for order in (0, 1):
            eng.set_flag("batch_order", order)
            got, st = eng.filter_batch(chunks)
            for i, (g, w) in enumerate(zip(got, want)):
                assert same_intervals(g, w), "chunk %d differs in order %d" % (i, order)
        eng.swap_query(*chunks[-1])
        plain, _ = eng.run_pass()
        assert same_intervals(plain, want[-1])


def test_batch_repeat_fixture_digest(gpu_available):
    _say("case batch: the repeat and multi fixtures as two chunks of one call")
    from helpers import Pair
    for name, want in (("repeat", (57, 1, "d01edd2118cf5fa5")), ("multi", (3, 2, "0da3bbda656e4f7b"))):
        p = Pair.fixture(name)
        o = p.oracle()
        with E.Engine(o.index, o.pos, o.ref_enc, p.sub, max_hits=16711680) as eng:
            got, _ = eng.filter_batch([(o.q_enc, o.q_rc_enc), (o.q_enc, o.q_rc_enc)])
        assert p.segments_digest(got[0]) == want and p.segments_digest(got[1]) == want


def test_batch_graph_replays_match_eager(gpu_available):
    _say("case batch: hipGraph capture once, replays x3 with eager work between, vs eager")
    target = synth.random_bases(0xB004, 300_000).tobytes()
    blocks = _chunks_from(target, 10, 21)
    sub = O.sub_mat("", 910)
    index, pos = O.seed_table(target, O.shape("12of19"), 1)
    chunks = [(S.encode(f), S.encode(r)) for f, r in blocks]
    with E.Engine(index, pos, S.encode(target), sub, max_hits=16711680) as eng:
        eng.set_flag("batch_graph", 0)
        want, _ = eng.filter_batch(chunks)
        eng.set_flag("batch_graph", 1)
        c0, r0 = eng.stat("graph_captures"), eng.stat("graph_replays")
        for k in range(3):
            got, _ = eng.filter_batch(chunks)
            for g, w in zip(got, want):
                assert same_intervals(g, w), "replay %d differs from eager" % k
            eng.swap_query(*chunks[k])          # eager work between replays
            eng.run_pass()
        assert eng.stat("graph_captures") - c0 == 1
        assert eng.stat("graph_replays") - r0 == 2


def _engine_and_chunks(seed, n_chunks, target_len=300_000):
    target = synth.random_bases(seed, target_len).tobytes()
    blocks = _chunks_from(target, n_chunks, seed & 0xFF)
    sub = O.sub_mat("", 910)
    index, pos = O.seed_table(target, O.shape("12of19"), 1)
    chunks = [(S.encode(f), S.encode(r)) for f, r in blocks]
    return target, blocks, sub, index, pos, chunks


def _low_complexity_case():
    """Poly-A, dinucleotide and trinucleotide runs in a 120 kbp target, and
    query blocks around them with 3% substitutions, plus pure runs."""
    rng = np.random.default_rng(0xB00D)
    t = bytearray(synth.random_bases(0xB00C, 120_000).tobytes())
    for at, unit, n in ((10_000, b"A", 400), (40_000, b"AC", 200), (70_000, b"AAT", 100), (95_000, b"T", 350)):
        run = unit * n
        t[at:at + len(run)] = run
    target = bytes(t)
    blocks = []
    for i, (a, b) in enumerate(((9_000, 12_000), (39_500, 41_000), (69_000, 72_000), (94_000, 97_000))):
        q = bytearray(target[a:b])
        for j in rng.integers(0, len(q), len(q) // 30):  # 3% substitutions outside the runs too
            q[j] = int(rng.choice(list(b"ACGT")))
        blocks.append(_block([("lc%d" % i, bytes(q))]))
    blocks.append(_block([("polyA", b"A" * 600), ("AC", b"AC" * 300)]))
    return target, blocks


def test_batch_low_complexity_mass_deferral(gpu_available):
    """Poly-A and dinucleotide runs in both sequences: whole k-mer groups of
    hits outlive the gate's windows, so waves defer up to 64 hits at once and
    their LDS queues flush before and after pushes; every gate configuration
    gives the oracle's lists."""
    _say("case batch: low-complexity runs (mass deferral) vs the oracle")
    target, blocks = _low_complexity_case()
    sub = O.sub_mat("", 910)
    index, pos = O.seed_table(target, O.shape("12of19"), 1)
    want = []
    for f, r in blocks:
        o = O.Oracle(target, f, r, sub, max_hits=16711680, index=(index, pos))
        want.append(o.run_pass(threads=16)[0])
    chunks = [(S.encode(f), S.encode(r)) for f, r in blocks]
    with E.Engine(index, pos, S.encode(target), sub, max_hits=16711680) as eng:
        for step, threads, claim, order, ahead in [(96, 1024, 64, 1, 1), (128, 1024, 64, 1, 1), (128, 512, 256, 1, 1),
                                                   (96, 512, 256, 1, 1), (64, 1024, 16, 1, 1), (128, 512, 256, 0, 1)]:
            _say("  step %d threads %d claim %d order %d lookahead %d" % (step, threads, claim, order, ahead))
            for name, v in (("batch_step_cells", step), ("batch_block_threads", threads), ("batch_claim_seeds", claim),
                            ("batch_order", order), ("batch_lookahead", ahead), ("batch_gate_blocks", 0)):
                eng.set_flag(name, v)
            got, st = eng.filter_batch(chunks)
            assert eng.stat("gate_deferred") > 0
            for i, (g, w) in enumerate(zip(got, want)):
                assert same_intervals(g, w), "chunk %d differs at step %d threads %d" % (i, step, threads)
        for i in range(2):  # the plain path on the same blocks
            eng.swap_query(*chunks[i])
            g, _ = eng.run_pass()
            assert same_intervals(g, want[i]), "plain chunk %d differs" % i


def _entropy_band_case():
    """Short low-complexity runs between unrelated flanks, so the HSPs are the
    runs and their shifted-diagonal overlaps, with totals of roughly 1,500 to
    6,000: at a threshold of 1500 most land inside [T, 3T], where the
    materializer rescales them by the entropy of their matched bases."""
    rng = np.random.default_rng(0xE7)
    runs = [b"A" * 24, b"A" * 40, b"AC" * 14, b"AC" * 30, b"ACGT" * 8, b"ACGT" * 12, b"ACACG" * 9,
            b"AAT" * 15, b"TTTG" * 10, b"GC" * 20]
    t = bytearray(synth.random_bases(0xE8, 40_000).tobytes())
    blocks = []
    for i, run in enumerate(runs):
        at = 1_000 + 3_500 * i
        t[at:at + len(run)] = run
        q = bytearray(synth.random_bases(0xE9 + i, 600).tobytes())
        copy = bytearray(run)
        if i % 3 == 2:  # a few substitutions inside some copies
            for j in rng.integers(2, len(copy) - 2, 2):
                copy[j] = int(rng.choice(list(b"ACGT")))
        q[280:280 + len(copy)] = copy
        blocks.append(_block([("run%d" % i, bytes(q))]))
    return bytes(t), blocks


@pytest.mark.parametrize("noentropy", [False, True])
def test_batch_entropy_band(gpu_available, noentropy):
    """The entropy branch where it decides: batched (both hit orders, device
    and host dedup) and plain paths give the oracle's lists with and without
    the correction, and the correction changes those lists."""
    _say("case batch: low-complexity runs in the entropy band, noentropy=%s" % noentropy)
    target, blocks = _entropy_band_case()
    got, _ = _run(target, blocks, noentropy=noentropy, hspthresh=1500)
    sub = O.sub_mat("", 910)
    index, pos = O.seed_table(target, O.shape("12of19"), 1)
    other = [O.Oracle(target, f, r, sub, noentropy=not noentropy, hspthresh=1500,
                      index=(index, pos)).run_pass(threads=16)[0] for f, r in blocks]
    assert sum(len(fw) + len(rc) for w in got for fw, rc in w) > 20
    assert sum(not same_intervals(g, o) for g, o in zip(got, other)) >= 5


def test_batch_list_overflow_reruns(gpu_available):
    """Survivor and deferred-hit lists started at 64 records overflow (writes
    past the end are dropped and flagged); the pass doubles them and runs
    again, and every gate gives the oracle's lists."""
    _say("case batch: lists started at 64 records overflow and the pass reruns")
    target, blocks = _low_complexity_case()
    sub = O.sub_mat("", 910)
    index, pos = O.seed_table(target, O.shape("12of19"), 1)
    want = [O.Oracle(target, f, r, sub, max_hits=16711680, index=(index, pos)).run_pass(threads=16)[0]
            for f, r in blocks]
    chunks = [(S.encode(f), S.encode(r)) for f, r in blocks]
    with E.Engine(index, pos, S.encode(target), sub, max_hits=16711680) as eng:
        eng.set_flag("batch_list_cap", 64)
        assert eng.get_flag("batch_list_cap") == 64
        for step, threads, claim, order, ahead in [(96, 1024, 64, 1, 0), (128, 512, 256, 1, 1), (64, 512, 16, 0, 0)]:
            _say("  step %d threads %d claim %d order %d lookahead %d" % (step, threads, claim, order, ahead))
            for name, v in (("batch_step_cells", step), ("batch_block_threads", threads), ("batch_claim_seeds", claim),
                            ("batch_order", order), ("batch_lookahead", ahead)):
                eng.set_flag(name, v)
            r0 = eng.stat("batch_reruns")
            got, _ = eng.filter_batch(chunks)
            assert eng.stat("batch_reruns") > r0
            assert eng.stat("gate_deferred") > 64
            for i, (g, w) in enumerate(zip(got, want)):
                assert same_intervals(g, w), "chunk %d differs after reruns at step %d" % (i, step)
        eng.set_flag("batch_list_cap", 0)
        r0 = eng.stat("batch_reruns")
        got, _ = eng.filter_batch(chunks)
        assert eng.stat("batch_reruns") == r0
        for g, w in zip(got, want):
            assert same_intervals(g, w)
        with pytest.raises(E.HspzError):
            eng.set_flag("batch_list_cap", 63)


def test_batch_empty_and_hitless_chunks(gpu_available):
    """Chunks that give no batches or no seeds beside ordinary ones: an empty
    block, one shorter than the seed, all N, all soft-masked, and an unrelated
    random block; an empty call returns no chunks."""
    _say("case batch: empty, tiny, all-N, all-masked and unrelated chunks beside homologs")
    target = synth.random_bases(0xB0E0, 200_000).tobytes()
    blocks = _chunks_from(target, 3, 0xE0, lo=5_000, hi=30_000)
    blocks += [_block([("empty", b"")]), _block([("tiny", b"ACGTAC")]), _block([("n", b"N" * 5_000)]),
               _block([("masked", target[50_000:53_000].lower())]),
               _block([("unrelated", synth.random_bases(0xB0E1, 30_000).tobytes())])]
    blocks.append(_chunks_from(target, 1, 0xE1, lo=5_000, hi=30_000)[0])
    got, st = _run(target, blocks)
    assert all(len(w) == 0 for w in got[3:5])  # no intervals: no batches
    assert all(sum(len(fw) + len(rc) for fw, rc in w) == 0 for w in got[5:8])
    assert sum(len(fw) + len(rc) for fw, rc in got[0]) > 0
    sub = O.sub_mat("", 910)
    index, pos = O.seed_table(target, O.shape("12of19"), 1)
    with E.Engine(index, pos, S.encode(target), sub, max_hits=16711680) as eng:
        res, _ = eng.filter_batch([])
        assert res == []
        res, _ = eng.filter_batch([(S.encode(f), S.encode(r)) for f, r in blocks[3:6]])
        assert [len(w) for w in res] == [0, 0, 1]


def test_batch_packed_gate_limits(gpu_available):
    """xdrop 4000 and ACGT scores up to 4000 (HOXD70 x 40), the largest the
    packed gates take: a word's 16 cells then reach +-64 000, so each of the
    gate's int16 partial sums must cover at most 8 cells. Batched (both hit
    orders, merged and narrow steps) and plain lists equal the oracle's."""
    _say("case batch: packed-gate limits (xdrop 4000, scores x40)")
    hoxd70 = [91, -114, -31, -123, -114, 100, -125, -31, -31, -125, 100, -114, -123, -31, -114, 91]
    acgt = [40 * v for v in hoxd70]
    xdrop, thresh = 4000, 120000
    target = synth.random_bases(0xB00E, 150_000).tobytes()
    blocks = _chunks_from(target, 10, 17, lo=300, hi=12_000, div=(0.0, 0.15))
    blocks.append(_block([("same", target[20_000:21_000]), ("near", target[60_000:60_400])]))
    sub = O.sub_mat("", xdrop, acgt)
    index, pos = O.seed_table(target, O.shape("12of19"), 1)
    want, hsps = [], 0
    for f, r in blocks:
        o = O.Oracle(target, f, r, sub, xdrop=xdrop, hspthresh=thresh, max_hits=16711680, index=(index, pos))
        res, st, _ = o.run_pass(threads=16)
        want.append(res)
        hsps += st["hsps"]
    assert hsps > 100
    chunks = [(S.encode(f), S.encode(r)) for f, r in blocks]
    with E.Engine(index, pos, S.encode(target), sub, xdrop=xdrop, hspthresh=thresh, max_hits=16711680) as eng:
        for step, threads, order in [(96, 1024, 1), (128, 512, 1), (64, 1024, 1), (128, 512, 0), (32, 512, 1)]:
            _say("  step %d threads %d order %d" % (step, threads, order))
            for name, v in (("batch_step_cells", step), ("batch_block_threads", threads), ("batch_order", order)):
                eng.set_flag(name, v)
            got, _ = eng.filter_batch(chunks)
            for i, (g, w) in enumerate(zip(got, want)):
                assert same_intervals(g, w), "chunk %d differs at step %d threads %d order %d" % (i, step, threads, order)
        for i in range(3):
            eng.swap_query(*chunks[i])
            g, _ = eng.run_pass()
            assert same_intervals(g, want[i]), "plain chunk %d differs" % i


def test_batch_gate_configs_match(gpu_available):
    _say("case batch: gate step/block/claim/grid variants vs the oracle")
    target, blocks, sub, index, pos, chunks = _engine_and_chunks(0xB005, 12)
    want = []
    for f, r in blocks:
        o = O.Oracle(target, f, r, sub, max_hits=16711680, index=(index, pos))
        want.append(o.run_pass(threads=16)[0])
    # (step, threads, claim, grid, order, lookahead): 128-cell steps in 512-thread blocks run the
    # lookahead gate unless it is switched off
    variants = [(16, 1024, 1, 0, 0, 1), (64, 512, 4096, 0, 0, 1), (32, 1024, 7, 3, 0, 1), (64, 1024, 1, 1, 0, 1),
                (16, 512, 64, 5000, 0, 1), (64, 1024, 1, 0, 1, 1), (32, 512, 3, 2, 1, 1), (16, 1024, 4096, 0, 1, 1),
                (64, 512, 64, 7, 1, 1), (32, 1024, 1, 5000, 1, 1), (128, 1024, 1, 0, 1, 1), (128, 512, 256, 3, 1, 0),
                (128, 1024, 64, 0, 0, 1), (128, 512, 256, 0, 1, 1), (128, 512, 1, 5000, 1, 1),
                (128, 512, 4096, 1, 1, 1), (128, 512, 16, 7, 1, 1), (96, 512, 256, 0, 1, 1),
                (96, 512, 1, 3, 1, 1), (96, 1024, 64, 0, 1, 1), (96, 512, 64, 0, 0, 1), (96, 512, 256, 0, 1, 0),
                (96, 1024, 1, 5000, 1, 1), (96, 1024, 4096, 3, 1, 1)]
    with E.Engine(index, pos, S.encode(target), sub, max_hits=16711680) as eng:
        for step, threads, claim, blocks_n, order, ahead in variants:
            _say("  step %d threads %d claim %d grid %d order %d lookahead %d"
                 % (step, threads, claim, blocks_n, order, ahead))
            eng.set_flag("batch_lookahead", ahead)
            eng.set_flag("batch_order", order)
            eng.set_flag("batch_step_cells", step)
            eng.set_flag("batch_block_threads", threads)
            eng.set_flag("batch_claim_seeds", claim)
            eng.set_flag("batch_gate_blocks", blocks_n)
            got, st = eng.filter_batch(chunks)
            assert eng.get_flag("batch_order_used") == order
            assert eng.get_flag("batch_step_cells") == step
            if blocks_n:
                assert eng.get_flag("batch_gate_blocks") == blocks_n
            else:
                assert eng.get_flag("batch_gate_blocks") >= 256
            if step >= 96 and order == 1:
                assert eng.get_flag("ref_records") == 2  # a merged-step gate builds the records
            for i, (g, w) in enumerate(zip(got, want)):
                assert same_intervals(g, w), "chunk %d differs at step %d claim %d" % (i, step, claim)
        # merged steps with their reference windows read from the plane (no records)
        eng.set_flag("ref_records", -1)
        assert eng.get_flag("ref_records") == -1
        for step, threads, claim in [(128, 1024, 64), (96, 1024, 64), (96, 512, 256), (128, 512, 1)]:
            _say("  step %d threads %d claim %d without reference records" % (step, threads, claim))
            eng.set_flag("batch_order", 1)
            eng.set_flag("batch_step_cells", step)
            eng.set_flag("batch_block_threads", threads)
            eng.set_flag("batch_claim_seeds", claim)
            eng.set_flag("batch_gate_blocks", 0)
            got, _ = eng.filter_batch(chunks)
            assert eng.get_flag("ref_records") == -1
            for i, (g, w) in enumerate(zip(got, want)):
                assert same_intervals(g, w), "chunk %d differs at step %d (plane windows)" % (i, step)
        eng.set_flag("ref_records", 0)
        for name, bad in [("batch_step_cells", 24), ("batch_step_cells", 48), ("batch_block_threads", 256),
                          ("batch_claim_seeds", 0), ("batch_gate_blocks", -1), ("batch_order", 3),
                          ("batch_order_used", 1), ("ref_records", 2), ("ref_records", -2)]:
            with pytest.raises(E.HspzError) as ei:
                eng.set_flag(name, bad)
            assert ei.value.status == 2


def test_batch_autotune(gpu_available):
    _say("case batch: autotuner grid, chosen cell, output unchanged")
    target, blocks, sub, index, pos, chunks = _engine_and_chunks(0xB006, 8)
    with E.Engine(index, pos, S.encode(target), sub, max_hits=16711680) as eng:
        before, st0 = eng.filter_batch(chunks)
        step0 = eng.get_flag("batch_step_cells")
        cells, best = eng.autotune(chunks, reps=1)
        assert all(c["status"] == 0 for c in cells)
        assert len({c["hits"] for c in cells}) == 1 and cells[0]["hits"] == st0["hits"]
        assert [c["order"] for c in cells[:2]] == [0, 1]
        assert len({c["order"] for c in cells[2:]}) == 1
        if cells[2]["order"] == 1:  # behind the coarse bound: blocks x larger claims at the engine's step
            assert len(cells) == 12
            assert {(c["step_cells"], c["block_threads"], c["claim_seeds"]) for c in cells[2:8]} == {
                (step0, t, c) for t in (512, 1024) for c in (64, 256, 1024)}
        else:
            assert len(cells) == 36
            assert {(c["step_cells"], c["block_threads"], c["claim_seeds"]) for c in cells[2:32]} == {
                (s, t, c) for s in (16, 32, 64, 96, 128) for t in (512, 1024) for c in (16, 64, 256)}
        fastest = min(c["gate_ms"] + c["after_ms"] for c in cells)
        assert best["gate_ms"] + best["after_ms"] == fastest
        assert eng.get_flag("batch_step_cells") == best["step_cells"]
        assert eng.get_flag("batch_block_threads") == best["block_threads"]
        assert eng.get_flag("batch_claim_seeds") == best["claim_seeds"]
        assert eng.get_flag("batch_order") == best["order"]
        after, st1 = eng.filter_batch(chunks)
        assert eng.get_flag("batch_gate_blocks") == best["grid_blocks"]
        for g, w in zip(after, before):
            assert same_intervals(g, w)
        assert st1["hits"] == st0["hits"]


@pytest.mark.parametrize("scale,xdrop", [(1, 910), (2, 240), (4, 1800), (16, 3000)])
def test_batch_score_sum_forms(gpu_available, scale, xdrop):
    """The merged gates' score sums in each form the matrix selects, against
    the oracle: HOXD70 (scores within +-255: raw 32-bit sums of the table
    entries, read as int16 at the decision); x2 at xdrop 240 (scores to -250
    clamped at -241, so raw sums again); x4 at xdrop 1800 (scores within 2047:
    16-cell int16 trees); x16 at xdrop 3000 (beyond 2047: 8-cell trees). At
    96- and 128-cell steps in both block sizes, on homologs long enough to
    take the continuation iterations (right past 32 cells, left past 64)."""
    _say("case batch: score-sum form, HOXD70 x%d, xdrop %d" % (scale, xdrop))
    hoxd70 = [91, -114, -31, -123, -114, 100, -125, -31, -31, -125, 100, -114, -123, -31, -114, 91]
    acgt = [scale * v for v in hoxd70]
    thresh = 3000 * scale
    target = synth.random_bases(0xB5C0 + scale, 200_000).tobytes()
    blocks = _chunks_from(target, 12, 23 + scale, lo=500, hi=15_000, div=(0.0, 0.2))
    sub = O.sub_mat("", xdrop, acgt)
    index, pos = O.seed_table(target, O.shape("12of19"), 1)
    want, hsps = [], 0
    for f, r in blocks:
        o = O.Oracle(target, f, r, sub, xdrop=xdrop, hspthresh=thresh, max_hits=16711680, index=(index, pos))
        res, st, _ = o.run_pass(threads=16)
        want.append(res)
        hsps += st["hsps"]
    assert hsps > 10
    chunks = [(S.encode(f), S.encode(r)) for f, r in blocks]
    with E.Engine(index, pos, S.encode(target), sub, xdrop=xdrop, hspthresh=thresh, max_hits=16711680) as eng:
        for step, threads in [(96, 1024), (128, 512), (96, 512), (128, 1024)]:
            for name, v in (("batch_step_cells", step), ("batch_block_threads", threads), ("batch_order", 1),
                            ("batch_lookahead", 0)):
                eng.set_flag(name, v)
            got, _ = eng.filter_batch(chunks)
            for i, (g, w) in enumerate(zip(got, want)):
                assert same_intervals(g, w), "chunk %d differs at step %d threads %d" % (i, step, threads)


@pytest.mark.parametrize("name,acgt,xdrop,used", [
    ("hoxd70", [91, -114, -31, -123, -114, 100, -125, -31, -31, -125, 100, -114, -123, -31, -114, 91], 910, 3),
    # a transversion pair scoring +20: the transversion class adds to the positive part
    ("positive transversion", [91, 20, -31, -123, 20, 100, -125, -31, -31, -125, 100, -114, -123, -31, -114, 91],
     910, 3),
    # class bests 101 / -31 / -112: the count coefficients -294 and 81 share the factor 3, which
    # 16 x 101 lacks, so no count biases exist and the paired sums carry their constant
    ("no bias pair", [91, -112, -31, -123, -112, 101, -125, -31, -31, -125, 101, -112, -123, -31, -112, 91], 910, 3),
    # x2: class scores beyond 127, so the single-hit form even when the paired one is asked for
    ("hoxd70 x2", [2 * v for v in [91, -114, -31, -123, -114, 100, -125, -31, -31, -125, 100, -114, -123, -31,
                                   -114, 91]], 1820, 1),
])
def test_batch_coarse_forms(gpu_available, name, acgt, xdrop, used):
    """The coarse bound's forms (two hits per lane in packed 16-bit sums, over
    two reference positions or two query records; one hit per lane in 32-bit
    sums) and no bound, against the oracle, on matrices that select each form
    and both positive-part variants."""
    _say("case batch: coarse bound forms, %s" % name)
    thresh = 3000 * (2 if "x2" in name else 1)
    target = synth.random_bases(0xC0A5 + len(name), 200_000).tobytes()
    blocks = _chunks_from(target, 10, 0xC0 + len(name), lo=500, hi=15_000, div=(0.0, 0.2))
    sub = O.sub_mat("", xdrop, acgt)
    index, pos = O.seed_table(target, O.shape("12of19"), 1)
    want, hsps = [], 0
    for f, r in blocks:
        o = O.Oracle(target, f, r, sub, xdrop=xdrop, hspthresh=thresh, max_hits=16711680, index=(index, pos))
        res, st, _ = o.run_pass(threads=16)
        want.append(res)
        hsps += st["hsps"]
    assert hsps > 10
    chunks = [(S.encode(f), S.encode(r)) for f, r in blocks]
    with E.Engine(index, pos, S.encode(target), sub, xdrop=xdrop, hspthresh=thresh, max_hits=16711680) as eng:
        eng.set_flag("batch_order", 1)
        for threads in (1024, 512):
            eng.set_flag("batch_block_threads", threads)
            for coarse, u, planes in ((1, used, 1), (1, used, 0), (3, 2 if used == 3 else 1, 1), (2, 1, 1),
                                      (2, 1, 0), (0, 0, 1)):
                eng.set_flag("batch_coarse", coarse)
                eng.set_flag("batch_cand_planes", planes)
                d0 = eng.stat("gate_deferred")
                got, st = eng.filter_batch(chunks)
                assert eng.get_flag("batch_coarse_used") == u
                assert eng.get_flag("batch_cand_binned") == (1 if coarse and planes else 0)
                if coarse:  # homologous chunks: the bound keeps many hits (to the two-cell maps or,
                    # sides outliving the windows, the deferred-hit gate), never more than there are
                    assert eng.stat("batch_candidates") <= st["hits"]
                    assert eng.stat("batch_candidates") + eng.stat("gate_deferred") - d0 > 0
                for i, (g, w) in enumerate(zip(got, want)):
                    assert same_intervals(g, w), "chunk %d differs in coarse mode %d, planes %d, %d threads" % (
                        i, coarse, planes, threads)
        # the planes' candidates unbinned, then without the long-gate routing
        eng.set_flag("batch_coarse", 1)
        eng.set_flag("batch_cand_planes", 1)
        eng.set_flag("batch_cand_bins", 0)
        got, _ = eng.filter_batch(chunks)
        assert eng.get_flag("batch_cand_binned") == 0
        for i, (g, w) in enumerate(zip(got, want)):
            assert same_intervals(g, w), "chunk %d differs with unbinned candidates" % i
        eng.set_flag("batch_cand_bins", 1)
        routed = eng.stat("batch_candidates")
        eng.set_flag("batch_coarse_long", 0)
        got, _ = eng.filter_batch(chunks)
        # homologous chunks: in the two-reference-positions form some candidates went straight to
        # the deferred-hit gate (the other forms send none there)
        if used == 3:
            assert eng.stat("batch_candidates") > routed
        else:
            assert eng.stat("batch_candidates") == routed
        for i, (g, w) in enumerate(zip(got, want)):
            assert same_intervals(g, w), "chunk %d differs without the long-gate routing" % i


def test_batch_staged_read(gpu_available):
    """The device-deduplicated lists read through the pinned staging halves in
    many pieces (a 4 KiB piece for the test; 64 MiB in use) equal the direct
    read and the oracle."""
    _say("case batch: output read through staged pieces")
    target = synth.random_bases(0x57A6, 300_000).tobytes()
    blocks = _chunks_from(target, 24, 0x57A6, lo=5_000, hi=60_000, div=(0.0, 0.1))
    sub = O.sub_mat("", 910)
    index, pos = O.seed_table(target, O.shape("12of19"), 1)
    want = []
    for f, r in blocks:
        o = O.Oracle(target, f, r, sub, max_hits=16711680, index=(index, pos))
        want.append(o.run_pass(threads=16)[0])
    chunks = [(S.encode(f), S.encode(r)) for f, r in blocks]
    with E.Engine(index, pos, S.encode(target), sub, max_hits=16711680) as eng:
        direct, st = eng.filter_batch(chunks)
        assert st["hsps"] > 1000  # at least four 4 KiB pieces of 256 records
        assert E.lib().hspz_test_set_read_piece(eng._h, 4096) == 0
        for pmax in (0, 1):  # one device pass, and a pass per batch
            eng.set_flag("batch_max_positions", pmax)
            staged, _ = eng.filter_batch(chunks)
            for i, (a, b, w) in enumerate(zip(staged, direct, want)):
                assert same_intervals(a, b) and same_intervals(a, w), "chunk %d (passes %d)" % (i, pmax)

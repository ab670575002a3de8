"""Pins the CPU oracle: the reference's own known-answer tests, its frozen
end-to-end digests, closed forms, and an independent model of the reference's
32-lane tiled X-drop scan (kernels.rs:1279-1543) that the oracle's scalar loop
must agree with."""
import math

import numpy as np
import pytest

import oracle_ffi as O
from helpers import Pair

NUC = 8


# ------------------------------------------------------------- scoring.rs:295-322

def test_default_sub_mat_kats():
    m = O.sub_mat("", 910)
    assert m[0] == 91 and m[3] == -123 and m[2 * NUC + 2] == 100
    assert m[4 * NUC + 4] == -1000 and m[5 * NUC + 5] == -1000
    assert m[6 * NUC + 6] == -100
    assert m[7 * NUC + 7] == -9100 and m[7] == -9100


def test_ambiguous_n_and_iupac_kats():
    m = O.sub_mat("n", 910)
    assert m[5 * NUC + 5] == 0 and m[5] == 0 and m[6 * NUC + 6] == -100
    m = O.sub_mat("iupac,5,7", 910)
    assert m[5 * NUC + 5] == 5 and m[5] == -7 and m[6 * NUC + 6] == 5


def test_custom_acgt_block_lands_in_the_matrix():
    blk = [100 if i == j else -50 for i in range(4) for j in range(4)]
    m = O.sub_mat("", 910, blk)
    assert m[0] == 100 and m[1] == -50 and m[4 * NUC + 4] == -1000


# ------------------------------------------------------------- seed.rs:503-537

def test_shape_12of19():
    s = O.shape("12of19")
    assert s.size == 19 and s.kmer_size == 12
    assert list(s.pos[:12]) == [0, 1, 2, 4, 7, 8, 11, 13, 15, 16, 17, 18]


def test_shape_floor_and_kmer_packing():
    with pytest.raises(ValueError):
        O.shape("101")
    s = O.shape("1111")
    assert O.kmer_at(s, b"ACGT", 0) == 0b00_01_10_11
    assert O.kmer_at(s, b"ACG&", 0) == 1 << 31
    spaced = O.shape("11011")
    assert O.kmer_at(spaced, b"AC&GT", 0) == 1 << 31
    assert O.kmer_at(spaced, b"ACAGT", 0) == 0b00_01_10_11


def test_transition_seeds_flip_one_bit():
    s = O.shape("1111")
    out = O.chunk_seeds(b"NNNNNNNAAAA", s, True, 7, 8)
    assert len(out) == 5
    assert out[0] == 7
    assert out[1] >> 32 == 2 and out[4] >> 32 == 2 << 6
    assert all(int(o) & 0xFFFFFFFF == 7 for o in out)


def test_chunk_seeds_doc_example():
    s = O.shape("12of19")
    seq = b"ACGTACGTACGTACGTACGTACGT"
    assert len(O.chunk_seeds(seq, s, False, 0, 24 - 19 + 1)) == 24 - 19 + 1


# ------------------------------------------------------------- seed.rs:601-623

def test_seed_table_skips_position_zero():
    s = O.shape("1111")
    index, pos = O.seed_table(b"AAAAACGTA", s, 1)
    assert len(pos) == 5
    assert list(pos[:index[0]]) == [1]
    assert index[-1] == 5


@pytest.mark.parametrize("seq", [b"", b"ACGT", b"AAAAACGTA"])
def test_seed_table_tiny_and_empty(seq):
    index, pos = O.seed_table(seq, O.shape("12of19"), 1)
    assert len(pos) == 0 and index[-1] == 0


def test_seed_table_matches_brute_force_buckets():
    rng = np.random.default_rng(7)
    seq = bytearray(rng.choice(list(b"ACGT"), 5000).astype(np.uint8).tobytes())
    for i in rng.integers(0, 5000, 40):
        seq[i] = ord("N") if i % 2 else ord("a")
    seq = bytes(seq)
    s = O.shape("1111")
    for step in (1, 3):
        index, pos = O.seed_table(seq, s, step)
        offset = (s.size + 1) % step
        start = step - offset
        want = {}
        p = start
        while p + s.size <= len(seq):
            k = O.kmer_at(s, seq, p)
            if k != 1 << 31:
                want.setdefault(k, []).append(p)
            p += step
        for k in range(256):
            lo = index[k - 1] if k else 0
            assert list(pos[lo:index[k]]) == want.get(k, [])


# ------------------------------------------------------------- chunks / intervals

def test_chunks_kat():
    assert O.chunks(0, 376657, 250000) == [(0, 250000), (250000, 376658)]
    assert O.chunks(0, 10, 250000) == [(0, 11)]


def test_intervals_kat():
    assert O.intervals(376676, 19, 10_000_000) == [(0, 376657)]
    assert O.intervals(1019, 19, 500) == [(0, 500), (500, 1000)]


# ------------------------------------------------------------- gpu/mod.rs:2645-2669

def test_chunk_limits_kats():
    assert O.chunk_limits([2, 5, 9], 1000) == [(0, 2, 0, 9)]
    ch = O.chunk_limits([3, 7, 11, 15], 5)
    assert len(ch) == 4 and ch[0] == (0, 0, 0, 3)
    for a, b in zip(ch, ch[1:]):
        assert b[0] == a[1] + 1
    assert ch[-1][1] == 3


def test_chunk_limits_underflow_is_an_error():
    with pytest.raises(ValueError):
        O.chunk_limits([10], 5)


# ------------------------------------------------------------- hsp.rs:308-368

def test_dedup_kats():
    # one survivor per diagonal; the shortest sorts first and is kept
    got = O.dedup_and_order([(0, 0, 10, 50), (0, 0, 5, 40), (6, 6, 2, 30), (0, 1, 4, 20)])
    assert len(got) == 2 and (0, 1, 4, 20) in got and (0, 0, 5, 40) in got
    # compares against the previous INPUT element, not the last kept one
    assert O.dedup_and_order([(0, 0, 5, 10), (0, 0, 10, 20), (6, 6, 2, 30)]) == [(0, 0, 5, 10)]


def test_lastz_order_is_query_major():
    got = O.dedup_and_order([(9, 5, 1, 1), (1, 2, 1, 1), (0, 5, 1, 1)])
    assert got == [(1, 2, 1, 1), (0, 5, 1, 1), (9, 5, 1, 1)]


def test_diagonal_wraps_like_u32():
    # (10,4) has diagonal 6 and sorts before (4,10) whose diagonal wraps to 2^32-6;
    # they are on different diagonals, so both survive.
    got = O.dedup_and_order([(4, 10, 0, 0), (10, 4, 0, 0)])
    assert sorted(got) == [(4, 10, 0, 0), (10, 4, 0, 0)]


# ------------------------------------------------------------- gate / ln

def test_gate_rounding_kat():
    assert O.gate_keep(1400 + 1600, 3000)
    assert not O.gate_keep(2999, 3000)
    assert 16_777_219 < 16_777_220 and O.gate_keep(16_777_219, 16_777_220)


def test_reference_ln_matches_libm_within_its_truncation_bound():
    # The series stops at z^7/15, so |ln_ref - ln| <= 2|s| z^8 / (17 (1 - z)) plus
    # rounding. Dropping or mis-weighting any kept term breaks this bound by
    # more than an order of magnitude.
    rng = np.random.default_rng(3)
    xs = list(rng.uniform(1e-6, 1.0, 4000)) + [1.0, 0.5, 0.25, 1 / 3, 0.999999, 0.7071, 2e-300]
    for x in xs:
        x = float(x)
        m, e = math.frexp(x)
        m *= 2.0
        if m > math.sqrt(2.0):
            m *= 0.5
        s = (m - 1.0) / (m + 1.0)
        z = s * s
        bound = 2 * abs(s) * z ** 8 / (17 * (1 - z)) + 8e-16 * max(1.0, abs(math.log(x)))
        assert abs(O.ln(x) - math.log(x)) <= bound, x


# ------------------------------------------------------------- extension


def tiled_right(ref, qry, sub, X, R0, Q0, W=32):
    """The reference's tiled right extension (kernels.rs:1279-1407): per tile
    of W lanes an inclusive add-scan, a max-scan seeded with the inherited
    maximum, the first lane whose max - score exceeds X, and position recovery
    from the earliest lane reaching the maximum. Out-of-range lanes add 0."""
    tile, prev_score, prev_max, prev_pos = 0, 0, 0, -1
    while True:
        offs = np.arange(W) + tile
        sc = np.zeros(W, dtype=np.int64)
        ok = (R0 + offs < len(ref)) & (Q0 + offs < len(qry))
        idx = offs[ok]
        sc[ok] = sub[ref[R0 + idx].astype(np.int64) * 8 + qry[Q0 + idx]]
        ts = np.cumsum(sc) + prev_score
        cand = np.maximum(ts, prev_max)
        scanned = np.maximum.accumulate(cand)
        dropped = np.nonzero(scanned - ts > X)[0]
        first = dropped[0] if len(dropped) else W
        mx = prev_max if first == 0 else scanned[first - 1]
        if mx == prev_max:
            mpos = prev_pos
        else:
            win = np.nonzero((ts == mx) & (np.arange(W) < first))[0]
            mpos = int(win[0]) + tile
        edge = (R0 + tile + W - 1 >= len(ref)) or (Q0 + tile + W - 1 >= len(qry))
        if len(dropped) or edge:
            return int(mx), int(mpos)
        prev_score, prev_max, prev_pos = int(ts[-1]), int(mx), mpos
        tile += W


def tiled_left(ref, qry, sub, X, R0, Q0, W=32):
    """kernels.rs:1417-1543: lanes at offsets tile+1 .. tile+W to the left."""
    tile, prev_score, prev_max, prev_pos = 0, 0, 0, 0
    while True:
        offs = np.arange(W) + 1 + tile
        sc = np.zeros(W, dtype=np.int64)
        ok = (R0 >= offs) & (Q0 >= offs)
        idx = offs[ok]
        sc[ok] = sub[ref[R0 - idx].astype(np.int64) * 8 + qry[Q0 - idx]]
        ts = np.cumsum(sc) + prev_score
        cand = np.maximum(ts, prev_max)
        scanned = np.maximum.accumulate(cand)
        dropped = np.nonzero(scanned - ts > X)[0]
        first = dropped[0] if len(dropped) else W
        mx = prev_max if first == 0 else scanned[first - 1]
        if mx == prev_max:
            mpos = prev_pos
        else:
            win = np.nonzero((ts == mx) & (np.arange(W) < first))[0]
            mpos = int(win[0]) + 1 + tile
        last = tile + W
        edge = R0 < last or Q0 < last
        if len(dropped) or edge:
            return int(mx), int(mpos)
        prev_score, prev_max, prev_pos = int(ts[-1]), int(mx), mpos
        tile += W


def _codes(rng, n, p_special=0.0):
    a = rng.choice(4, n).astype(np.uint8)
    if p_special:
        m = rng.random(n) < p_special
        a[m] = rng.choice([4, 5, 6, 7], int(m.sum()))
    return a


@pytest.mark.parametrize("W", [32, 64])
def test_scalar_extension_equals_reference_tiles(W):
    rng = np.random.default_rng(0x243F6A88)
    sub = O.sub_mat("", 910).astype(np.int64)
    subi = O.sub_mat("", 910)
    for trial in range(400):
        n = int(rng.integers(20, 400))
        ref = _codes(rng, n, 0.01 if trial % 3 == 0 else 0.0)
        # similar query so extensions run long
        qry = ref.copy()
        flip = rng.random(n) < rng.uniform(0.0, 0.4)
        qry[flip] = rng.choice(4, int(flip.sum()))
        X = int(rng.choice([0, 50, 300, 910, 5000]))
        R0 = int(rng.integers(0, n + 1))
        Q0 = int(min(n, max(0, R0 + rng.integers(-3, 4))))
        rs, re, ls, le, _ = O.extend(ref, qry, subi, X, R0, Q0)
        assert (rs, re) == tiled_right(ref, qry, sub, X, R0, Q0, W), (trial, R0, Q0, X)
        assert (ls, le) == tiled_left(ref, qry, sub, X, R0, Q0, W), (trial, R0, Q0, X)


def test_extension_closed_forms():
    rng = np.random.default_rng(11)
    sub = O.sub_mat("", 910)
    ref = _codes(rng, 300)
    # identical sequences: every cell scores positive, no drop, runs to the end
    rs, re, ls, le, _ = O.extend(ref, ref.copy(), sub, 910, 100, 100)
    diag = sub[ref.astype(np.int64) * 8 + ref]
    assert rs == int(diag[100:].sum()) and re == 199
    assert ls == int(diag[:100].sum()) and le == 100
    # infinite X: right score is the maximum prefix sum (earliest argmax)
    qry = _codes(rng, 300)
    rs, re, *_ = O.extend(ref, qry, sub, 10**9, 0, 0)
    pref = np.cumsum(sub[ref.astype(np.int64) * 8 + qry])
    best = max(0, int(pref.max()))
    assert rs == best
    assert re == (int(np.argmax(pref)) if best > 0 else -1)
    # a separator forces a drop at its cell: nothing beyond it is reached
    q2 = ref.copy()
    q2[150] = 7
    rs, re, *_ = O.extend(ref, q2, sub, 910, 100, 100)
    assert re == 49 and rs == int(diag[100:150].sum())


def _materialize(ref, qry, thresh, noentropy=False, r0=None):
    r, q = np.array(ref, np.uint8), np.array(qry, np.uint8)
    r0 = len(r) // 2 if r0 is None else r0
    return O.materialize(r, q, O.sub_mat("", 910), 910, thresh, noentropy, r0, r0)


def test_entropy_closed_forms():
    """The materializer's entropy branch (kernels.rs:1556-1702) on self-hits
    whose base composition is fixed by construction. The score is rescaled by
    the Shannon entropy of the matched-base composition over the extent's
    positions, in units of ln 4, only for totals inside [T, 3T] with at least
    20 matched bases; the divisor is float32(ln 4) widened (gpu/mod.rs:2349-2352)."""
    A, Cc, G = 0, 1, 2
    # one letter: p = 1, entropy 0, so an in-band HSP is dropped; 30 x A/A = 2730
    assert _materialize([A] * 30, [A] * 30, 910) is None          # 2730 = 3T: in band
    assert _materialize([A] * 30, [A] * 30, 909) == (0, 0, 29, 2730)  # above 3T: untouched
    assert _materialize([A] * 30, [A] * 30, 910, noentropy=True) == (0, 0, 29, 2730)
    # the matched-base count must reach 20 for the entropy to apply
    assert _materialize([A] * 19, [A] * 19, 800) == (0, 0, 18, 1729)
    assert _materialize([A] * 20, [A] * 20, 800) is None
    # two letters, half each: ln(1/2) = -ln 2 exactly, so entropy = ln 2 / float32(ln 4),
    # a hair below 1/2: the even total 16 x (91 + 100) = 3056 scores 1527, not 1528,
    # and a threshold of 1528 drops it (a double-precision ln 4 would keep it)
    ac = [A, Cc] * 16
    assert _materialize(ac, ac, 1100) == (0, 0, 31, 1527)
    assert _materialize(ac, ac, 1527) == (0, 0, 31, 1527)
    assert _materialize(ac, ac, 1528) is None
    # four letters, a quarter each: ln(1/4) = -2 ln 2 exactly, entropy = 2 ln 2 /
    # float32(ln 4) = 1 - 2.7e-9 since float32(ln 4) rounds up: 3056 scores 3055
    acgt = [0, 1, 2, 3] * 8
    total = 8 * (91 + 100 + 100 + 91)
    assert _materialize(acgt, acgt, 1100) == (0, 0, 31, total - 1)
    assert _materialize(acgt, acgt, 1100, noentropy=True) == (0, 0, 31, total)
    # mismatched positions count in the denominator (extent + 1) but not in the
    # counts: 18 A, 18 C and 8 G/A transitions over 44 positions, hit at 20
    ref = [A, Cc, A, Cc, G] * 8 + [A, Cc, A, Cc]
    qry = [A, Cc, A, Cc, A] * 8 + [A, Cc, A, Cc]
    total = 18 * 91 + 18 * 100 - 8 * 31
    p = 18 / 44
    h = -2 * p * math.log(p) / float(np.float32(math.log(4)))
    assert abs(total * h - round(total * h)) > 1e-3  # far from a rounding edge
    got = _materialize(ref, qry, 1200, r0=20)
    assert got == (0, 0, 43, int(total * h)) == (0, 0, 43, 1682)
    for wrong in (36, 43):  # matched count or extent as the denominator
        pw = 18 / wrong
        assert int(total * (-2 * pw * math.log(pw) / math.log(4))) != got[3]


# ------------------------------------------------------------- end to end

def test_frozen_digest_repeat():
    p = Pair.fixture("repeat")
    res, st, _ = p.oracle().run_pass()
    assert p.segments_digest(res) == (57, 1, "d01edd2118cf5fa5")


def test_frozen_digest_multi():
    p = Pair.fixture("multi")
    res, st, _ = p.oracle().run_pass()
    assert p.segments_digest(res) == (3, 2, "0da3bbda656e4f7b")


def test_run_pass_is_thread_invariant():
    from hspz_amd import synth
    t, q = synth.c1_pair(300_000)
    p = Pair.raw(t, q)
    a, sa, ha = p.oracle().run_pass(threads=1, wga_chunk=50_000, interval_size=120_000)
    b, sb, hb = p.oracle().run_pass(threads=4, wga_chunk=50_000, interval_size=120_000)
    assert ha == hb and sa == sb
    assert sum(len(f) + len(r) for f, r in a) > 0


def test_max_hits_changes_only_duplicates():
    # parity.sh:96-126: raw-HSP count is partition-invariant under max_hits
    p1 = Pair.fixture("repeat", max_hits=16711680)
    p2 = Pair.fixture("repeat", max_hits=1000)
    a, sa, _ = p1.oracle().run_pass()
    b, sb, _ = p2.oracle().run_pass()
    assert sa["raw_hsps"] == sb["raw_hsps"] and sa["hits"] == sb["hits"]
    assert sb["hsps"] >= sa["hsps"]
    fa = {tuple(int(x) for x in r) for f, rc in a for r in list(f) + list(rc)}
    fb = {tuple(int(x) for x in r) for f, rc in b for r in list(f) + list(rc)}
    assert fa <= fb

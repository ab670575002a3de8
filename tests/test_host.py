"""Host-side path functions of the native library against the oracle (CPU)."""
import numpy as np
import pytest

import oracle_ffi as O
from hspz_amd import engine as E
from hspz_amd import sequence as S
from hspz_amd import synth


def _seq_with_noise(n, seed):
    rng = np.random.default_rng(seed)
    s = bytearray(synth.random_bases(seed, n).tobytes())
    for i in rng.integers(0, n, n // 200):
        s[i] = ord("N") if i % 3 == 0 else (ord("a") if i % 3 == 1 else ord("&"))
    return bytes(s)


@pytest.mark.parametrize("threads,step", [(1, 1), (2, 1), (3, 1), (8, 1), (4, 3)])
def test_seed_table_parallel_is_byte_identical_to_serial(threads, step):
    # seed.rs:546-597: every thread count and step gives the serial tables
    seq = _seq_with_noise(400_000, 0x2545F491)
    want_i, want_p = O.seed_table(seq, O.shape("12of19"), step)
    got_i, got_p = E.seed_table_build(seq, "12of19", step, threads)
    assert np.array_equal(got_i, want_i)
    assert np.array_equal(got_p, want_p)


def test_seed_table_tiny_inputs():
    for seq in (b"", b"ACGT", b"AAAAACGTA"):
        i, p = E.seed_table_build(seq, "12of19", 1, 4)
        assert len(p) == 0 and i[-1] == 0
    i, p = E.seed_table_build(b"AAAAACGTA", "1111", 1, 4)
    assert list(p[:i[0]]) == [1] and i[-1] == 5


@pytest.mark.parametrize("threads", [1, 2, 3, 8, 16])
def test_chunk_seeds_parallel_is_bit_identical(threads):
    # seed.rs:626-646
    seq = _seq_with_noise(60_000, 99)
    hi = len(seq) - 19
    want = O.chunk_seeds(seq, O.shape("12of19"), True, 0, hi)
    got = E.chunk_seeds(seq, 0, hi, "12of19", True, threads)
    assert np.array_equal(got, want)
    assert np.array_equal(E.chunk_seeds(seq, 100, 5000, "12of19", False, threads),
                          O.chunk_seeds(seq, O.shape("12of19"), False, 100, 5000))


def test_chunk_seeds_rejects_a_range_past_the_block():
    with pytest.raises(E.HspzError) as ei:
        E.chunk_seeds(b"ACGT" * 10, 0, 30)
    assert ei.value.status == 4


def test_dedup_matches_oracle_on_random_sets():
    rng = np.random.default_rng(5)
    for _ in range(50):
        n = int(rng.integers(0, 300))
        rows = [(int(rng.integers(0, 60)), int(rng.integers(0, 60)), int(rng.integers(0, 20)),
                 int(rng.integers(-5, 5))) for _ in range(n)]
        want = O.dedup_and_order(rows)
        got = E.dedup_and_order(np.array(rows, dtype=E.SP_DTYPE)) if n else []
        assert [tuple(int(x) for x in r) for r in got] == want


def test_dedup_reference_kats():
    got = E.dedup_and_order(np.array([(0, 0, 5, 10), (0, 0, 10, 20), (6, 6, 2, 30)],
                                     dtype=E.SP_DTYPE))
    assert [tuple(int(x) for x in r) for r in got] == [(0, 0, 5, 10)]


def test_default_max_hits():
    # gpu/mod.rs:2423-2432: 4194304 * GiB as f32
    assert E.default_max_hits(0) == 4194304
    assert E.default_max_hits(1 << 30) == 4194304
    assert E.default_max_hits(288 * (1 << 30)) == 4194304 * 288


def test_shape_parse():
    assert E.shape_parse("12of19")[0:2] == (19, 12)
    assert E.shape_parse("14of22")[0:2] == (22, 14)
    with pytest.raises(E.HspzError):
        E.shape_parse("101")


def test_sequence_helpers():
    buf, chrs, bl = S.pack([("a", b"ACGT"), ("b", b"GGa")])
    assert buf == b"ACGT&GGa&" and bl == 8
    assert [(c.name, c.start, c.len) for c in chrs] == [("a", 0, 4), ("b", 5, 3)]
    rc, rcc = S.reverse_complement(buf, chrs, bl)
    assert rc == b"tCC&ACGT"
    assert [(c.name, c.start) for c in rcc] == [("b", 0), ("a", 4)]
    assert list(S.encode(b"Acgt&NnRx")) == [0, 4, 4, 4, 7, 5, 5, 6, 6]
    chrs2 = [S.Chr("a", 0, 10), S.Chr("b", 11, 10)]
    assert [S.chr_at(chrs2, o) for o in (0, 10, 11, 20)] == [0, 0, 1, 1]
    raw = bytes(range(256))
    want = np.zeros(len(raw), dtype=np.uint8)
    O.lib().orc_revcomp(O._ptr(np.frombuffer(raw, dtype=np.uint8).copy()), len(raw), O._ptr(want))
    assert S.reverse_complement(raw + b"&", [S.Chr("x", 0, 256)], 256)[0] == want.tobytes()


@pytest.mark.parametrize("amb", ["", "n", "iupac", "iupac,5,7", "n,3,4", "x"])
@pytest.mark.parametrize("xdrop", [910, 50])
def test_build_sub_mat_matches_oracle(amb, xdrop):
    blk = [100 if i == j else -50 for i in range(4) for j in range(4)]
    assert np.array_equal(E.build_sub_mat(amb, xdrop), O.sub_mat(amb, xdrop))
    assert np.array_equal(E.build_sub_mat(amb, xdrop, blk), O.sub_mat(amb, xdrop, blk))


def test_build_sub_mat_rejects_malformed_triplet():
    with pytest.raises(E.HspzError):
        E.build_sub_mat("iupac,a,7")

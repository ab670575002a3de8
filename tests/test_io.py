"""Input and output paths of the native library (include/hspz_io.h), on the CPU.

Input (hspZ src/sequence.rs): the three formats load to the same packed set
(sequence.rs:469-526: FASTA, gzipped FASTA and a 2bit file written here from
the UCSC format description), reader edge cases, the complement table and
interval KATs (sequence.rs:433-452), encode's doc example (sequence.rs:63).

Output (src/hsp.rs, src/partition.rs, src/sink.rs, src/run.rs): the frozen
segment digests of the repeat and multi fixtures through the native emitter,
the partitioner KATs of partition.rs:197-303, Python's statistics.quantiles as
the definition the partitioner restates, the -D naming, and the .tar.gz round
trip and reproducibility of sink.rs:127-195.
"""
import gzip
import hashlib
import os
import statistics
import struct
import tarfile

import numpy as np
import pytest

from hspz_amd import seqio as X
from hspz_amd import sequence as S
from helpers import Pair

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures")


# ---- a 2bit writer from the UCSC format description -------------------------

def write_2bit(path, records, big_endian=False, version=0):
    """header {magic, version, count, reserved}; index {u8 name length, name,
    offset (u32, or u64 for version 1)}; records {dnaSize, nBlockCount,
    starts..., sizes..., maskBlockCount, starts..., sizes..., reserved, bases
    2 bits each, T0 C1 A2 G3, first base in the top bits}."""
    e = ">" if big_endian else "<"
    u32 = lambda v: struct.pack(e + "I", v)
    off = u32 if version == 0 else (lambda v: struct.pack(e + "Q", v))

    def runs(seq, pred):
        out, i = [], 0
        while i < len(seq):
            if pred(seq[i]):
                s = i
                while i < len(seq) and pred(seq[i]):
                    i += 1
                out.append((s, i - s))
            else:
                i += 1
        return out

    bodies = []
    for name, seq in records:
        nb = runs(seq, lambda b: b in b"Nn")
        mb = runs(seq, lambda b: 97 <= b <= 122)
        body = u32(len(seq)) + u32(len(nb)) + b"".join(u32(s) for s, _ in nb) + \
            b"".join(u32(l) for _, l in nb) + u32(len(mb)) + b"".join(u32(s) for s, _ in mb) + \
            b"".join(u32(l) for _, l in mb) + u32(0)
        code = {ord("T"): 0, ord("C"): 1, ord("A"): 2, ord("G"): 3}
        packed = bytearray()
        for i in range(0, len(seq), 4):
            v = 0
            for k in range(4):
                b = seq[i + k] if i + k < len(seq) else ord("T")
                v |= code.get(b & ~0x20, 0) << (6 - 2 * k)
            packed.append(v)
        bodies.append(body + bytes(packed))
    head = u32(0x1A412743) + u32(version) + u32(len(records)) + u32(0)
    index_len = sum(1 + len(n) + (4 if version == 0 else 8) for n, _ in records)
    at = len(head) + index_len
    index = b""
    for (name, _), body in zip(records, bodies):
        index += bytes([len(name)]) + name.encode() + off(at)
        at += len(body)
    with open(path, "wb") as fh:
        fh.write(head + index + b"".join(bodies))


RECORDS = [("chrA", b"ACGTacgtNNAACG"), ("chrB", b"ttttGCttttN")]


def write_fasta(path, records, width=60):
    with open(path, "wb") as fh:
        for n, s in records:
            fh.write(b">" + n.encode() + b"\n")
            for i in range(0, len(s), width):
                fh.write(s[i:i + width] + b"\n")


def test_three_formats_load_to_the_same_set(tmp_path):
    fa, gz, tb = tmp_path / "x.fa", tmp_path / "x.fa.gz", tmp_path / "x.2bit"
    write_fasta(fa, RECORDS)
    gz.write_bytes(gzip.compress(fa.read_bytes()))
    write_2bit(tb, RECORDS)
    sets = [X.SeqSet.read(str(p)) for p in (fa, gz, tb)]
    assert [s.format for s in sets] == ["fasta", "fasta.gz", "2bit"]
    for s in sets:
        assert s.buf == b"ACGTacgtNNAACG&ttttGCttttN&"
        assert s.block_len == len(s.buf) - 1
        assert [(c.name, c.start, c.len) for c in s.chrs] == [("chrA", 0, 14), ("chrB", 15, 11)]
        assert s.digest == sets[0].digest
    assert sets[2].bytes_read == tb.stat().st_size


@pytest.mark.parametrize("big_endian,version", [(False, 0), (True, 0), (False, 1), (True, 1)])
def test_2bit_byte_orders_versions_and_masked_n(tmp_path, big_endian, version):
    rng = np.random.default_rng(7 + version + 2 * big_endian)
    recs = []
    for i in range(4):
        s = bytearray(rng.choice(list(b"ACGT"), size=int(rng.integers(1, 400))).astype(np.uint8).tobytes())
        for _ in range(5):  # N runs, soft-masked runs (which may cover Ns: 'n')
            a = int(rng.integers(0, len(s)))
            b = min(len(s), a + int(rng.integers(1, 30)))
            s[a:b] = b"N" * (b - a)
            a = int(rng.integers(0, len(s)))
            b = min(len(s), a + int(rng.integers(1, 50)))
            s[a:b] = bytes(s[a:b]).lower()
        recs.append(("r%d" % i, bytes(s)))
    recs.append(("empty", b""))
    fa, tb = tmp_path / "x.fa", tmp_path / "x.2bit"
    write_fasta(fa, recs, width=70)
    write_2bit(tb, recs, big_endian=big_endian, version=version)
    a, b = X.SeqSet.read(str(fa)), X.SeqSet.read(str(tb))
    assert b.format == "2bit" and a.buf == b.buf and a.digest == b.digest
    assert b"n" in b.buf


def test_2bit_truncated_is_an_error(tmp_path):
    tb = tmp_path / "x.2bit"
    write_2bit(tb, RECORDS)
    data = tb.read_bytes()
    tb.write_bytes(data[:-2])
    with pytest.raises(X.HspzError) as e:
        X.SeqSet.read(str(tb))
    assert e.value.status == 8


def test_multi_member_gzip_and_fixtures_match_the_python_reader(tmp_path):
    fa = os.path.join(FIX, "multi.qry.fa")
    raw = open(fa, "rb").read()
    cut = len(raw) // 3
    gz = tmp_path / "m.fa.gz"
    gz.write_bytes(gzip.compress(raw[:cut]) + gzip.compress(raw[cut:]))
    assert X.SeqSet.read(str(gz)).buf == X.SeqSet.read(fa).buf
    for name in sorted(os.listdir(FIX)):
        path = os.path.join(FIX, name)
        s = X.SeqSet.read(path)
        buf, chrs, bl = S.pack(S.read_fasta(path))
        assert s.buf == buf and s.block_len == bl
        assert [(c.name, c.start, c.len) for c in s.chrs] == [(c.name, c.start, c.len) for c in chrs]


def test_fasta_reader_edge_cases(tmp_path):
    # before the first header: dropped; CRLF; interior whitespace incl. form
    # feed; a vertical tab is not ASCII whitespace in Rust and stays a byte;
    # empty name; name = first token; no final newline; invalid UTF-8 names
    body = (b"ACGT\n>  first desc here\r\nAC GT\r\n\tac\x0cgt \n>\nNN\x0bN\n>third\n"
            b">bad\xff\xe2\x82name x\nACG")
    p = tmp_path / "e.fa"
    p.write_bytes(body)
    s = X.SeqSet.read(str(p), prefix="q.")
    names = [c.name for c in s.chrs]
    assert names == ["q.first", "q.", "q.third", "q.bad��name"]
    assert names[3] == "q." + b"bad\xff\xe2\x82name".decode("utf-8", "replace")
    assert s.buf == b"ACGTacgt&NN\x0bN&&ACG&"
    assert [(c.start, c.len) for c in s.chrs] == [(0, 8), (9, 4), (14, 0), (15, 3)]
    buf, chrs, bl = S.pack(S.read_fasta(str(p)), prefix="q.")
    assert buf == s.buf and [c.name for c in chrs] == names


def test_empty_input_and_pack_digest_distinguishes(tmp_path):
    p = tmp_path / "none.fa"
    p.write_bytes(b"")
    s = X.SeqSet.read(str(p))
    assert s.buf == b"" and s.block_len == 0 and s.chrs == []
    a = X.SeqSet.pack([("x", b"ACGT"), ("y", b"GG")])
    assert a.buf == b"ACGT&GG&" and a.block_len == 7
    for other in ([("x", b"ACGT"), ("y", b"Gg")], [("y", b"GG"), ("x", b"ACGT")],
                  [("x", b"ACGT"), ("z", b"GG")], [("x", b"ACGTG"), ("y", b"G")]):
        assert X.SeqSet.pack(other).digest != a.digest
    assert X.SeqSet.pack([("x", b"ACGT"), ("y", b"GG")]).digest == a.digest


def test_reverse_complement_table_and_remap():
    rc = X.SeqSet.pack([("a", bytes(range(1, 256)))]).reverse_complement()
    comp = rc.buf[::-1]  # complement of bytes 1..255 in order
    table = {b: comp[b - 1] for b in range(1, 256)}
    for a, b in ((b"A", b"T"), (b"C", b"G"), (b"G", b"C"), (b"T", b"A"), (b"&", b"&"), (b"a", b"t"),
                 (b"N", b"N"), (b"U", b"A"), (b"j", b"n"), (b"J", b"N")):
        assert table[a[0]] == b[0]
    iupac = b"ACGTBVDHKMRYSWNacgtbvdhkmryswn"
    for c in iupac:  # complement is an involution on IUPAC codes
        assert table[table[c]] == c
    assert all(table[c] == ord("N") for c in range(128, 256))
    s = X.SeqSet.pack([("a", b"AACG"), ("b", b"TTT"), ("c", b"GacN")])
    r = s.reverse_complement()
    assert r.buf == b"NgtC&AAA&CGTT" and r.block_len == 13
    assert [(c.name, c.start, c.len) for c in r.chrs] == [("c", 0, 4), ("b", 5, 3), ("a", 9, 4)]
    buf, chrs, bl = S.pack([("a", b"AACG"), ("b", b"TTT"), ("c", b"GacN")])
    prc, pchrs = S.reverse_complement(buf, chrs, bl)
    assert prc == r.buf


def test_encode_and_intervals():
    assert list(X.encode(b"Acgt&")) == [0, 4, 4, 4, 7]
    allb = bytes(range(256))
    assert np.array_equal(X.encode(allb), S.encode(allb))
    big = np.random.default_rng(1).integers(0, 256, size=5_000_000, dtype=np.uint8).tobytes()
    assert np.array_equal(X.encode(big, threads=4), S.encode(big))
    assert X.intervals(376676, 19, 10_000_000) == [(0, 376657)]
    assert X.intervals(1019, 19, 500) == [(0, 500), (500, 1000)]
    assert X.intervals(19, 19, 500) == []
    for bl in (20, 999, 10_000_019, 25_000_000):
        assert X.intervals(bl, 19, 10_000_000) == S.intervals(bl, 19, 10_000_000)


# ---- output ------------------------------------------------------------------

def _dir_digest(d):
    lines, files = [], 0
    for f in sorted(os.listdir(d)):
        files += 1
        lines += open(os.path.join(d, f), "rb").read().splitlines(keepends=True)
    lines.sort()
    return len(lines), files, hashlib.sha256(b"".join(lines)).hexdigest()[:16]


@pytest.mark.parametrize("name,want", [("repeat", (57, 1, "d01edd2118cf5fa5")),
                                       ("multi", (3, 2, "0da3bbda656e4f7b"))])
def test_emitter_reproduces_the_frozen_digests(tmp_path, name, want):
    p = Pair.fixture(name)
    res, _, h = p.oracle().run_pass()
    assert X.hash_pass(res) == h
    em = X.Emitter(str(tmp_path / "out"))
    em.emit_unit(res, p.rchrs, p.qchrs, p.rcchrs)
    rep = em.finish()
    assert _dir_digest(tmp_path / "out") == want
    assert rep["files"] == want[1]
    # hash_segments (benchmark.rs:425-436) = FNV-1a over the rendered text
    hv = 0xcbf29ce484222325
    for fw, rc in res:
        for hsps, chrs, strand in ((fw, p.qchrs, "+"), (rc, p.rcchrs, "-")):
            for byte in S.render_segments([tuple(int(x) for x in r) for r in hsps], p.rchrs, chrs,
                                          strand).encode():
                hv = ((hv ^ byte) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    assert X.hash_segments(res, p.rchrs, p.qchrs, p.rcchrs) == hv


def _rec(rc, qc, rs, qs, ln, score=100, r_end=None, q_end=None):
    return (rc, qc, rs, rs + ln if r_end is None else r_end, qs, qs + ln if q_end is None else q_end, score, 0)


def _recs(rows):
    return np.array(rows, dtype=X.REC_DTYPE)


def _keys(parts):
    return [[(int(r["r_chr"]), int(r["q_chr"]), int(r["r_start"]), int(r["q_start"])) for r in p] for p in parts]


def test_records_and_render_match_the_python_writer():
    chrs = [X.Chr("a", 0, 10), X.Chr("b", 11, 10)]
    qchrs = [X.Chr("q", 0, 30)]
    hsps = [(0, 3, 5, 77), (10, 4, 1, -5), (11, 0, 9, 3000), (20, 29, 1, 1)]
    for strand in "+-":
        recs = X.records(hsps, chrs, qchrs, strand)
        order = hsps if strand == "+" else hsps[::-1]
        assert [int(r["r_chr"]) for r in recs] == [S.chr_at(chrs, h[0]) for h in order]
        text = X.render_records(recs, chrs, qchrs, strand)
        assert text == S.render_segments(hsps, chrs, qchrs, strand)
        assert X.render_segments(hsps, chrs, qchrs, strand) == text
    assert X.render_segments([], [], [], "+") == ""
    with pytest.raises(X.HspzError):
        X.records([(0, 0, 1, 1)], [X.Chr("a", 5, 10)], qchrs, "+")


def test_diagonal_keys_order_by_sum_or_difference():
    a, b, c = _rec(0, 0, 100, 10, 0), _rec(0, 0, 10, 100, 0), _rec(0, 0, 1, 1, 0)
    # one pair is always split; chunk 1 exposes the sort order
    plus = _keys(X.partition_split(_recs([a, b, c]), "+", 1))
    assert plus == [[(0, 0, 1, 1)], [(0, 0, 10, 100)], [(0, 0, 100, 10)]]  # sums 2, 110, 110; r_mid breaks the tie
    minus = _keys(X.partition_split(_recs([a, b, c]), "-", 1))
    assert minus == [[(0, 0, 100, 10)], [(0, 0, 1, 1)], [(0, 0, 10, 100)]]  # differences -90, 0, 90


def test_midpoint_uses_the_reference_half_length():
    x = _rec(0, 0, 10, 100, 0, r_end=17, q_end=100)  # (13 + 103, 13) with the reference half
    y = _rec(0, 0, 0, 115, 0)                           # (115, 0)
    z = _rec(0, 0, 0, 117, 0)                           # (117, 0)
    got = _keys(X.partition_split(_recs([x, y, z]), "+", 1))
    assert got == [[(0, 0, 0, 115)], [(0, 0, 10, 100)], [(0, 0, 0, 117)]]


def test_chunk_size_kats_and_quantiles():
    assert X.partition_chunk_size([], 4999) is None
    assert X.partition_chunk_size([], 50000) is None
    assert X.partition_chunk_size([], 50001) == 50000
    assert X.partition_chunk_size([7], 50001) == 50000
    rng = np.random.default_rng(3)
    for n in range(2, 30):
        hist = [int(v) for v in rng.integers(1, 80000, size=n)]
        q = statistics.quantiles(hist, n=4)  # the exclusive method
        pick = q[1] if n < 7 else q[2]
        chunk = min(int(pick), 50000)
        for count in (4999, 5000, chunk, chunk + 1, 120000):
            want = None if count < 5000 or count <= chunk else chunk
            assert X.partition_chunk_size(hist, count) == want, (hist, count)
    assert X.partition_chunk_size([1, 2, 3, 4], 5000) == 2  # median of [1, 2, 3, 4] is 2.5


def test_split_kats():
    recs = _recs([_rec(0, 0, i * 10, i * 10, 5) for i in range(10)])
    files = X.partition_split(recs, "+", 4)
    assert [len(f) for f in files] == [4, 4, 2]
    rows = [_rec(0, 0, i * 10, i * 10, 1) for i in range(6)]
    rows += [_rec(1, 1, i * 10, i * 10, 1) for i in range(2)]
    rows += [_rec(2, 2, i * 10, i * 10, 1) for i in range(2)]
    files = X.partition_split(_recs(rows), "+", 4)
    assert sorted(len(f) for f in files) == [2, 4, 4]
    rows = [_rec(c, c, i * 3, i * 5, 2) for c in range(3) for i in range(7)]
    files = X.partition_split(_recs(rows), "-", 3)
    after = sorted(tuple(int(v) for v in r) for f in files for r in f)
    assert after == sorted(tuple(int(v) for v in r) for r in _recs(rows))


def test_split_packs_small_pairs_in_size_order_then_query_order():
    # pairs (first appearance): (0,5) x3, (1,2) x1, (2,9) x2, (3,2) x1, (0,0) x6; chunk 4
    rows = [_rec(0, 5, i, i, 1) for i in range(3)] + [_rec(1, 2, 0, 0, 1)] + \
        [_rec(2, 9, i, i, 1) for i in range(2)] + [_rec(3, 2, 0, 0, 1)] + [_rec(0, 0, i, 40 - i, 1) for i in range(6)]
    got = X.partition_split(_recs(rows), "+", 4)
    pairs = [[(int(r["r_chr"]), int(r["q_chr"])) for r in f] for f in got]
    # (0,0) is split first (sorted by q+r sum, all 40 + half: r_mid ascending), then the packs:
    # ascending (size, pair): (1,2)1 (3,2)1 (2,9)2 | (0,5)3; each pack ordered by query first appearance
    assert pairs[0] == [(0, 0)] * 4 and pairs[1] == [(0, 0)] * 2
    assert pairs[2] == [(1, 2), (3, 2), (2, 9), (2, 9)]
    assert pairs[3] == [(0, 5)] * 3


def test_emitter_diagonal_partition_names_and_contents(tmp_path):
    r_chrs = [X.Chr("refA", 0, 100000), X.Chr("refB", 100001, 100000)]
    q_chrs = [X.Chr("qryA", 0, 100000)]
    rc_chrs = [X.Chr("qryA", 0, 100000)]
    big = [(100 + i, 50 + i // 2, 50 + i % 10, 3000 + i) for i in range(5500)]
    small = [(100001 + i, 60 + i, 40, 4000) for i in range(7)]
    fw = np.array(big + small, dtype=X.SP_DTYPE)
    rc = np.array(big[:100], dtype=X.SP_DTYPE)
    em = X.Emitter(str(tmp_path / "d"), diagonal=True)
    em.emit_unit([(fw, rc), (rc, fw[:0])], r_chrs, q_chrs, rc_chrs, ref_bin=2, query_bin=5)
    rep = em.finish()
    names = sorted(os.listdir(tmp_path / "d"))
    # history [5507]: cap 50000 -> whole; [5507, 100]: 100 < 5000 -> whole; [.., 100]
    assert names == ["tmp1.block5.r2.minus.segments", "tmp1.block5.r2.plus.segments",
                     "tmp2.block5.r2.plus.segments"]
    assert rep["files"] == 3
    # with a history the median decides: after two small files it is small
    em = X.Emitter(str(tmp_path / "f"), diagonal=True)
    tiny = fw[5500:5503]
    em.emit_unit([(tiny, tiny[:0])], r_chrs, q_chrs, rc_chrs, query_bin=0)
    em.emit_unit([(tiny, tiny[:0])], r_chrs, q_chrs, rc_chrs, query_bin=1)
    em.emit_unit([(fw, fw[:0])], r_chrs, q_chrs, rc_chrs, query_bin=2)
    em.finish()
    names = sorted(os.listdir(tmp_path / "f"))
    chunk = X.partition_chunk_size([3, 3, 5507], 5507)
    recs = X.records(fw, r_chrs, q_chrs, "+")
    parts = X.partition_split(recs, "+", chunk)
    want = ["tmp1.block2.r0.plus.split%d.segments" % (k + 1) for k in range(len(parts))]
    assert sorted(n for n in names if "block2" in n) == sorted(want)
    for k, part in enumerate(parts):
        text = open(tmp_path / "f" / ("tmp1.block2.r0.plus.split%d.segments" % (k + 1))).read()
        assert text == X.render_records(part, r_chrs, q_chrs, "+")
    union = sorted(l for k in range(len(parts)) for l in
                   open(tmp_path / "f" / want[k]).read().splitlines())
    assert union == sorted(X.render_segments(fw, r_chrs, q_chrs, "+").splitlines())


def test_tarball_round_trip_matches_directory_and_is_reproducible(tmp_path):
    p = Pair.fixture("multi")
    res, _, _ = p.oracle().run_pass()
    big = np.array([(i, i, 20, 3000 + i) for i in range(12000)], dtype=X.SP_DTYPE)
    units = [(res, 0), ([(big, big[:100])], 1)]
    r_chrs = p.rchrs
    em = X.Emitter(str(tmp_path / "out"))
    for u, qb in units:
        em.emit_unit(u, r_chrs, p.qchrs, p.rcchrs, query_bin=qb)
    em.finish()
    archives = []
    for i in range(2):
        em = X.Emitter(str(tmp_path / "out"), tarball=("" if i == 0 else str(tmp_path / "sub" / "o2.tgz")))
        for u, qb in units:
            em.emit_unit(u, r_chrs, p.qchrs, p.rcchrs, query_bin=qb)
        rep = em.finish()
        archives.append(tmp_path / "out.tar.gz" if i == 0 else tmp_path / "sub" / "o2.tgz")
        assert rep["bytes_out"] == archives[-1].stat().st_size
    assert archives[0].read_bytes() == archives[1].read_bytes()
    with tarfile.open(archives[0], "r:gz") as t:
        members = t.getmembers()
        names = [m.name for m in members]
        assert sorted(names) == sorted(os.listdir(tmp_path / "out"))
        for m in members:
            assert not m.name.startswith("/") and not m.name.startswith("./")
            assert m.mode == 0o644 and m.mtime == 0 and m.isfile()
            assert t.extractfile(m).read() == (tmp_path / "out" / m.name).read_bytes()
    assert gzip.decompress(archives[0].read_bytes())[-1024:] == bytes(1024)

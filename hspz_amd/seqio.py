"""ctypes binding of the native input/output paths (include/hspz_io.h).

Input: `SeqSet.read(path)` loads FASTA, gzipped FASTA or 2bit (by magic) into
the packed block the engine takes (hspZ src/sequence.rs:238-248 read_records +
pack), `SeqSet.pack(records)`, `.reverse_complement()`, `.digest`, and
`encode` / `intervals`.

Output: `records` / `render_records` / `render_segments` (src/hsp.rs:213-292),
`partition_chunk_size` / `partition_split` (src/partition.rs), `hash_pass` /
`hash_segments` (src/benchmark.rs:411-436), and `Emitter` (run.rs:535-666:
`.segments` files into a directory or one .tar.gz, `-D` diagonal partitions).
All of it runs in libhspz.so; nothing here falls back to Python.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .engine import SP_DTYPE, HspzError, lib as _engine_lib

FORMAT = {0: "fasta", 1: "fasta.gz", 2: "2bit"}

REC_DTYPE = np.dtype([("r_chr", "<u4"), ("q_chr", "<u4"), ("r_start", "<u8"), ("r_end", "<u8"),
                      ("q_start", "<u8"), ("q_end", "<u8"), ("score", "<i4"), ("_pad", "<u4")])

# Every symbol include/hspz_io.h declares (checked by tests/test_abi.py).
EXPORTS = [
    "hspz_seqset_read", "hspz_seqset_pack", "hspz_seqset_reverse_complement", "hspz_seqset_get",
    "hspz_seqset_free", "hspz_seqset_digest", "hspz_encode", "hspz_intervals", "hspz_records", "hspz_render_records",
    "hspz_render_segments", "hspz_partition_chunk_size", "hspz_partition_split", "hspz_hash_pass",
    "hspz_hash_segments", "hspz_emitter_new", "hspz_emit_unit", "hspz_emitter_finish",
]


class ChrT(C.Structure):
    _fields_ = [("name", C.c_char_p), ("start", C.c_uint64), ("len", C.c_uint32)]


class SeqSetInfo(C.Structure):
    _fields_ = [("buf", C.c_void_p), ("buf_len", C.c_size_t), ("block_len", C.c_size_t),
                ("chrs", C.POINTER(ChrT)), ("n_chrs", C.c_size_t), ("format", C.c_int32),
                ("bytes_read", C.c_uint64)]


class OutputReport(C.Structure):
    _fields_ = [("partition_ms", C.c_double), ("format_ms", C.c_double), ("archive_ms", C.c_double),
                ("files", C.c_uint64), ("bytes_in", C.c_uint64), ("bytes_out", C.c_uint64)]

    def as_dict(self):
        return {k: (float(getattr(self, k)) if t is C.c_double else int(getattr(self, k)))
                for k, t in self._fields_}


_ready = False


def lib():
    global _ready
    L = _engine_lib()
    if not _ready:
        vp, sz, u32, u64 = C.c_void_p, C.c_size_t, C.c_uint32, C.c_uint64
        P = C.POINTER
        L.hspz_seqset_read.argtypes = [C.c_char_p, C.c_char_p, P(vp)]
        L.hspz_seqset_pack.argtypes = [sz, vp, vp, vp, C.c_char_p, P(vp)]
        L.hspz_seqset_reverse_complement.argtypes = [vp, P(vp)]
        L.hspz_seqset_get.argtypes = [vp, P(SeqSetInfo)]
        L.hspz_seqset_free.argtypes = [vp]
        L.hspz_seqset_digest.argtypes = [vp]
        L.hspz_seqset_digest.restype = u64
        L.hspz_encode.argtypes = [vp, sz, C.c_int, vp]
        L.hspz_intervals.argtypes = [sz, u32, u32, vp, sz, P(sz)]
        L.hspz_records.argtypes = [vp, sz, vp, sz, vp, sz, C.c_char, vp]
        L.hspz_render_records.argtypes = [vp, sz, vp, sz, vp, sz, C.c_char, P(vp), P(sz)]
        L.hspz_render_segments.argtypes = [vp, sz, vp, sz, vp, sz, C.c_char, P(vp), P(sz)]
        L.hspz_partition_chunk_size.argtypes = [vp, sz, u64, P(u64)]
        L.hspz_partition_split.argtypes = [vp, sz, C.c_char, u64, vp, vp, P(sz)]
        L.hspz_hash_pass.argtypes = [vp, vp, sz]
        L.hspz_hash_pass.restype = u64
        L.hspz_hash_segments.argtypes = [vp, vp, sz, vp, sz, vp, sz, vp, sz, P(u64)]
        L.hspz_emitter_new.argtypes = [C.c_char_p, C.c_char_p, C.c_int, P(vp)]
        L.hspz_emit_unit.argtypes = [vp, u32, u32, vp, sz, vp, sz, vp, sz, vp, vp, sz]
        L.hspz_emitter_finish.argtypes = [vp, P(OutputReport)]
        L.hspz_free.argtypes = [vp]
        _ready = True
    return L


def _check(rc: int):
    if rc != 0:
        msg = _engine_lib().hspz_last_error()
        raise HspzError(rc, msg.decode(errors="replace") if msg else "")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p) if a.size else None


@dataclass
class Chr:
    name: str
    start: int
    len: int


class ChrTable:
    """A record table laid out as the C ABI's hspz_chr array (names kept alive)."""

    def __init__(self, chrs):
        self._names = [(c.name if isinstance(c.name, bytes) else c.name.encode()) for c in chrs]
        self.arr = (ChrT * max(len(chrs), 1))()
        for i, c in enumerate(chrs):
            self.arr[i] = ChrT(self._names[i], int(c.start), int(c.len) & 0xFFFFFFFF)
        self.n = len(chrs)

    @property
    def ptr(self):
        return C.cast(self.arr, C.c_void_p)


def _table(chrs):
    return chrs if isinstance(chrs, ChrTable) else ChrTable(chrs)


class SeqSet:
    """A packed sequence set owned by the native library (hspz_seqset)."""

    def __init__(self, handle):
        self._h = handle
        info = SeqSetInfo()
        _check(lib().hspz_seqset_get(self._h, C.byref(info)))
        self._info = info
        self._buf = None
        self.block_len = int(info.block_len)
        self.chrs = [Chr(info.chrs[i].name.decode("utf-8"), int(info.chrs[i].start), int(info.chrs[i].len))
                     for i in range(info.n_chrs)]
        self.format = FORMAT[int(info.format)]
        self.bytes_read = int(info.bytes_read)

    @property
    def buf(self) -> bytes:
        """The packed bytes (copied out of the library on first use)."""
        if self._buf is None:
            i = self._info
            self._buf = C.string_at(i.buf, i.buf_len) if i.buf_len else b""
        return self._buf

    @property
    def digest(self) -> int:
        return int(lib().hspz_seqset_digest(self._h))

    @classmethod
    def read(cls, path: str, prefix: str = "") -> "SeqSet":
        h = C.c_void_p()
        _check(lib().hspz_seqset_read(path.encode(), prefix.encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def pack(cls, records, prefix: str = "") -> "SeqSet":
        names = [(n if isinstance(n, bytes) else n.encode()) for n, _ in records]
        seqs = [bytes(s) if not isinstance(s, str) else s.encode() for _, s in records]
        n = len(records)
        cn = (C.c_char_p * max(n, 1))(*names)
        cs = (C.c_char_p * max(n, 1))(*seqs)
        cl = (C.c_size_t * max(n, 1))(*[len(s) for s in seqs])
        h = C.c_void_p()
        _check(lib().hspz_seqset_pack(n, C.cast(cn, C.c_void_p), C.cast(cs, C.c_void_p),
                                      C.cast(cl, C.c_void_p), prefix.encode(), C.byref(h)))
        return cls(h)

    def reverse_complement(self) -> "SeqSet":
        h = C.c_void_p()
        _check(lib().hspz_seqset_reverse_complement(self._h, C.byref(h)))
        return SeqSet(h)

    @property
    def raw(self) -> bytes:
        """The block's bytes without the trailing separator."""
        return self.buf[:self.block_len]

    def close(self):
        if self._h:
            lib().hspz_seqset_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def encode(raw: bytes, threads: int = 0) -> np.ndarray:
    a = np.frombuffer(bytes(raw), dtype=np.uint8)
    out = np.empty(len(a), dtype=np.uint8)
    _check(lib().hspz_encode(_ptr(a), len(a), threads, _ptr(out)))
    return out


def intervals(block_len: int, seed_size: int, interval_size: int):
    n = C.c_size_t()
    _check(lib().hspz_intervals(block_len, seed_size, interval_size, None, 0, C.byref(n)))
    pairs = np.zeros(2 * max(n.value, 1), dtype=np.uint32)
    _check(lib().hspz_intervals(block_len, seed_size, interval_size, _ptr(pairs), n.value, C.byref(n)))
    return [(int(pairs[2 * i]), int(pairs[2 * i + 1])) for i in range(n.value)]


def _pairs(hsps) -> np.ndarray:
    if isinstance(hsps, np.ndarray) and hsps.dtype == SP_DTYPE:
        return np.ascontiguousarray(hsps)
    rows = [tuple(int(x) for x in r) for r in hsps]
    return np.array(rows, dtype=SP_DTYPE) if rows else np.zeros(0, dtype=SP_DTYPE)


def _text(p, n) -> str:
    s = C.string_at(p, n.value).decode("utf-8") if n.value else ""
    lib().hspz_free(p)
    return s


def records(hsps, r_chrs, q_chrs, strand: str) -> np.ndarray:
    a = _pairs(hsps)
    r, q = _table(r_chrs), _table(q_chrs)
    out = np.zeros(len(a), dtype=REC_DTYPE)
    _check(lib().hspz_records(_ptr(a), len(a), r.ptr, r.n, q.ptr, q.n, strand.encode(), _ptr(out)))
    return out


def render_records(recs: np.ndarray, r_chrs, q_chrs, strand: str) -> str:
    recs = np.ascontiguousarray(recs, dtype=REC_DTYPE)
    r, q = _table(r_chrs), _table(q_chrs)
    p, n = C.c_void_p(), C.c_size_t()
    _check(lib().hspz_render_records(_ptr(recs), len(recs), r.ptr, r.n, q.ptr, q.n, strand.encode(),
                                     C.byref(p), C.byref(n)))
    return _text(p, n)


def render_segments(hsps, r_chrs, q_chrs, strand: str) -> str:
    a = _pairs(hsps)
    r, q = _table(r_chrs), _table(q_chrs)
    p, n = C.c_void_p(), C.c_size_t()
    _check(lib().hspz_render_segments(_ptr(a), len(a), r.ptr, r.n, q.ptr, q.n, strand.encode(),
                                      C.byref(p), C.byref(n)))
    return _text(p, n)


def partition_chunk_size(history, count: int):
    """None = emit whole, else the chunk size."""
    h = np.asarray(history, dtype=np.uint64)
    c = C.c_uint64()
    _check(lib().hspz_partition_chunk_size(_ptr(h), len(h), count, C.byref(c)))
    return int(c.value) or None


def partition_split(recs: np.ndarray, strand: str, chunk: int):
    recs = np.ascontiguousarray(recs, dtype=REC_DTYPE)
    out = np.zeros(len(recs), dtype=REC_DTYPE)
    sizes = np.zeros(max(len(recs), 1), dtype=np.uint64)
    n = C.c_size_t()
    _check(lib().hspz_partition_split(_ptr(recs), len(recs), strand.encode(), chunk, _ptr(out),
                                      _ptr(sizes), C.byref(n)))
    parts, at = [], 0
    for k in sizes[:n.value]:
        parts.append(out[at:at + int(k)])
        at += int(k)
    return parts


def _flatten(intervals_out):
    lists, counts = [], []
    for fw, rc in intervals_out:
        for x in (fw, rc):
            a = _pairs(x)
            lists.append(a)
            counts.append(len(a))
    flat = np.concatenate(lists) if lists else np.zeros(0, dtype=SP_DTYPE)
    return np.ascontiguousarray(flat, dtype=SP_DTYPE), np.array(counts, dtype=np.uintp)


def hash_pass(intervals_out) -> int:
    flat, counts = _flatten(intervals_out)
    return int(lib().hspz_hash_pass(_ptr(flat), _ptr(counts), len(counts) // 2))


def hash_segments(intervals_out, r_chrs, q_chrs, rc_chrs) -> int:
    flat, counts = _flatten(intervals_out)
    r, q, rc = _table(r_chrs), _table(q_chrs), _table(rc_chrs)
    h = C.c_uint64()
    _check(lib().hspz_hash_segments(_ptr(flat), _ptr(counts), len(counts) // 2, r.ptr, r.n, q.ptr, q.n,
                                    rc.ptr, rc.n, C.byref(h)))
    return int(h.value)


class Emitter:
    """run.rs Emitter: `Emitter(dir)` writes files, `Emitter(dir, tarball="")`
    writes <dir>.tar.gz, `diagonal=True` is -D."""

    def __init__(self, output_dir: str | None, tarball: str | None = None, diagonal: bool = False):
        h = C.c_void_p()
        _check(lib().hspz_emitter_new(output_dir.encode() if output_dir is not None else None,
                                      tarball.encode() if tarball is not None else None, int(diagonal),
                                      C.byref(h)))
        self._h = h

    def emit_unit(self, intervals_out, ref_chrs, query_chrs, rc_chrs, ref_bin: int = 0, query_bin: int = 0):
        flat, counts = _flatten(intervals_out)
        r, q, rc = _table(ref_chrs), _table(query_chrs), _table(rc_chrs)
        _check(lib().hspz_emit_unit(self._h, ref_bin, query_bin, r.ptr, r.n, q.ptr, q.n, rc.ptr, rc.n,
                                    _ptr(flat), _ptr(counts), len(counts) // 2))

    def finish(self) -> dict:
        rep = OutputReport()
        h, self._h = self._h, None
        _check(lib().hspz_emitter_finish(h, C.byref(rep)))
        return rep.as_dict()

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().hspz_emitter_finish(self._h, None)
            except Exception:
                pass

"""ctypes bindings to the CPU oracle (oracle/liboracle.so).

Test infrastructure: used by tests/, the smoke entry and bench.py's
cpu_baseline leg as the checker, never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

SP_DTYPE = np.dtype([("ref_start", "<u4"), ("query_start", "<u4"), ("len", "<u4"),
                     ("score", "<i4")])


class OrcShape(C.Structure):
    _fields_ = [("size", C.c_uint32), ("kmer_size", C.c_uint32), ("pos", C.c_uint32 * 32),
                ("transition", C.c_uint8 * 32)]


class OrcEngine(C.Structure):
    _fields_ = [("index_table", C.c_void_p), ("pos_table", C.c_void_p),
                ("ref_enc", C.c_void_p), ("ref_len", C.c_uint32),
                ("q_enc", C.c_void_p), ("q_rc_enc", C.c_void_p), ("query_len", C.c_uint32),
                ("sub", C.c_void_p), ("seed_size", C.c_uint32), ("xdrop", C.c_int32),
                ("hspthresh", C.c_int32), ("noentropy", C.c_int), ("max_hits", C.c_uint32)]


class OrcStats(C.Structure):
    _fields_ = [("seeds", C.c_uint64), ("hits", C.c_uint64), ("gate_pass", C.c_uint64),
                ("raw_hsps", C.c_uint64), ("hsps", C.c_uint64), ("cells", C.c_uint64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, u32, i32, sz = C.c_void_p, C.c_uint32, C.c_int32, C.c_size_t
        L.orc_encode.argtypes = [vp, sz, vp]
        L.orc_revcomp.argtypes = [vp, sz, vp]
        L.orc_build_sub_mat.argtypes = [C.c_char_p, i32, vp, vp]
        L.orc_shape_parse.argtypes = [C.c_char_p, C.POINTER(OrcShape)]
        L.orc_kmer_at.argtypes = [C.POINTER(OrcShape), vp, sz]
        L.orc_kmer_at.restype = u32
        L.orc_seed_table_build.argtypes = [vp, sz, C.POINTER(OrcShape), u32,
                                           C.POINTER(vp), C.POINTER(vp), C.POINTER(sz)]
        L.orc_free.argtypes = [vp]
        L.orc_chunk_seeds.argtypes = [vp, C.POINTER(OrcShape), C.c_int, u32, u32, vp]
        L.orc_chunk_seeds.restype = sz
        L.orc_intervals.argtypes = [sz, u32, u32, vp, sz]
        L.orc_intervals.restype = sz
        L.orc_chunks.argtypes = [u32, u32, u32, vp, sz]
        L.orc_chunks.restype = sz
        for f in (L.orc_extend_right, L.orc_extend_left):
            f.argtypes = [vp, u32, vp, u32, vp, i32, u32, u32, C.POINTER(i32), C.POINTER(i32),
                          C.POINTER(C.c_uint64)]
        L.orc_gate_keep.argtypes = [i32, i32]
        L.orc_ln.argtypes = [C.c_double]
        L.orc_ln.restype = C.c_double
        L.orc_materialize.argtypes = [vp, u32, vp, u32, vp, i32, i32, C.c_int, u32, u32, vp]
        L.orc_chunk_limits.argtypes = [vp, u32, u32, vp, sz]
        L.orc_chunk_limits.restype = C.c_long
        L.orc_dedup_and_order.argtypes = [vp, sz]
        L.orc_dedup_and_order.restype = sz
        L.orc_seed_and_filter.argtypes = [C.POINTER(OrcEngine), vp, u32, C.c_int,
                                          C.POINTER(vp), C.POINTER(sz), C.POINTER(OrcStats)]
        L.orc_run_pass.argtypes = [C.POINTER(OrcEngine), vp, vp, u32, C.POINTER(OrcShape),
                                   C.c_int, C.c_int, C.c_int, u32, u32, C.c_int,
                                   C.POINTER(vp), C.POINTER(vp), C.POINTER(sz),
                                   C.POINTER(OrcStats)]
        L.orc_hash_pass.argtypes = [vp, vp, sz]
        L.orc_hash_pass.restype = C.c_uint64
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def sub_mat(ambiguous: str = "", xdrop: int = 910, acgt=None) -> np.ndarray:
    out = np.zeros(64, dtype=np.int32)
    blk = None if acgt is None else np.ascontiguousarray(acgt, dtype=np.int32)
    rc = lib().orc_build_sub_mat(ambiguous.encode(), xdrop,
                                 None if blk is None else _ptr(blk), _ptr(out))
    if rc != 0:
        raise ValueError("malformed ambiguous spec %r" % ambiguous)
    return out


def shape(spec: str = "12of19") -> OrcShape:
    s = OrcShape()
    if lib().orc_shape_parse(spec.encode(), C.byref(s)) != 0:
        raise ValueError("bad seed shape %r" % spec)
    return s


def kmer_at(s: OrcShape, raw: bytes, pos: int) -> int:
    a = np.frombuffer(raw, dtype=np.uint8).copy()
    return int(lib().orc_kmer_at(C.byref(s), _ptr(a), pos))


def seed_table(raw: bytes, s: OrcShape, step: int = 1):
    a = np.frombuffer(raw, dtype=np.uint8).copy()
    ip, pp, n = C.c_void_p(), C.c_void_p(), C.c_size_t()
    if lib().orc_seed_table_build(_ptr(a), len(a), C.byref(s), step, C.byref(ip), C.byref(pp),
                                  C.byref(n)) != 0:
        raise MemoryError
    nk = 1 << (2 * s.kmer_size)
    index = np.ctypeslib.as_array(C.cast(ip, C.POINTER(C.c_uint32)), (nk,)).copy()
    pos = (np.ctypeslib.as_array(C.cast(pp, C.POINTER(C.c_uint32)), (max(n.value, 1),))
           [:n.value].copy())
    lib().orc_free(ip)
    lib().orc_free(pp)
    return index, pos


def chunk_seeds(raw: bytes, s: OrcShape, transitions: bool, lo: int, hi: int) -> np.ndarray:
    a = np.frombuffer(raw, dtype=np.uint8).copy()
    out = np.zeros(max(hi - lo, 1) * (1 + s.kmer_size), dtype=np.uint64)
    n = lib().orc_chunk_seeds(_ptr(a), C.byref(s), int(transitions), lo, hi, _ptr(out))
    return out[:n].copy()


def intervals(block_len: int, seed_size: int, interval_size: int):
    n = lib().orc_intervals(block_len, seed_size, interval_size, None, 0)
    out = np.zeros(max(2 * n, 2), dtype=np.uint32)
    lib().orc_intervals(block_len, seed_size, interval_size, _ptr(out), n)
    return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(n)]


def chunks(start: int, end: int, wga: int):
    n = lib().orc_chunks(start, end, wga, None, 0)
    out = np.zeros(max(2 * n, 2), dtype=np.uint32)
    lib().orc_chunks(start, end, wga, _ptr(out), n)
    return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(n)]


def extend(ref: np.ndarray, qry: np.ndarray, sub: np.ndarray, xdrop: int, r0: int, q0: int):
    """Returns (right_score, right_ext, left_score, left_ext, cells)."""
    rs, re_, ls, le = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    cells = C.c_uint64(0)
    lib().orc_extend_right(_ptr(ref), len(ref), _ptr(qry), len(qry), _ptr(sub), xdrop, r0, q0,
                           C.byref(rs), C.byref(re_), C.byref(cells))
    lib().orc_extend_left(_ptr(ref), len(ref), _ptr(qry), len(qry), _ptr(sub), xdrop, r0, q0,
                          C.byref(ls), C.byref(le), C.byref(cells))
    return rs.value, re_.value, ls.value, le.value, cells.value


def gate_keep(total: int, thresh: int) -> bool:
    return bool(lib().orc_gate_keep(total, thresh))


def ln(x: float) -> float:
    return lib().orc_ln(x)


def materialize(ref, qry, sub, xdrop, thresh, noentropy, r0, q0):
    out = np.zeros(1, dtype=SP_DTYPE)
    kept = lib().orc_materialize(_ptr(ref), len(ref), _ptr(qry), len(qry), _ptr(sub), xdrop,
                                 thresh, int(noentropy), r0, q0, _ptr(out))
    return tuple(int(v) for v in out[0]) if kept else None


def chunk_limits(hit_num, max_hits: int):
    h = np.ascontiguousarray(hit_num, dtype=np.uint32)
    n = lib().orc_chunk_limits(_ptr(h), len(h), max_hits, None, 0)
    if n < 0:
        raise ValueError("chunk_limits underflow")
    out = np.zeros(4 * n, dtype=np.uint32)
    lib().orc_chunk_limits(_ptr(h), len(h), max_hits, _ptr(out), n)
    return [tuple(int(v) for v in out[4 * i:4 * i + 4]) for i in range(n)]


def dedup_and_order(rows):
    a = np.array([tuple(r) for r in rows], dtype=SP_DTYPE)
    n = lib().orc_dedup_and_order(_ptr(a), len(a)) if len(a) else 0
    return [tuple(int(v) for v in a[i]) for i in range(n)]


class Oracle:
    """One engine-equivalent: a reference block, its index and a query block."""

    def __init__(self, ref_raw: bytes, q_raw: bytes, q_rc_raw: bytes, sub: np.ndarray,
                 seed: str = "12of19", step: int = 1, xdrop: int = 910, hspthresh: int = 3000,
                 noentropy: bool = False, max_hits: int = 16711680, index=None, ref_enc=None):
        """index: (index_table, pos_table) when already built; ref_enc: the
        encoded reference when already encoded (shared, not copied)."""
        from hspz_amd import sequence as S
        self.shape = shape(seed)
        self.ref_raw, self.q_raw, self.q_rc_raw = ref_raw, q_raw, q_rc_raw
        self.ref_enc = (np.ascontiguousarray(ref_enc, dtype=np.uint8) if ref_enc is not None
                        else S.encode(ref_raw).copy())
        self.q_enc = S.encode(q_raw).copy()
        self.q_rc_enc = S.encode(q_rc_raw).copy()
        self.sub = np.ascontiguousarray(sub, dtype=np.int32)
        self.index, self.pos = index if index is not None else seed_table(ref_raw, self.shape,
                                                                          step)
        self.pos_keep = self.pos if len(self.pos) else np.zeros(1, dtype=np.uint32)
        e = OrcEngine()
        e.index_table = _ptr(self.index).value
        e.pos_table = _ptr(self.pos_keep).value
        e.ref_enc = _ptr(self.ref_enc).value
        e.ref_len = len(self.ref_enc)
        e.q_enc = _ptr(self.q_enc).value
        e.q_rc_enc = _ptr(self.q_rc_enc).value
        e.query_len = len(self.q_enc)
        e.sub = _ptr(self.sub).value
        e.seed_size = self.shape.size
        e.xdrop = xdrop
        e.hspthresh = hspthresh
        e.noentropy = int(noentropy)
        e.max_hits = max_hits
        self.e = e

    def seed_and_filter(self, seeds: np.ndarray, rev: bool):
        s = np.ascontiguousarray(seeds, dtype=np.uint64)
        out, n, st = C.c_void_p(), C.c_size_t(), OrcStats()
        rc = lib().orc_seed_and_filter(C.byref(self.e), _ptr(s), len(s), int(rev), C.byref(out),
                                       C.byref(n), C.byref(st))
        if rc != 0:
            raise RuntimeError("oracle seed_and_filter failed: %d" % rc)
        res = np.zeros(n.value, dtype=SP_DTYPE)
        if n.value:
            C.memmove(_ptr(res), out, n.value * 16)
        lib().orc_free(out)
        return res, st.as_dict()

    def run_pass(self, transitions=True, plus=True, minus=True, wga_chunk=250000,
                 interval_size=10000000, threads=1):
        """Returns ([(fw, rc) per interval], stats, hash_pass)."""
        fwd = np.frombuffer(self.q_raw, dtype=np.uint8).copy()
        rc = np.frombuffer(self.q_rc_raw, dtype=np.uint8).copy()
        out, counts, ni, st = C.c_void_p(), C.c_void_p(), C.c_size_t(), OrcStats()
        status = lib().orc_run_pass(C.byref(self.e), _ptr(fwd), _ptr(rc), len(fwd),
                                    C.byref(self.shape), int(transitions), int(plus), int(minus),
                                    wga_chunk, interval_size, threads, C.byref(out),
                                    C.byref(counts), C.byref(ni), C.byref(st))
        if status != 0:
            raise RuntimeError("oracle run_pass failed: %d" % status)
        cnt = np.ctypeslib.as_array(C.cast(counts, C.POINTER(C.c_size_t)),
                                    (max(2 * ni.value, 1),))[:2 * ni.value].copy()
        total = int(cnt.sum())
        allsp = np.zeros(total, dtype=SP_DTYPE)
        if total:
            C.memmove(_ptr(allsp), out, total * 16)
        h = lib().orc_hash_pass(out, counts, ni.value)
        lib().orc_free(out)
        lib().orc_free(counts)
        res, at = [], 0
        for i in range(ni.value):
            fw = allsp[at:at + cnt[2 * i]]
            at += cnt[2 * i]
            rcl = allsp[at:at + cnt[2 * i + 1]]
            at += cnt[2 * i + 1]
            res.append((fw, rcl))
        return res, st.as_dict(), int(h)

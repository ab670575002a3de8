"""ctypes binding of the native library (include/hspz.h).

`Engine` mirrors the reference's `Engine` (hspZ src/gpu/mod.rs:817-2307):
`Engine(...)` = Engine::new, `swap_query`, `upload_seeds`, `generate_seeds`,
`seed_and_filter`, accessors. `run_pass` is seed_and_filter_all
(src/run.rs:266-495). Errors raise `HspzError` carrying the library's status
and message. There is no fallback: if libhspz.so is missing or no GPU is
visible, construction fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhspz.so")

SP_DTYPE = np.dtype([("ref_start", "<u4"), ("query_start", "<u4"), ("len", "<u4"),
                     ("score", "<i4")])

STAT = {"max_hits": 0, "launches": 1, "pipeline_syncs": 2, "reference_uploads": 3,
        "query_swaps": 4, "peak_device_bytes": 5, "gate_ns": 6, "gate_hits": 7,
        "gate_launches": 8, "gate_deferred": 9, "graph_captures": 10, "graph_replays": 11,
        "batch_prep_ns": 12, "batch_device_ns": 13, "batch_read_ns": 14, "batch_post_ns": 15,
        "index_build_ns": 16, "pos_len": 17, "index_len": 18,
        "batch_reruns": 19, "batch_passes": 20, "device_bytes": 21,
        "batch_candidates": 22, "seed_uploads": 23, "pinned_active": 24}
FLAG = {"batch_graph": 0, "batch_gate_blocks": 1, "batch_step_cells": 2, "batch_block_threads": 3,
        "batch_claim_seeds": 4, "batch_order": 5, "batch_order_used": 6, "batch_device_dedup": 7,
        "ref_records": 8, "batch_lookahead": 9, "batch_list_cap": 10, "noalloc_tile_hits": 11,
        "batch_max_positions": 12, "batch_coarse": 13, "batch_coarse_used": 14, "dump_raw": 15,
        "batch_cand_planes": 16, "batch_lane_materialize": 17,
        "batch_cand_bins": 18, "batch_cand_binned": 19, "batch_coarse_long": 20}

# Every symbol include/hspz.h declares (checked by tests/test_abi.py).
EXPORTS = [
    "hspz_engine_new", "hspz_engine_free", "hspz_swap_query", "hspz_upload_seeds",
    "hspz_generate_seeds", "hspz_seed_and_filter", "hspz_filter_output_free", "hspz_get_stat",
    "hspz_default_max_hits", "hspz_device_count", "hspz_device_memory", "hspz_last_error",
    "hspz_seed_table_build", "hspz_shape_parse", "hspz_chunk_seeds", "hspz_dedup_and_order",
    "hspz_run_pass", "hspz_free", "hspz_build_sub_mat", "hspz_filter_batch",
    "hspz_batch_output_free", "hspz_server_new", "hspz_server_submit", "hspz_server_wait",
    "hspz_request_result_free", "hspz_server_get_stats", "hspz_server_free", "hspz_set_flag",
    "hspz_get_flag", "hspz_autotune_batch", "hspz_engine_new_device_index", "hspz_engine_index",
    "hspz_swap_query_raw", "hspz_engine_query", "hspz_seed_and_filter_noalloc", "hspz_engine_share",
    "hspz_upload_seeds_async", "hspz_sync_seeds", "hspz_pinned_take", "hspz_pinned_give",
    "hspz_check_seed_bytes", "hspz_launch_overhead_ms",
]


class HspzError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__("hspz status %d: %s" % (status, message))
        self.status = status


class EngineConfig(C.Structure):
    _fields_ = [("index_table", C.c_void_p), ("index_len", C.c_size_t),
                ("pos_table", C.c_void_p), ("pos_len", C.c_size_t),
                ("ref_seq", C.c_void_p), ("ref_len", C.c_size_t),
                ("sub_mat", C.c_void_p), ("seed_size", C.c_uint32), ("xdrop", C.c_int32),
                ("hspthresh", C.c_int32), ("noentropy", C.c_int32), ("max_hits", C.c_uint32),
                ("hsp_blocks", C.c_uint32)]


class FilterOutput(C.Structure):
    _fields_ = [("hsps", C.c_void_p), ("n_hsps", C.c_size_t), ("num_hits", C.c_uint32),
                ("raw_hsps", C.c_uint32), ("raw", C.c_void_p), ("n_raw", C.c_size_t)]


class PassParams(C.Structure):
    _fields_ = [("pattern", C.c_char_p), ("transitions", C.c_int32), ("plus", C.c_int32),
                ("minus", C.c_int32), ("wga_chunk", C.c_uint32), ("interval_size", C.c_uint32)]


class PassStats(C.Structure):
    _fields_ = [("seeds", C.c_uint64), ("hits", C.c_uint64), ("raw_hsps", C.c_uint64),
                ("hsps", C.c_uint64), ("batches", C.c_uint64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class QueryChunk(C.Structure):
    _fields_ = [("fwd", C.c_void_p), ("rc", C.c_void_p), ("len", C.c_size_t)]


class BatchOutput(C.Structure):
    _fields_ = [("hsps", C.c_void_p), ("n_hsps", C.c_size_t), ("counts", C.c_void_p),
                ("chunk_intervals", C.c_void_p), ("n_chunks", C.c_size_t)]


class TuneCell(C.Structure):
    _fields_ = [("step_cells", C.c_uint32), ("block_threads", C.c_uint32),
                ("grid_blocks", C.c_uint32), ("claim_seeds", C.c_uint32), ("order", C.c_uint32),
                ("gate_ms", C.c_double), ("after_ms", C.c_double), ("hits", C.c_uint64),
                ("survivors", C.c_uint64), ("deferred", C.c_uint64), ("status", C.c_int32)]

    def as_dict(self):
        return {k: (float(getattr(self, k)) if t is C.c_double else int(getattr(self, k)))
                for k, t in self._fields_}


class ServerConfig(C.Structure):
    _fields_ = [("params", PassParams), ("hbm_budget", C.c_uint64),
                ("max_batch_requests", C.c_uint32), ("max_batch_bases", C.c_uint64)]


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("hsps", C.c_void_p), ("n_hsps", C.c_size_t),
                ("counts", C.c_void_p), ("n_intervals", C.c_size_t), ("t_submit", C.c_double),
                ("t_start", C.c_double), ("t_done", C.c_double), ("batch_id", C.c_uint32),
                ("batch_requests", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("batches", "requests", "bases", "hits", "rejected",
                                          "failed_batches", "max_batch_requests", "queued",
                                          "in_flight_bytes")]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


_lib = None


def lib():
    """Loads libhspz.so (built in-tree by hspz_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HspzError(-1, "native library missing: %s (run hspz_amd.build)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        vp, u32, i32, sz = C.c_void_p, C.c_uint32, C.c_int32, C.c_size_t
        L.hspz_last_error.restype = C.c_char_p
        L.hspz_engine_new.argtypes = [C.c_int, C.POINTER(EngineConfig), C.POINTER(vp)]
        L.hspz_engine_free.argtypes = [vp]
        L.hspz_engine_new_device_index.argtypes = [C.c_int, C.POINTER(EngineConfig), C.c_char_p, u32,
                                                   C.POINTER(vp)]
        L.hspz_engine_index.argtypes = [vp, vp, sz, C.POINTER(vp), C.POINTER(sz)]
        L.hspz_engine_free.restype = None
        L.hspz_swap_query.argtypes = [vp, vp, vp, sz]
        L.hspz_swap_query_raw.argtypes = [vp, vp, sz]
        L.hspz_engine_query.argtypes = [vp, vp, vp, sz]
        L.hspz_upload_seeds.argtypes = [vp, C.c_int, vp, sz]
        L.hspz_generate_seeds.argtypes = [vp, C.c_int, C.c_int, u32, u32, vp, u32, C.c_int,
                                          C.POINTER(u32)]
        L.hspz_seed_and_filter.argtypes = [vp, C.c_int, C.c_int, C.POINTER(FilterOutput)]
        L.hspz_seed_and_filter_noalloc.argtypes = [vp, C.c_int, C.c_int, vp, sz, C.POINTER(sz),
                                                   C.POINTER(u32), C.POINTER(u32)]
        L.hspz_filter_output_free.argtypes = [C.POINTER(FilterOutput)]
        L.hspz_filter_output_free.restype = None
        L.hspz_get_stat.argtypes = [vp, C.c_int, C.POINTER(C.c_uint64)]
        L.hspz_set_flag.argtypes = [vp, C.c_int, C.c_int]
        L.hspz_get_flag.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
        L.hspz_autotune_batch.argtypes = [vp, C.POINTER(PassParams), C.POINTER(QueryChunk), sz,
                                          u32, C.POINTER(TuneCell), sz, C.POINTER(sz),
                                          C.POINTER(TuneCell)]
        L.hspz_default_max_hits.argtypes = [C.c_uint64]
        L.hspz_default_max_hits.restype = u32
        L.hspz_device_memory.argtypes = [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.hspz_seed_table_build.argtypes = [vp, sz, C.c_char_p, u32, C.c_int, vp, sz,
                                            C.POINTER(vp), C.POINTER(sz)]
        L.hspz_shape_parse.argtypes = [C.c_char_p, C.POINTER(u32), C.POINTER(u32), vp]
        L.hspz_chunk_seeds.argtypes = [vp, sz, C.c_char_p, C.c_int, u32, u32, C.c_int, vp,
                                       C.POINTER(sz)]
        L.hspz_dedup_and_order.argtypes = [vp, sz, C.POINTER(sz)]
        L.hspz_run_pass.argtypes = [vp, C.POINTER(PassParams), C.POINTER(vp), C.POINTER(vp),
                                    C.POINTER(sz), C.POINTER(PassStats)]
        L.hspz_build_sub_mat.argtypes = [C.c_char_p, i32, vp, vp]
        L.hspz_filter_batch.argtypes = [vp, C.POINTER(PassParams), C.POINTER(QueryChunk), sz,
                                        C.POINTER(BatchOutput), C.POINTER(PassStats)]
        L.hspz_batch_output_free.argtypes = [C.POINTER(BatchOutput)]
        L.hspz_batch_output_free.restype = None
        L.hspz_server_new.argtypes = [C.POINTER(vp), sz, C.POINTER(ServerConfig), C.POINTER(vp)]
        L.hspz_server_submit.argtypes = [vp, vp, vp, sz, C.POINTER(C.c_uint64)]
        L.hspz_server_wait.argtypes = [vp, C.c_uint64, C.POINTER(RequestResult)]
        L.hspz_request_result_free.argtypes = [C.POINTER(RequestResult)]
        L.hspz_request_result_free.restype = None
        L.hspz_server_get_stats.argtypes = [vp, C.POINTER(ServerStats)]
        L.hspz_server_free.argtypes = [vp]
        L.hspz_server_free.restype = None
        L.hspz_free.argtypes = [vp]
        L.hspz_free.restype = None
        L.hspz_engine_share.argtypes = [vp, C.POINTER(vp)]
        L.hspz_upload_seeds_async.argtypes = [vp, C.c_int, vp, sz]
        L.hspz_sync_seeds.argtypes = [vp, C.c_int]
        L.hspz_pinned_take.argtypes = [vp, sz, C.POINTER(vp)]
        L.hspz_pinned_give.argtypes = [vp, vp]
        L.hspz_check_seed_bytes.argtypes = [vp, C.c_int, vp, sz]
        L.hspz_launch_overhead_ms.argtypes = [vp, u32, C.POINTER(C.c_double)]
        L.hspz_test_set_read_piece.argtypes = [vp, sz]
        _lib = L
    return _lib


def _check(status: int):
    if status != 0:
        msg = lib().hspz_last_error()
        raise HspzError(status, msg.decode() if msg else "")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ------------------------------------------------------------------ host side

def shape_parse(pattern: str = "12of19"):
    size, k = C.c_uint32(), C.c_uint32()
    pos = np.zeros(32, dtype=np.uint32)
    _check(lib().hspz_shape_parse(pattern.encode(), C.byref(size), C.byref(k), _ptr(pos)))
    return int(size.value), int(k.value), pos[:k.value].copy()


def seed_table_build(raw: bytes, pattern: str = "12of19", step: int = 1, threads: int = 8):
    """SeedTable::build_parallel (seed.rs:207-361): (index_table, pos_table)."""
    _, k, _ = shape_parse(pattern)
    a = np.frombuffer(raw, dtype=np.uint8)
    index = np.zeros(1 << (2 * k), dtype=np.uint32)
    pp, n = C.c_void_p(), C.c_size_t()
    _check(lib().hspz_seed_table_build(_ptr(a), len(a), pattern.encode(), step, threads,
                                       _ptr(index), len(index), C.byref(pp), C.byref(n)))
    pos = np.empty(n.value, dtype=np.uint32)
    if n.value:
        C.memmove(_ptr(pos), pp, n.value * 4)
    lib().hspz_free(pp)
    return index, pos


def chunk_seeds(raw: bytes, lo: int, hi: int, pattern: str = "12of19", transitions=True,
                threads: int = 1) -> np.ndarray:
    _, k, _ = shape_parse(pattern)
    a = np.frombuffer(raw, dtype=np.uint8)
    out = np.zeros(max(hi - lo, 1) * (1 + k), dtype=np.uint64)
    n = C.c_size_t()
    _check(lib().hspz_chunk_seeds(_ptr(a), len(a), pattern.encode(), int(transitions), lo, hi,
                                  threads, _ptr(out), C.byref(n)))
    return out[:n.value].copy()


def build_sub_mat(ambiguous: str = "", xdrop: int = 910, acgt=None) -> np.ndarray:
    """scoring.rs:41-120; acgt = 16 ints (row-major ACGT block) or None for HOXD70."""
    out = np.zeros(64, dtype=np.int32)
    blk = None if acgt is None else np.ascontiguousarray(acgt, dtype=np.int32)
    _check(lib().hspz_build_sub_mat(ambiguous.encode(), xdrop,
                                    None if blk is None else _ptr(blk), _ptr(out)))
    return out


def dedup_and_order(rows: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(rows, dtype=SP_DTYPE).copy()
    n = C.c_size_t()
    _check(lib().hspz_dedup_and_order(_ptr(a), len(a), C.byref(n)))
    return a[:n.value].copy()


def default_max_hits(total_bytes: int) -> int:
    return int(lib().hspz_default_max_hits(total_bytes))


def device_count() -> int:
    return int(lib().hspz_device_count())


# ------------------------------------------------------------------ engine

class Engine:
    """Reference-resident engine on one device (gpu/mod.rs:1024-1128)."""

    def __init__(self, index_table, pos_table, ref_enc: np.ndarray,
                 sub_mat: np.ndarray, seed_size: int = 19, xdrop: int = 910,
                 hspthresh: int = 3000, noentropy: bool = False, max_hits: int = 0,
                 device: int = 0, device_index: str = None, index_step: int = 1):
        """index_table/pos_table: the host index (hspz_engine_new); or None and
        device_index = the seed pattern, to build the index on the device
        (hspz_engine_new_device_index)."""
        ref = np.ascontiguousarray(ref_enc, dtype=np.uint8)
        sub = np.ascontiguousarray(sub_mat, dtype=np.int32)
        if sub.size != 64:
            raise ValueError("sub_mat must have 64 entries")
        h = C.c_void_p()
        if device_index is not None:
            cfg = EngineConfig(None, 0, None, 0, _ptr(ref).value, len(ref), _ptr(sub).value, seed_size,
                               xdrop, hspthresh, int(noentropy), max_hits, 0)
            _check(lib().hspz_engine_new_device_index(device, C.byref(cfg), device_index.encode(),
                                                      index_step, C.byref(h)))
        else:
            idx = np.ascontiguousarray(index_table, dtype=np.uint32)
            pos = np.ascontiguousarray(pos_table, dtype=np.uint32)
            cfg = EngineConfig(_ptr(idx).value, len(idx), _ptr(pos).value if len(pos) else None,
                               len(pos), _ptr(ref).value, len(ref), _ptr(sub).value, seed_size,
                               xdrop, hspthresh, int(noentropy), max_hits, 0)
            _check(lib().hspz_engine_new(device, C.byref(cfg), C.byref(h)))
        self._h = h
        self.seed_size = seed_size

    def index(self):
        """The engine's resident index copied to the host: (index_table, pos_table)."""
        idx = np.zeros(self.stat("index_len"), dtype=np.uint32)
        p, ln = C.c_void_p(), C.c_size_t()
        _check(lib().hspz_engine_index(self._h, _ptr(idx), len(idx), C.byref(p), C.byref(ln)))
        pos = np.zeros(ln.value, dtype=np.uint32)
        if ln.value:
            C.memmove(_ptr(pos), p, ln.value * 4)
        lib().hspz_free(p)
        return idx, pos

    def close(self):
        if getattr(self, "_h", None):
            lib().hspz_engine_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def swap_query(self, fwd_enc: np.ndarray, rc_enc: np.ndarray):
        f = np.ascontiguousarray(fwd_enc, dtype=np.uint8)
        r = np.ascontiguousarray(rc_enc, dtype=np.uint8)
        if len(f) != len(r):
            raise ValueError("forward and reverse-complement blocks differ in length")
        _check(lib().hspz_swap_query(self._h, _ptr(f), _ptr(r), len(f)))
        self._query_len = len(f)

    def swap_query_raw(self, raw):
        """hspz_swap_query_raw: the raw query block; both strands are encoded
        on the device."""
        a = _u8(raw)
        _check(lib().hspz_swap_query_raw(self._h, _ptr(a), len(a)))
        self._query_len = len(a)

    def query(self):
        """The engine's encoded (fwd, rc) query strands, copied back."""
        n = int(getattr(self, "_query_len", 0))
        f, r = np.empty(n, dtype=np.uint8), np.empty(n, dtype=np.uint8)
        _check(lib().hspz_engine_query(self._h, _ptr(f), _ptr(r), n))
        return f, r

    def upload_seeds(self, slot: int, seeds: np.ndarray):
        s = np.ascontiguousarray(seeds, dtype=np.uint64)
        _check(lib().hspz_upload_seeds(self._h, slot, _ptr(s), len(s)))

    def generate_seeds(self, slot: int, rev: bool, start: int, end: int,
                       pattern: str = "12of19", transitions: bool = True) -> int:
        size, k, pos = shape_parse(pattern)
        n = C.c_uint32()
        _check(lib().hspz_generate_seeds(self._h, slot, int(rev), start, end, _ptr(pos), k,
                                         int(transitions), C.byref(n)))
        return int(n.value)

    def seed_and_filter(self, slot: int, rev: bool, with_raw: bool = False):
        """Returns (hsps structured array, num_hits, raw_hsps), and with
        with_raw the pre-dedup records as a fourth item (dump_raw must be set,
        else it is empty)."""
        out = FilterOutput()
        _check(lib().hspz_seed_and_filter(self._h, slot, int(rev), C.byref(out)))
        res = np.zeros(out.n_hsps, dtype=SP_DTYPE)
        if out.n_hsps:
            C.memmove(_ptr(res), out.hsps, out.n_hsps * 16)
        raw = np.zeros(out.n_raw, dtype=SP_DTYPE)
        if out.n_raw:
            C.memmove(_ptr(raw), out.raw, out.n_raw * 16)
        lib().hspz_filter_output_free(C.byref(out))
        if with_raw:
            return res, int(out.num_hits), int(out.raw_hsps), raw
        return res, int(out.num_hits), int(out.raw_hsps)

    def upload_seeds_async(self, slot: int, seeds):
        """hspz_upload_seeds_async: `seeds` must stay alive and unchanged until
        the slot's copy completes (the next seed_and_filter, generate_seeds or
        upload on the slot, or sync_seeds(slot)). Pass a PinnedSeeds array for
        a truly asynchronous copy."""
        s = seeds if isinstance(seeds, np.ndarray) and seeds.dtype == np.uint64 and seeds.flags.c_contiguous \
            else np.ascontiguousarray(seeds, dtype=np.uint64)
        _check(lib().hspz_upload_seeds_async(self._h, slot, _ptr(s), len(s)))
        self._inflight = getattr(self, "_inflight", {})
        self._inflight[slot] = s  # kept alive until the slot's copy is known complete

    def sync_seeds(self, slot: int):
        _check(lib().hspz_sync_seeds(self._h, slot))

    def pinned_take(self, n: int) -> np.ndarray:
        """hspz_pinned_take: a uint64 array of n seeds over a pinned pool buffer
        (valid until pinned_give or the engine is closed)."""
        p = C.c_void_p()
        _check(lib().hspz_pinned_take(self._h, n, C.byref(p)))
        buf = (C.c_uint64 * max(n, 1)).from_address(p.value)
        return np.ctypeslib.as_array(buf)[:n]

    def pinned_give(self, arr: np.ndarray):
        _check(lib().hspz_pinned_give(self._h, C.c_void_p(arr.ctypes.data)))

    def check_seed_bytes(self, slot: int, expected):
        e = np.ascontiguousarray(expected, dtype=np.uint64)
        _check(lib().hspz_check_seed_bytes(self._h, slot, _ptr(e), len(e)))

    def launch_overhead_ms(self, iters: int = 1000) -> float:
        ms = C.c_double()
        _check(lib().hspz_launch_overhead_ms(self._h, iters, C.byref(ms)))
        return float(ms.value)

    def share(self) -> "Engine":
        """hspz_engine_share: a second engine over this engine's resident target."""
        h = C.c_void_p()
        _check(lib().hspz_engine_share(self._h, C.byref(h)))
        e = Engine.__new__(Engine)
        e._h = h
        e.seed_size = self.seed_size
        return e

    def seed_and_filter_noalloc(self, slot: int, rev: bool, cap: int):
        """hspz_seed_and_filter_noalloc into a caller array of cap records.
        Returns (hsps structured array, num_hits, raw_hsps)."""
        out = np.zeros(max(cap, 1), dtype=SP_DTYPE)
        n, nh, raw = C.c_size_t(), C.c_uint32(), C.c_uint32()
        _check(lib().hspz_seed_and_filter_noalloc(self._h, slot, int(rev), _ptr(out), cap, C.byref(n),
                                                  C.byref(nh), C.byref(raw)))
        return out[:n.value].copy(), int(nh.value), int(raw.value)

    def set_flag(self, name: str, value: int):
        _check(lib().hspz_set_flag(self._h, FLAG[name], int(value)))

    def get_flag(self, name: str) -> int:
        v = C.c_int()
        _check(lib().hspz_get_flag(self._h, FLAG[name], C.byref(v)))
        return int(v.value)

    def autotune(self, chunks, reps: int = 1, pattern: str = "12of19", transitions: bool = True,
                 plus: bool = True, minus: bool = True, wga_chunk: int = 250000,
                 interval_size: int = 10000000):
        """hspz_autotune_batch over chunks = [(fwd_enc, rc_enc)]: measures every
        cell of the batched gate's launch grid, keeps the fastest on the engine.
        Returns (cells, best) as dicts."""
        arr, keep = _chunk_array(chunks)
        p = PassParams(pattern.encode(), int(transitions), int(plus), int(minus), wga_chunk,
                       interval_size)
        cap = 64
        cells = (TuneCell * cap)()
        n, best = C.c_size_t(), TuneCell()
        _check(lib().hspz_autotune_batch(self._h, C.byref(p), arr, len(chunks), int(reps), cells,
                                         cap, C.byref(n), C.byref(best)))
        del keep
        return [cells[i].as_dict() for i in range(min(n.value, cap))], best.as_dict()

    def stat(self, name: str) -> int:
        v = C.c_uint64()
        _check(lib().hspz_get_stat(self._h, STAT[name], C.byref(v)))
        return int(v.value)

    def run_pass(self, pattern: str = "12of19", transitions: bool = True, plus: bool = True,
                 minus: bool = True, wga_chunk: int = 250000, interval_size: int = 10000000):
        """seed_and_filter_all (run.rs:266-495). Returns ([(fw, rc)], stats)."""
        p = PassParams(pattern.encode(), int(transitions), int(plus), int(minus), wga_chunk,
                       interval_size)
        hs, cnt, ni, st = C.c_void_p(), C.c_void_p(), C.c_size_t(), PassStats()
        _check(lib().hspz_run_pass(self._h, C.byref(p), C.byref(hs), C.byref(cnt),
                                   C.byref(ni), C.byref(st)))
        counts = (np.ctypeslib.as_array(C.cast(cnt, C.POINTER(C.c_size_t)),
                                        (max(2 * ni.value, 1),))[:2 * ni.value].copy())
        total = int(counts.sum())
        allsp = np.zeros(total, dtype=SP_DTYPE)
        if total:
            C.memmove(_ptr(allsp), hs, total * 16)
        lib().hspz_free(hs)
        lib().hspz_free(cnt)
        res, at = [], 0
        for i in range(ni.value):
            fw = allsp[at:at + counts[2 * i]]
            at += counts[2 * i]
            rc = allsp[at:at + counts[2 * i + 1]]
            at += counts[2 * i + 1]
            res.append((fw, rc))
        return res, st.as_dict()

    def filter_batch(self, chunks, pattern: str = "12of19", transitions: bool = True,
                     plus: bool = True, minus: bool = True, wga_chunk: int = 250000,
                     interval_size: int = 10000000):
        """Batched path over many query blocks: chunks = [(fwd_enc, rc_enc)].
        Returns ([per chunk: [(fw, rc) per interval]], stats); each chunk's
        lists equal run_pass on that chunk alone."""
        arr, keep = _chunk_array(chunks)
        p = PassParams(pattern.encode(), int(transitions), int(plus), int(minus), wga_chunk,
                       interval_size)
        out, st = BatchOutput(), PassStats()
        _check(lib().hspz_filter_batch(self._h, C.byref(p), arr, len(chunks), C.byref(out),
                                       C.byref(st)))
        try:
            nci = (np.ctypeslib.as_array(C.cast(out.chunk_intervals, C.POINTER(C.c_size_t)),
                                         (max(out.n_chunks, 1),))[:out.n_chunks].copy())
            nl = 2 * int(nci.sum())
            counts = (np.ctypeslib.as_array(C.cast(out.counts, C.POINTER(C.c_size_t)),
                                            (max(nl, 1),))[:nl].copy())
            allsp = np.zeros(out.n_hsps, dtype=SP_DTYPE)
            if out.n_hsps:
                C.memmove(_ptr(allsp), out.hsps, out.n_hsps * 16)
        finally:
            lib().hspz_batch_output_free(C.byref(out))
        res, at, li = [], 0, 0
        for c in range(len(chunks)):
            ivs = []
            for _ in range(int(nci[c])):
                fw = allsp[at:at + counts[li]]
                at += counts[li]
                rc = allsp[at:at + counts[li + 1]]
                at += counts[li + 1]
                li += 2
                ivs.append((fw, rc))
            res.append(ivs)
        return res, st.as_dict()


def _u8(x) -> np.ndarray:
    if isinstance(x, (bytes, bytearray, memoryview)):
        return np.frombuffer(bytes(x), dtype=np.uint8)
    return np.ascontiguousarray(x, dtype=np.uint8)


def _chunk_array(chunks):
    """QueryChunk[] over [(fwd_enc, rc_enc)] or [(raw, None)] (the device then
    encodes both strands), plus the arrays it points into."""
    keep = []
    arr = (QueryChunk * max(len(chunks), 1))()
    for i, (f, r) in enumerate(chunks):
        f = _u8(f)
        keep.append(f)
        if r is None:
            arr[i] = QueryChunk(_ptr(f).value, None, len(f))
            continue
        r = _u8(r)
        if len(f) != len(r):
            raise ValueError("chunk %d: forward and reverse-complement lengths differ" % i)
        keep.append(r)
        arr[i] = QueryChunk(_ptr(f).value, _ptr(r).value, len(f))
    return arr, keep


class Server:
    """Continuous-batching loop (hspz_server_*) over one or more engines that
    hold the same reference on the same device."""

    def __init__(self, engines, hbm_budget: int = 0, max_batch_requests: int = 0,
                 max_batch_bases: int = 0, pattern: str = "12of19", transitions: bool = True,
                 plus: bool = True, minus: bool = True, wga_chunk: int = 250000,
                 interval_size: int = 10000000):
        self._engines = list(engines)  # keep the handles alive
        self._pattern = pattern.encode()
        arr = (C.c_void_p * len(self._engines))(*[e._h.value for e in self._engines])
        cfg = ServerConfig(PassParams(self._pattern, int(transitions), int(plus), int(minus),
                                      wga_chunk, interval_size),
                           hbm_budget, max_batch_requests, max_batch_bases)
        h = C.c_void_p()
        _check(lib().hspz_server_new(arr, len(self._engines), C.byref(cfg), C.byref(h)))
        self._h = h

    def submit(self, fwd_enc, rc_enc=None) -> int:
        """(fwd_enc, rc_enc) encoded strands, or (raw, None): the raw block."""
        f = _u8(fwd_enc)
        r = None if rc_enc is None else _u8(rc_enc)
        rid = C.c_uint64()
        _check(lib().hspz_server_submit(self._h, _ptr(f), None if r is None else _ptr(r), len(f),
                                        C.byref(rid)))
        return int(rid.value)

    def wait(self, rid: int):
        """Returns ([(fw, rc)] per interval, timing dict)."""
        out = RequestResult()
        st = lib().hspz_server_wait(self._h, rid, C.byref(out))
        try:
            _check(st)
            counts = (np.ctypeslib.as_array(C.cast(out.counts, C.POINTER(C.c_size_t)),
                                            (max(2 * out.n_intervals, 1),))[:2 * out.n_intervals].copy())
            allsp = np.zeros(out.n_hsps, dtype=SP_DTYPE)
            if out.n_hsps:
                C.memmove(_ptr(allsp), out.hsps, out.n_hsps * 16)
            timing = {"t_submit": out.t_submit, "t_start": out.t_start, "t_done": out.t_done,
                      "batch_id": int(out.batch_id), "batch_requests": int(out.batch_requests)}
        finally:
            lib().hspz_request_result_free(C.byref(out))
        res, at = [], 0
        for i in range(len(counts) // 2):
            fw = allsp[at:at + counts[2 * i]]
            at += counts[2 * i]
            rc = allsp[at:at + counts[2 * i + 1]]
            at += counts[2 * i + 1]
            res.append((fw, rc))
        return res, timing

    def stats(self):
        st = ServerStats()
        _check(lib().hspz_server_get_stats(self._h, C.byref(st)))
        return st.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            lib().hspz_server_free(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hash_pass(intervals) -> int:
    """benchmark.rs:401-421: FNV-1a over (ref_start, query_start, len, score)
    little-endian, intervals in order, fw then rc."""
    h = 0xcbf29ce484222325
    for fw, rc in intervals:
        for arr in (fw, rc):
            b = np.ascontiguousarray(arr, dtype=SP_DTYPE).tobytes()
            for byte in b:
                h ^= byte
                h = (h * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h

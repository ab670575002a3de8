"""Synthetic workloads (include/hspz_synth.h) shaped like the benchmark
configurations: a 41%-GC iid target, a mutated query with reverse-complemented
blocks, interspersed soft-masked repeats, and lognormal query chunks.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .engine import lib

GC = 0.41


def _L():
    L = lib()
    if not getattr(L, "_synth_ready", False):
        vp, sz, u64, dbl = C.c_void_p, C.c_size_t, C.c_uint64, C.c_double
        L.hspz_synth_random.argtypes = [u64, sz, dbl, vp]
        L.hspz_synth_random.restype = None
        L.hspz_synth_mutate.argtypes = [vp, sz, u64, dbl, dbl, dbl, vp, sz]
        L.hspz_synth_mutate.restype = sz
        L.hspz_synth_revcomp.argtypes = [vp, sz]
        L.hspz_synth_revcomp.restype = None
        L.hspz_synth_revcomp_blocks.argtypes = [vp, sz, sz]
        L.hspz_synth_revcomp_blocks.restype = None
        L.hspz_synth_repeats.argtypes = [vp, sz, u64, C.c_uint32, C.c_uint32, dbl, dbl, dbl,
                                         C.c_int, dbl]
        L.hspz_synth_repeats.restype = sz
        L.hspz_synth_lognormal.argtypes = [u64, sz, dbl, dbl, dbl, dbl, vp]
        L.hspz_synth_lognormal.restype = sz
        L._synth_ready = True
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def random_bases(seed: int, n: int, gc: float = GC) -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    _L().hspz_synth_random(seed, n, gc, _p(out))
    return out


def mutate(src: np.ndarray, seed: int, sub: float, indel: float, gc: float = GC) -> np.ndarray:
    src = np.ascontiguousarray(src, dtype=np.uint8)
    cap = int(len(src) * (1 + 4 * indel)) + 1024
    out = np.empty(cap, dtype=np.uint8)
    n = _L().hspz_synth_mutate(_p(src), len(src), seed, sub, indel, gc, _p(out), cap)
    return out[:n].copy()


def revcomp(seq: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(seq, dtype=np.uint8).copy()
    _L().hspz_synth_revcomp(_p(s), len(s))
    return s


def revcomp_blocks(seq: np.ndarray, block: int) -> np.ndarray:
    s = np.ascontiguousarray(seq, dtype=np.uint8).copy()
    _L().hspz_synth_revcomp_blocks(_p(s), len(s), block)
    return s


def add_repeats(seq: np.ndarray, seed: int, families: int = 50, consensus_len: int = 300,
                div_lo: float = 0.10, div_hi: float = 0.20, fraction: float = 0.10,
                lowercase: bool = True) -> np.ndarray:
    s = np.ascontiguousarray(seq, dtype=np.uint8).copy()
    _L().hspz_synth_repeats(_p(s), len(s), seed, families, consensus_len, div_lo, div_hi,
                            fraction, int(lowercase), GC)
    return s


def lognormal_lengths(seed: int, count: int, median: float, sigma: float, lo: float,
                      hi: float) -> np.ndarray:
    out = np.empty(count, dtype=np.uint64)
    _L().hspz_synth_lognormal(seed, count, median, sigma, lo, hi, _p(out))
    return out


def c1_pair(target_len: int = 4_000_000, seed: int = 0xC0FFEE01):
    """C1 shape: iid 41%-GC target; query = target with 15% substitutions
    (ts:tv 2:1) and 1% indels (mean length 3), reverse-complemented in
    alternating 100 kbp blocks. Returns (target, query) raw bytes."""
    t = random_bases(seed, target_len)
    q = mutate(t, seed ^ 0x5EED, 0.15, 0.01)
    q = revcomp_blocks(q, 100_000)
    return t.tobytes(), q.tobytes()


def c2_pair(target_len: int = 100_000_000, seed: int = 0xC0FFEE02):
    """C2 shape: C1 plus 10% soft-masked interspersed repeats (50 families,
    300 bp consensus, 10-20% divergent copies); query is the C1-style mutation
    of the whole target with lowercase preserved."""
    t = random_bases(seed, target_len)
    t = add_repeats(t, seed ^ 0xAB, 50, 300, 0.10, 0.20, 0.10, True)
    q = mutate(t, seed ^ 0x5EED, 0.15, 0.01)
    q = revcomp_blocks(q, 100_000)
    return t.tobytes(), q.tobytes()


C4_RECORDS, C4_RECORD_LEN = 20, 150_000_000


def c4_target(seed: int = 0xC0FFEE04, records: int = C4_RECORDS, record_len: int = C4_RECORD_LEN):
    """C4 shape, mammal-scale target: `records` x `record_len` bases (20 x
    150 Mbp = 3 Gbp), 41% GC, 45% repeat-derived (500 families of 300 bp
    consensus, copies 10-20% divergent), all uppercase so that every window
    seeds (about 3e9 indexed positions). Returns one array of records *
    record_len bases; record r is [r * record_len, (r + 1) * record_len)."""
    t = random_bases(seed, records * record_len)
    return add_repeats(t, seed ^ 0xAB, 500, 300, 0.10, 0.20, 0.45, False)


def c4_query(target: np.ndarray, seed: int, length: int = 50_000_000, record_len: int = C4_RECORD_LEN):
    """One C4 query record: `length` bases sampled from one target record at
    12% divergence (substitutions 12%, indels 0.8%); then segmental
    duplications (10-100 kbp copies of other target regions at divergence
    U(1%, 5%), written over the query until they make 10% of it) and
    reverse-complemented segments (10-100 kbp, reversed in place until 10% of
    the query). Returns raw bytes."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nrec = len(target) // record_len
    r = int(rng.integers(0, nrec))
    start = r * record_len + int(rng.integers(0, record_len - length + 1))
    q = mutate(target[start:start + length], seed ^ 0x5EED, 0.12, 0.008)
    covered = 0
    while covered < 0.10 * len(q):
        n = int(rng.integers(10_000, 100_001))
        src = int(rng.integers(0, len(target) - n))
        d = float(rng.uniform(0.01, 0.05))
        seg = mutate(target[src:src + n], int(rng.integers(0, 1 << 62)), d, d / 15.0)
        at = int(rng.integers(0, len(q) - len(seg)))
        q[at:at + len(seg)] = seg
        covered += len(seg)
    covered = 0
    while covered < 0.10 * len(q):
        n = int(rng.integers(10_000, 100_001))
        at = int(rng.integers(0, len(q) - n))
        q[at:at + n] = revcomp(q[at:at + n])
        covered += n
    return q.tobytes()


def sample_chunks(target: np.ndarray, seed: int, count: int, median: float, sigma: float,
                  lo: float, hi: float, div_lo: float, div_hi: float):
    """C3/C5 query chunks: clipped lognormal lengths, uniform start,
    divergence d ~ U(div_lo, div_hi) (substitutions d, indels d/15), half of
    them reverse-complemented. Returns a list of raw byte strings."""
    t = np.ascontiguousarray(target, dtype=np.uint8)
    lens = lognormal_lengths(seed, count, median, sigma, lo, hi)
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i, L in enumerate(lens):
        L = int(min(L, len(t) - 1))
        start = int(rng.integers(0, len(t) - L))
        d = float(rng.uniform(div_lo, div_hi))
        q = mutate(t[start:start + L], seed * 7919 + i, d, d / 15.0)
        if rng.uniform() < 0.5:
            q = revcomp(q)
        out.append(q.tobytes())
    return out

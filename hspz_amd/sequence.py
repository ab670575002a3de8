"""Host-side sequence handling: FASTA records, packing, reverse complement,
encoding and `.segments` rendering.

Mirrors the reference's src/sequence.rs (pack :185-205, reverse_complement
:212-230, encode :66-79, chr_at :316-318, intervals :322-334) and the
`.segments` writer in src/hsp.rs:179-292. These run once per input block and
are not on the timed path; the per-batch work lives in the native library.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

SEP = ord("&")

# Scoring alphabet (sequence.rs:35-44).
A_NT, C_NT, G_NT, T_NT, L_NT, N_NT, X_NT, E_NT = range(8)

_ENCODE = np.full(256, X_NT, dtype=np.uint8)
for _ch, _code in ((b"A", A_NT), (b"C", C_NT), (b"G", G_NT), (b"T", T_NT)):
    _ENCODE[_ch[0]] = _code
for _ch in b"acgt":
    _ENCODE[_ch] = L_NT
_ENCODE[ord("n")] = N_NT
_ENCODE[ord("N")] = N_NT
_ENCODE[SEP] = E_NT


def _complement_table() -> bytes:
    # IUPAC complement pairs; unknown uppercase -> 'N', unknown lowercase -> 'n',
    # '&' to itself, every other byte (including >= 128) -> 'N'
    # (sequence.rs:83-86).
    t = bytearray(b"N" * 256)
    for c in range(ord("a"), ord("z") + 1):
        t[c] = ord("n")
    t[SEP] = SEP
    pairs = "ATBVCGDHGCHDKMMKRYSSTAUAVBWWYR"
    for i in range(0, len(pairs), 2):
        src, dst = ord(pairs[i]), ord(pairs[i + 1])
        t[src] = dst
        t[src + 32] = dst + 32
    return bytes(t)


_RC = _complement_table()


@dataclass
class Chr:
    name: str
    start: int
    len: int


# ASCII whitespace as Rust's u8::is_ascii_whitespace defines it (no \x0b).
_WS = b" \t\n\x0c\r"


def read_fasta(path: str) -> list[tuple[str, bytes]]:
    """sequence.rs:346-414: name = first whitespace token after '>' (invalid
    UTF-8 replaced), whitespace dropped from sequence lines, lines before the
    first header ignored."""
    records: list[tuple[str, bytes]] = []
    name, parts = None, []
    with open(path, "rb") as fh:
        data = fh.read()
    for line in data.split(b"\n"):
        if line.startswith(b">"):
            if name is not None:
                records.append((name, b"".join(parts)))
            head = line[1:]
            for w in _WS:
                head = head.replace(bytes([w]), b" ")
            toks = [t for t in head.split(b" ") if t]
            name, parts = (toks[0].decode("utf-8", "replace") if toks else ""), []
        elif name is not None:
            parts.append(line.translate(None, _WS))
    if name is not None:
        records.append((name, b"".join(parts)))
    return records


def pack(records, prefix: str = ""):
    """Joins records with '&'; the trailing separator is in the buffer but not
    in block_len. Returns (buf, chrs, block_len)."""
    buf = bytearray()
    chrs = []
    for name, seq in records:
        seq = seq.encode() if isinstance(seq, str) else bytes(seq)
        chrs.append(Chr(prefix + name, len(buf), len(seq)))
        buf += seq
        buf.append(SEP)
    block_len = max(len(buf) - 1, 0)
    return bytes(buf), chrs, block_len


def reverse_complement(buf: bytes, chrs: list[Chr], block_len: int):
    rc = bytes(buf[:block_len])[::-1].translate(_RC)
    rc_chrs = [Chr(c.name, block_len - c.start - c.len, c.len) for c in reversed(chrs)]
    return rc, rc_chrs


def encode(raw: bytes) -> np.ndarray:
    return _ENCODE[np.frombuffer(raw, dtype=np.uint8)]


def intervals(block_len: int, seed_size: int, interval_size: int):
    out = []
    if block_len <= seed_size:
        return out
    end_pos = block_len - seed_size
    curr = 0
    while curr < end_pos:
        out.append((curr, min(end_pos, curr + interval_size)))
        curr += interval_size
    return out


def chr_at(chrs: list[Chr], off: int) -> int:
    lo, hi = 0, len(chrs)
    while lo < hi:
        mid = (lo + hi) // 2
        if chrs[mid].start <= off:
            lo = mid + 1
        else:
            hi = mid
    return lo - 1


def render_segments(hsps, r_chrs, q_chrs, strand: str) -> str:
    """hsps: iterable of (ref_start, query_start, len, score). The minus list is
    printed in reverse (hsp.rs:213-218)."""
    rows = list(hsps)
    if strand == "-":
        rows = rows[::-1]
    out = []
    for rs, qs, ln, score in rows:
        r = r_chrs[chr_at(r_chrs, rs)]
        q = q_chrs[chr_at(q_chrs, qs)]
        out.append("%s\t%d\t%d\t%s\t%d\t%d\t%s\t%d\n" % (
            r.name, rs + 1 - r.start, rs + ln + 1 - r.start,
            q.name, qs + 1 - q.start, qs + ln + 1 - q.start, strand, score))
    return "".join(out)


def segment_files(intervals_out, r_chrs, q_chrs, rc_chrs, query_bin=0, ref_bin=0):
    """{file name: text} as run.rs:577-606 names them; empty lists skipped."""
    files = {}
    for n, (fw, rc) in enumerate(intervals_out):
        base = "tmp%d.block%d.r%d" % (n + 1, query_bin, ref_bin)
        for hsps, chrs, strand in ((fw, q_chrs, "+"), (rc, rc_chrs, "-")):
            if len(hsps) == 0:
                continue
            stem = base + (".minus" if strand == "-" else ".plus")
            files[stem + ".segments"] = render_segments(hsps, r_chrs, chrs, strand)
    return files

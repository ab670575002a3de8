"""Shared parity-test plumbing: build one pair (reference block, query block),
run it through the oracle and through the native engine."""
from __future__ import annotations

import hashlib
import os

import numpy as np

import oracle_ffi as O
from hspz_amd import sequence as S

FIXTURES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures")


class Pair:
    def __init__(self, ref_records, qry_records, acgt=None, ambiguous="", xdrop=910,
                 hspthresh=3000, noentropy=False, max_hits=16711680, seed="12of19"):
        self.rbuf, self.rchrs, self.rbl = S.pack(ref_records)
        self.qbuf, self.qchrs, self.qbl = S.pack(qry_records)
        self.ref_raw = self.rbuf[:self.rbl]
        self.q_raw = self.qbuf[:self.qbl]
        self.q_rc_raw, self.rcchrs = S.reverse_complement(self.qbuf, self.qchrs, self.qbl)
        self.sub = O.sub_mat(ambiguous, xdrop, acgt)
        self.xdrop, self.hspthresh, self.noentropy = xdrop, hspthresh, noentropy
        self.max_hits, self.seed = max_hits, seed
        self._oracle = None

    @classmethod
    def fixture(cls, name, **kw):
        return cls(S.read_fasta(os.path.join(FIXTURES, name + ".ref.fa")),
                   S.read_fasta(os.path.join(FIXTURES, name + ".qry.fa")), **kw)

    @classmethod
    def raw(cls, ref: bytes, qry: bytes, **kw):
        return cls([("target", ref)], [("query", qry)], **kw)

    def oracle(self):
        if self._oracle is None:
            self._oracle = O.Oracle(self.ref_raw, self.q_raw, self.q_rc_raw, self.sub,
                                    seed=self.seed, xdrop=self.xdrop, hspthresh=self.hspthresh,
                                    noentropy=self.noentropy, max_hits=self.max_hits)
        return self._oracle

    def engine(self):
        from hspz_amd import engine as E
        o = self.oracle()
        size, _, _ = E.shape_parse(self.seed)
        eng = E.Engine(o.index, o.pos, o.ref_enc, self.sub, seed_size=size, xdrop=self.xdrop,
                       hspthresh=self.hspthresh, noentropy=self.noentropy,
                       max_hits=self.max_hits)
        eng.swap_query(o.q_enc, o.q_rc_enc)
        return eng

    def segments_digest(self, intervals):
        rows = [([tuple(int(x) for x in r) for r in fw], [tuple(int(x) for x in r) for r in rc])
                for fw, rc in intervals]
        files = S.segment_files(rows, self.rchrs, self.qchrs, self.rcchrs)
        lines = []
        for t in files.values():
            lines += t.splitlines(keepends=True)
        lines.sort(key=lambda s: s.encode())
        return len(lines), len(files), hashlib.sha256("".join(lines).encode()).hexdigest()[:16]


def same_intervals(a, b):
    if len(a) != len(b):
        return False
    for (fa, ra), (fb, rb) in zip(a, b):
        if not (np.array_equal(np.asarray(fa), np.asarray(fb)) and
                np.array_equal(np.asarray(ra), np.asarray(rb))):
            return False
    return True

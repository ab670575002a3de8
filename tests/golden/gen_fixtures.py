#!/usr/bin/env python3
"""Writes the ten tiny parity fixtures as FASTA files.

Restates the fixture recipe of the reference's assets/tests/gen_fixtures.py:32-94
(stdlib ``random.Random(20260811)``; a 600 bp random base drawn first, then the
20 bp repeat unit, then the 60 bp divergent insert; 60-column FASTA). Two of
these fixtures carry frozen answers published by the reference
(assets/container/smoke.sh:11-12,78-79): ``repeat`` gives 57 HSPs with sorted
digest d01edd2118cf5fa5 and ``multi`` gives 3 HSPs in 2 files with 0da3bbda656e4f7b,
both at max_hits 16711680.

Usage: python tests/golden/gen_fixtures.py [outdir]   (default: tests/golden/fixtures)
"""
import os
import random
import sys

_COMPLEMENT = {"A": "T", "C": "G", "G": "C", "T": "A", "a": "t", "c": "g", "g": "c",
               "t": "a", "N": "N", "n": "n"}


def revcomp(seq):
    return "".join(_COMPLEMENT[c] for c in reversed(seq))


def random_bases(rng, n):
    return "".join([rng.choice("ACGT") for _ in range(n)])


def build_cases():
    """Returns {name: (ref_records, query_records)} plus the custom 4x4 matrix."""
    rng = random.Random(20260811)
    base = random_bases(rng, 600)
    swap = {"A": "C", "C": "A", "G": "T", "T": "G"}
    cases = {}
    cases["identical"] = ([("ref", base)], [("qry", base)])
    point = base[:300] + swap[base[300]] + base[301:]
    cases["mismatch"] = ([("ref", base)], [("qry", point)])
    cases["indel"] = ([("ref", base)], [("qry", base[:300] + base[305:])])
    cases["revcomp"] = ([("ref", base)], [("qry", revcomp(base))])
    unit = random_bases(rng, 20)
    cases["repeat"] = ([("ref", unit * 30)], [("qry", unit * 30)])
    cases["ambiguous"] = ([("ref", base)], [("qry", base[:250] + "N" * 30 + base[280:])])
    diverged = base[:250] + random_bases(rng, 60) + base[310:]
    cases["xdrop"] = ([("ref", base)], [("qry", diverged)])
    cases["multi"] = (
        [("r1", base[:200]), ("r2", base[200:400]), ("r3", base[400:])],
        [("q1", base[:200]), ("q2", revcomp(base[200:400])), ("q3", base[400:])],
    )
    cases["softmask"] = ([("ref", base)], [("qry", base[:250].lower() + base[250:])])
    cases["scoring"] = ([("ref", base)], [("qry", diverged)])
    return cases


# The `scoring` fixture's ACGT block (+100 match, -50 mismatch).
SCORING_ACGT = [100 if i == j else -50 for i in range(4) for j in range(4)]


def write_fasta(path, records):
    with open(path, "w") as fh:
        for name, seq in records:
            fh.write(">%s\n" % name)
            for i in range(0, len(seq), 60):
                fh.write(seq[i:i + 60] + "\n")


def main(outdir):
    os.makedirs(outdir, exist_ok=True)
    for name, (ref, qry) in sorted(build_cases().items()):
        write_fasta(os.path.join(outdir, name + ".ref.fa"), ref)
        write_fasta(os.path.join(outdir, name + ".qry.fa"), qry)
    print(" ".join(sorted(build_cases())))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else
         os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures"))

/*
 * hspz_synth.h — deterministic synthetic sequences for tests and benchmarks.
 *
 * No counterpart in the reference (which reads FASTA/2bit inputs); these
 * generators fix the synthetic workloads the parity tests and the benchmark
 * run on. Bases are raw ASCII (ACGT, lowercase = soft-masked).
 */
#ifndef HSPZ_SYNTH_H
#define HSPZ_SYNTH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* n iid bases, P(C) = P(G) = gc/2, P(A) = P(T) = (1-gc)/2. */
void hspz_synth_random(uint64_t seed, size_t n, double gc, uint8_t *out);

/* Substitutions (transition:transversion 2:1) at rate sub, indels at rate
 * indel (1:1 insertion:deletion, Geometric(1/3) lengths). Returns the output
 * length (at most cap bytes are written). */
size_t hspz_synth_mutate(const uint8_t *src, size_t n, uint64_t seed, double sub, double indel,
                         double gc, uint8_t *out, size_t cap);

/* In-place reverse complement of a whole sequence / of every odd block. */
void hspz_synth_revcomp(uint8_t *seq, size_t n);
void hspz_synth_revcomp_blocks(uint8_t *seq, size_t n, size_t block);

/* Interspersed repeat families; returns the repeat-derived base count. */
size_t hspz_synth_repeats(uint8_t *seq, size_t n, uint64_t seed, uint32_t families,
                          uint32_t consensus_len, double div_lo, double div_hi, double fraction,
                          int lowercase, double gc);

/* count clipped lognormal lengths; returns their sum. */
size_t hspz_synth_lognormal(uint64_t seed, size_t count, double median, double sigma, double lo,
                            double hi, uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif

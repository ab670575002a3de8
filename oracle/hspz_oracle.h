/*
 * hspz_oracle.h — CPU oracle for the seed-and-filter path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference
 * semantics (hspZ, src/seed.rs, src/hsp.rs, src/scoring.rs, src/sequence.rs,
 * src/gpu/kernels.rs, src/gpu/mod.rs, src/run.rs). Only tests/, the smoke entry
 * and bench.py's cpu_baseline leg may load it; the product library never links
 * or calls it.
 *
 * Pinning (what checks each function beyond its own formula):
 *   orc_build_sub_mat      scoring.rs:295-322 KATs (tests/test_oracle.py)
 *   orc_shape_parse/kmer   seed.rs:503-537 KATs
 *   orc_seed_table_build   seed.rs:615-623 KAT + brute-force bucket check
 *   orc_chunk_seeds        seed.rs:540-545 doc KAT + brute force
 *   orc_intervals/chunks   seed.rs:649-655, sequence.rs:438-443 KATs
 *   orc_extend_*           closed forms (identical seqs, separator drop,
 *                          infinite X = max prefix) + brute-force all-k loop
 *   orc_gate               kernels.rs:2028-2042 rounding KAT
 *   orc_ln                 libm log() to 1e-15 relative
 *   orc_chunk_limits       gpu/mod.rs:2645-2669 KATs
 *   orc_dedup_and_order    hsp.rs:308-368 KATs
 *   end to end             smoke.sh:11-12,78-79 frozen digests (repeat 57 HSPs
 *                          d01edd2118cf5fa5, multi 3 HSPs 0da3bbda656e4f7b)
 *   entropy branch         parity unpinned beyond what the frozen fixtures
 *                          exercise (no in-repo vector isolates it)
 */
#ifndef HSPZ_ORACLE_H
#define HSPZ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_INVALID_KMER (1u << 31)

typedef struct {
    uint32_t ref_start;
    uint32_t query_start;
    uint32_t len;
    int32_t score;
} orc_sp;

typedef struct {
    uint32_t size;       /* span of the pattern */
    uint32_t kmer_size;  /* care positions */
    uint32_t pos[32];
    uint8_t transition[32];
} orc_shape;

/* sequence.rs:66-79 scoring alphabet; sequence.rs:212-230 reverse complement */
void orc_encode(const uint8_t *raw, size_t n, uint8_t *out);
void orc_revcomp(const uint8_t *raw, size_t n, uint8_t *out);

/* scoring.rs:41-120. acgt: 16 entries row-major or NULL for HOXD70.
 * Returns 0 on success, -1 on a malformed ambiguous spec. */
int orc_build_sub_mat(const char *ambiguous, int32_t xdrop, const int32_t *acgt,
                      int32_t out[64]);

/* seed.rs:33-123 */
int orc_shape_parse(const char *spec, orc_shape *s);
uint32_t orc_kmer_at(const orc_shape *s, const uint8_t *raw, size_t pos);

/* seed.rs:135-180. index_table has 4^kmer_size entries. Caller frees both with
 * orc_free. */
int orc_seed_table_build(const uint8_t *raw, size_t n, const orc_shape *s, uint32_t step,
                         uint32_t **index_table, uint32_t **pos_table, size_t *n_pos);
void orc_free(void *p);

/* seed.rs:385-398: seeds for j in [lo, hi). out must hold (hi-lo)*(1+k). */
size_t orc_chunk_seeds(const uint8_t *raw, const orc_shape *s, int transitions, uint32_t lo,
                       uint32_t hi, uint64_t *out);

/* sequence.rs:322-334 / seed.rs:488-496. Return count; out holds pairs. */
size_t orc_intervals(size_t block_len, uint32_t seed_size, uint32_t interval_size,
                     uint32_t *out, size_t cap);
size_t orc_chunks(uint32_t start, uint32_t end, uint32_t wga_chunk, uint32_t *out, size_t cap);

/* Appendix A.4 scalar X-drop (kernels.rs:1279-1543). cells counts evaluated
 * cells including the terminating one. */
void orc_extend_right(const uint8_t *ref, uint32_t ref_len, const uint8_t *qry,
                      uint32_t qry_len, const int32_t *sub, int32_t xdrop, uint32_t r0,
                      uint32_t q0, int32_t *score, int32_t *ext, uint64_t *cells);
void orc_extend_left(const uint8_t *ref, uint32_t ref_len, const uint8_t *qry,
                     uint32_t qry_len, const int32_t *sub, int32_t xdrop, uint32_t r0,
                     uint32_t q0, int32_t *score, int32_t *ext, uint64_t *cells);

/* kernels.rs:1048-1051: keep iff trunc(f64(f32(total))) >= thresh */
int orc_gate_keep(int32_t total, int32_t hspthresh);
/* kernels.rs:1885-1908 */
double orc_ln(double x);
/* gpu/mod.rs:2349-2352 */
#define ORC_LOG4 1.3862943649291992

/* kernels.rs:1063-1702 per anchor; returns 1 and fills *out when kept. */
int orc_materialize(const uint8_t *ref, uint32_t ref_len, const uint8_t *qry, uint32_t qry_len,
                    const int32_t *sub, int32_t xdrop, int32_t hspthresh, int noentropy,
                    uint32_t r0, uint32_t q0, orc_sp *out);

/* gpu/mod.rs:2397-2420. out holds 4-tuples. Returns count or -1 on the
 * reference's underflow case (a seed alone above max_hits at position 0). */
long orc_chunk_limits(const uint32_t *hit_num, uint32_t n, uint32_t max_hits, uint32_t *out,
                      size_t cap);

/* hsp.rs:94-126. In place; returns kept count. */
size_t orc_dedup_and_order(orc_sp *h, size_t n);

typedef struct {
    const uint32_t *index_table;  /* 4^k entries */
    const uint32_t *pos_table;
    const uint8_t *ref_enc;
    uint32_t ref_len;
    const uint8_t *q_enc;
    const uint8_t *q_rc_enc;
    uint32_t query_len;
    const int32_t *sub;
    uint32_t seed_size;
    int32_t xdrop;
    int32_t hspthresh;
    int noentropy;
    uint32_t max_hits;
} orc_engine;

typedef struct {
    uint64_t seeds;
    uint64_t hits;
    uint64_t gate_pass;
    uint64_t raw_hsps;
    uint64_t hsps;
    uint64_t cells;  /* scalar cells evaluated by the gate loops over all hits */
} orc_stats;

/* gpu/mod.rs:1499-2178 for one batch of seeds. *out is malloc'd (orc_free). */
int orc_seed_and_filter(const orc_engine *e, const uint64_t *seeds, uint32_t nseeds, int rev,
                        orc_sp **out, size_t *nout, orc_stats *st);

/* run.rs:266-495 over one query pass. raw_fwd/raw_rc are raw bytes (seeding
 * alphabet). Output: per interval, fw list then rc list, flattened into *out
 * with counts[2*i] (fw) and counts[2*i+1] (rc). Parallel over batches with
 * `threads` workers; output is independent of the thread count. */
int orc_run_pass(const orc_engine *e, const uint8_t *raw_fwd, const uint8_t *raw_rc,
                 uint32_t block_len, const orc_shape *shape, int transitions, int plus,
                 int minus, uint32_t wga_chunk, uint32_t interval_size, int threads,
                 orc_sp **out, size_t **counts, size_t *n_intervals, orc_stats *st);

/* benchmark.rs:401-421 */
uint64_t orc_hash_pass(const orc_sp *all, const size_t *counts, size_t n_intervals);

#ifdef __cplusplus
}
#endif
#endif

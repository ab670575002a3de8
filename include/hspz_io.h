/*
 * hspz_io.h — C ABI of the host input and output paths around the engine:
 * reading FASTA / gzipped FASTA / 2bit into the packed block the engine takes
 * (the reference's src/sequence.rs), and writing the ordered HSP lists as
 * `.segments` files, optionally partitioned along the diagonal and optionally
 * archived as one .tar.gz (src/hsp.rs records/render_records, src/partition.rs,
 * src/sink.rs, and the Emitter of src/run.rs).
 *
 * Same conventions as hspz.h: status 0 = ok, > 0 = misuse (HSPZ_E_*), < 0 =
 * I/O or runtime failure; hspz_last_error() holds the message.
 */
#ifndef HSPZ_IO_H
#define HSPZ_IO_H

#include <stddef.h>
#include <stdint.h>

#include "hspz.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HSPZ_E_IO 8 /* file missing, unreadable or malformed */

/* sequence.rs:89-94 Chr: a record of a packed block. start is the offset of
 * its first base in the block, len its base count. */
typedef struct hspz_chr {
    const char *name;
    uint64_t start;
    uint32_t len;
} hspz_chr;

/* ---- Input (sequence.rs) ------------------------------------------------- */

/* sequence.rs:116-120 Format, chosen from magic bytes (sequence.rs:125-145). */
#define HSPZ_FORMAT_FASTA 0
#define HSPZ_FORMAT_FASTA_GZ 1
#define HSPZ_FORMAT_TWOBIT 2

/* A packed sequence set (sequence.rs:97-110 Genome): records joined by '&'
 * with the trailing separator in buf but not in block_len, plus the record
 * table. A reverse-complement set holds block_len bytes and no trailing
 * separator (sequence.rs:212-230). */
typedef struct hspz_seqset hspz_seqset;
typedef struct hspz_seqset_info {
    const uint8_t *buf; /* raw bytes, owned by the set */
    size_t buf_len;
    size_t block_len;
    const hspz_chr *chrs; /* owned by the set */
    size_t n_chrs;
    int32_t format;       /* HSPZ_FORMAT_* of the file it was read from */
    uint64_t bytes_read;  /* on-disk size of that file, 0 for packed records */
} hspz_seqset_info;

/* read_records + pack (sequence.rs:238-248, :185-205): FASTA (the reference's
 * kseq-style reader, sequence.rs:346-414), gzipped FASTA (concatenated members
 * allowed) or UCSC 2bit with soft masks kept (sequence.rs:161-174), by magic.
 * Record names get `prefix` (may be NULL). */
int hspz_seqset_read(const char *path, const char *prefix, hspz_seqset **out);
/* pack (sequence.rs:185-205) of n in-memory records. */
int hspz_seqset_pack(size_t n, const char *const *names, const uint8_t *const *seqs, const size_t *lens,
                     const char *prefix, hspz_seqset **out);
/* reverse_complement (sequence.rs:212-230): the IUPAC-aware complement of the
 * block reversed, and the record table remapped to start' = n - start - len in
 * reverse order. */
int hspz_seqset_reverse_complement(const hspz_seqset *s, hspz_seqset **out);
int hspz_seqset_get(const hspz_seqset *s, hspz_seqset_info *info);
/* Genome::digest (sequence.rs:289-306): FNV-1a over the record table (name,
 * start as u64, len as u32, little-endian), the buffer and block_len. Equal
 * sets from different formats have equal digests. */
uint64_t hspz_seqset_digest(const hspz_seqset *s);
void hspz_seqset_free(hspz_seqset *s);

/* encode (sequence.rs:66-79): raw bytes to the scoring alphabet (A0 C1 G2 T3,
 * acgt L4, nN N5, '&' E7, other X6) on `threads` host threads (0 = all). */
int hspz_encode(const uint8_t *raw, size_t n, int threads, uint8_t *codes);

/* intervals (sequence.rs:322-334) of a block: up to cap (start, end) pairs,
 * end inclusive for seeding; *n_out counts them all. */
int hspz_intervals(size_t block_len, uint32_t seed_size, uint32_t interval_size, uint32_t *pairs, size_t cap,
                   size_t *n_out);

/* ---- Output (hsp.rs, partition.rs, sink.rs, run.rs) ---------------------- */

/* hsp.rs:179-187 Record: a printed `.segments` row, still numeric: 1-based,
 * record-relative, both ends as printed. */
typedef struct hspz_record {
    uint32_t r_chr;
    uint32_t q_chr;
    uint64_t r_start;
    uint64_t r_end;
    uint64_t q_start;
    uint64_t q_end;
    int32_t score;
} hspz_record;

/* records (hsp.rs:213-234): the rows render_segments prints, in printed order
 * (the minus list reversed). out holds n records. */
int hspz_records(const hspz_segment_pair *hsps, size_t n, const hspz_chr *r_chrs, size_t n_r,
                 const hspz_chr *q_chrs, size_t n_q, char strand, hspz_record *out);
/* render_records (hsp.rs:237-254) / render_segments (hsp.rs:263-292): the
 * `.segments` text; *text is library-allocated (hspz_free), not NUL-counted
 * in *len. */
int hspz_render_records(const hspz_record *recs, size_t n, const hspz_chr *r_chrs, size_t n_r,
                        const hspz_chr *q_chrs, size_t n_q, char strand, char **text, size_t *len);
int hspz_render_segments(const hspz_segment_pair *hsps, size_t n, const hspz_chr *r_chrs, size_t n_r,
                         const hspz_chr *q_chrs, size_t n_q, char strand, char **text, size_t *len);

/* Partitioner::chunk_size_for (partition.rs:84-102) given the sizes observed
 * so far (history, in observation order): *chunk = 0 means "emit whole". */
int hspz_partition_chunk_size(const uint64_t *history, size_t n_history, uint64_t count, uint64_t *chunk);
/* split (partition.rs:122-192): groups by (r_chr, q_chr), splits busy pairs
 * along the strand's diagonal key, bin-packs the rest. out holds n records
 * (the partitions concatenated), sizes receives the partition sizes (capacity
 * n; *n_parts counts them). */
int hspz_partition_split(const hspz_record *recs, size_t n, char strand, uint64_t chunk, hspz_record *out,
                         uint64_t *sizes, size_t *n_parts);

/* benchmark.rs:411-436 hash_pass / hash_segments: FNV-1a over a pass's lists
 * (intervals in order, fw then rc; counts holds 2 per interval), over the
 * fields as little-endian u32, or over the rendered `.segments` text. */
uint64_t hspz_hash_pass(const hspz_segment_pair *hsps, const size_t *counts, size_t n_intervals);
int hspz_hash_segments(const hspz_segment_pair *hsps, const size_t *counts, size_t n_intervals,
                       const hspz_chr *r_chrs, size_t n_r, const hspz_chr *q_chrs, size_t n_q,
                       const hspz_chr *rc_chrs, size_t n_rc, uint64_t *hash);

/* run.rs:535-666 Emitter: one output pass into a directory or a .tar.gz,
 * sharing one -D history across work units. */
typedef struct hspz_emitter hspz_emitter;
typedef struct hspz_output_report {
    double partition_ms, format_ms, archive_ms;
    uint64_t files, bytes_in, bytes_out;
} hspz_output_report;
/* Emitter::new (run.rs:547-562): tarball NULL writes files into output_dir
 * (created); tarball "" or "-" writes "<output_dir>.tar.gz" (run.rs:506-517);
 * otherwise that path. diagonal = -D. Entries are bare file names; the
 * archive's members carry mode 0644 and mtime 0, so it is reproducible. */
int hspz_emitter_new(const char *output_dir, const char *tarball, int diagonal, hspz_emitter **out);
/* Emitter::emit_unit (run.rs:568-626): every logical file of one reference bin
 * x query bin work unit: for interval n (1-based) the fw list as
 * tmp<n>.block<query_bin>.r<ref_bin>.plus[.split<k>].segments against
 * query_chrs, then the rc list as .minus against rc_chrs; empty lists write
 * nothing. counts holds 2 per interval (hspz_run_pass's layout). */
int hspz_emit_unit(hspz_emitter *em, uint32_t ref_bin, uint32_t query_bin, const hspz_chr *ref_chrs, size_t n_ref,
                   const hspz_chr *query_chrs, size_t n_query, const hspz_chr *rc_chrs, size_t n_rc,
                   const hspz_segment_pair *hsps, const size_t *counts, size_t n_intervals);
/* Emitter::finish (run.rs:645-666): writes the archive trailers, releases
 * the emitter (also on failure) and fills *report when not NULL. */
int hspz_emitter_finish(hspz_emitter *em, hspz_output_report *report);

#ifdef __cplusplus
}
#endif
#endif

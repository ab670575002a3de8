/*
 * hspz.h — C ABI of the MI355X seed-and-filter engine.
 *
 * Each entry point replaces one method of the reference's `Engine`
 * (hspZ src/gpu/mod.rs) or one host function the reference's callers use
 * (src/seed.rs, src/hsp.rs, src/run.rs). Plain pointers and sizes only; no
 * exceptions cross this boundary. Status codes: 0 = ok, < 0 = HIP/runtime
 * error, > 0 = API misuse (see HSPZ_E_*). hspz_last_error() returns a
 * thread-local message for the last failing call on the calling thread.
 *
 * Threading: a handle is not thread-safe. One handle per thread; several
 * handles per process (and per device) are allowed (gpu/mod.rs:817-950,
 * run.rs:782).
 */
#ifndef HSPZ_H
#define HSPZ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HSPZ_OK 0
#define HSPZ_E_NO_QUERY 1      /* seed_and_filter on an empty or absent query, run_pass before any
                                  swap_query (gpu/mod.rs:1515-1519) */
#define HSPZ_E_ARG 2           /* bad argument or size */
#define HSPZ_E_SHAPE 3         /* seed shape mismatch (gpu/mod.rs:1320-1341) */
#define HSPZ_E_RANGE 4         /* seed range beyond the query (gpu/mod.rs:1326-1331) */
#define HSPZ_E_CHUNK 5         /* chunk_limits underflow (gpu/mod.rs:2405) */
#define HSPZ_E_NO_DEVICE 6     /* no HIP device */
#define HSPZ_E_OOM 7           /* device or host allocation failed / admission refused */

/* hsp.rs:31-41: 16 bytes, align 16, same field order. */
typedef struct hspz_segment_pair {
    uint32_t ref_start;
    uint32_t query_start;
    uint32_t len;
    int32_t score;
} __attribute__((aligned(16))) hspz_segment_pair;

/* gpu/mod.rs:1005-1021 EngineConfig. Host arrays are borrowed for the call
 * and copied to the device (the reference uploads them in Engine::new,
 * gpu/mod.rs:1039-1047). ref_seq is the encoded reference block
 * (sequence.rs:66-79 alphabet, codes 0..7). */
typedef struct hspz_engine_config {
    const uint32_t *index_table; /* 4^kmer_size entries, inclusive bucket ends */
    size_t index_len;
    const uint32_t *pos_table;
    size_t pos_len;
    const uint8_t *ref_seq;
    size_t ref_len;
    const int32_t *sub_mat; /* 64 entries, 8x8 row-major (scoring.rs:41-120) */
    uint32_t seed_size;     /* pattern span, 19 for 12of19 */
    int32_t xdrop;
    int32_t hspthresh;
    int32_t noentropy;
    uint32_t max_hits; /* 0 = derive from device memory (gpu/mod.rs:2423-2432) */
    uint32_t hsp_blocks; /* accepted for API parity; 0 = default */
} hspz_engine_config;

/* gpu/mod.rs:57-67 FilterOutput. hsps is library-allocated; release with
 * hspz_filter_output_free. */
typedef struct hspz_filter_output {
    hspz_segment_pair *hsps;
    size_t n_hsps;
    uint32_t num_hits;
    uint32_t raw_hsps;
    /* FilterOutput::raw (gpu/mod.rs:55-67): the records surviving the score
     * threshold before dedup, in the order each max_hits chunk produced them,
     * kept only when HSPZ_FLAG_DUMP_RAW is set (else NULL / 0). */
    hspz_segment_pair *raw;
    size_t n_raw;
} hspz_filter_output;

typedef struct hspz_engine hspz_engine;

/* ---- Engine (gpu/mod.rs) ---------------------------------------------- */

/* Engine::new (gpu/mod.rs:1024-1128). */
int hspz_engine_new(int device, const hspz_engine_config *cfg, hspz_engine **out);
/* Engine::new with the reference index built on the device and kept there
 * (SeedTable::build, seed.rs:126-373: the same positions, buckets and order as
 * hspz_seed_table_build with the same pattern and step). cfg->index_table,
 * index_len, pos_table and pos_len are ignored; cfg->seed_size must equal the
 * pattern's span. */
int hspz_engine_new_device_index(int device, const hspz_engine_config *cfg, const char *pattern, uint32_t step,
                                 hspz_engine **out);
/* A second engine on base's device over the same resident target (the
 * reference bin the reference keeps resident across work units, gpu/mod.rs:
 * 1130-1148): it has its own stream, query side, seed slots and work buffers
 * and base's scoring and gate configuration, and reads base's device
 * reference, index, pos table, packed planes, clean bits and reference
 * records (built now if the engine would use them) instead of holding a
 * copy. The target's device memory is held once, by whichever of the
 * engines sharing it is freed last; HSPZ_STAT_DEVICE_BYTES counts it in base
 * only. Engines sharing a target may run on different host threads. */
int hspz_engine_share(hspz_engine *base, hspz_engine **out);
/* Copies the engine's resident index to the host: index_len must be 4^k;
 * *pos_table is library-allocated (hspz_free). No counterpart in the
 * reference; it lets a caller reuse or check a device-built index. */
int hspz_engine_index(const hspz_engine *h, uint32_t *index_table, size_t index_len, uint32_t **pos_table,
                      size_t *pos_len);
void hspz_engine_free(hspz_engine *h);

/* Engine::swap_query (gpu/mod.rs:1142-1148): encoded forward and reverse-
 * complement query blocks of equal length; copied, no ownership transfer. */
int hspz_swap_query(hspz_engine *h, const uint8_t *fwd, const uint8_t *rc, size_t len);
/* Engine::swap_query from the raw query block (records joined by '&', as
 * pack makes it): the reverse complement and the encoding of both strands
 * (the caller's per-bin host step, run.rs:92-97, sequence.rs:212-230 and
 * :66-79) run on the device from one upload of the raw bytes. Same result as
 * hspz_swap_query(encode(raw), encode(reverse_complement(raw))). */
int hspz_swap_query_raw(hspz_engine *h, const uint8_t *raw, size_t len);
/* Copies the engine's encoded query strands back to the host (len must be the
 * swapped length). No counterpart in the reference; it lets a caller check
 * the device encoding. */
int hspz_engine_query(const hspz_engine *h, uint8_t *fwd, uint8_t *rc, size_t len);

/* Engine::upload_seeds (gpu/mod.rs:1246-1309): slot 0 or 1; seeds are
 * (kmer << 32) | query_position in seed-stream order (seed.rs:385-398).
 * Copied synchronously, so the host array may be reused on return. */
int hspz_upload_seeds(hspz_engine *h, int slot, const uint64_t *seeds, size_t n);
/* Engine::upload_seeds with async_seed_copy (gpu/mod.rs:1246-1309): the copy
 * runs on the engine's copy stream, ordered after the compute stream's last
 * work on the slot, and the next hspz_seed_and_filter (or its no-alloc
 * variant) on the slot waits for it on the device; the call returns at once.
 * Slot lifetime: the host array must stay allocated and unchanged until that
 * slot's copy has completed, which it has once hspz_seed_and_filter /
 * hspz_seed_and_filter_noalloc / hspz_generate_seeds / hspz_upload_seeds* on
 * the same slot or hspz_sync_seeds(h, slot) has returned. Issue it one batch
 * ahead (the other slot) for the copy to hide behind compute. A pinned host
 * array (hspz_pinned_take) is what makes the copy asynchronous. */
int hspz_upload_seeds_async(hspz_engine *h, int slot, const uint64_t *seeds, size_t n);
/* Waits on the host for the slot's pending asynchronous copy. No counterpart
 * in the reference (its caller synchronises the copy stream itself). */
int hspz_sync_seeds(hspz_engine *h, int slot);
/* Engine::take_pinned / give_pinned (gpu/mod.rs:1165-1187): a pinned host
 * staging buffer of at least n seeds from the engine's pool (page-locked
 * once, reused across passes), and its return to the pool; buffers are freed
 * with the engine. Give a buffer back only after its copy has completed
 * (the slot lifetime above). HSPZ_STAT_PINNED_ACTIVE reports whether one was
 * handed out (Engine::pinned_seeds_active). */
int hspz_pinned_take(hspz_engine *h, size_t n, uint64_t **buf);
int hspz_pinned_give(hspz_engine *h, uint64_t *buf);
/* Engine::check_seed_bytes (gpu/mod.rs:1464-1495): compares the slot's seeds
 * (as hspz_generate_seeds or an upload left them) with `expected` byte for
 * byte; HSPZ_E_ARG with the count or first differing index otherwise. */
int hspz_check_seed_bytes(hspz_engine *h, int slot, const uint64_t *expected, size_t n);
/* Engine::launch_overhead_ms (gpu/mod.rs:1210-1236): mean wall-clock cost of
 * one launch of an empty kernel over `iters` launches after 16 warm-up ones. */
int hspz_launch_overhead_ms(hspz_engine *h, uint32_t iters, double *ms);

/* Engine::generate_seeds (gpu/mod.rs:1312-1463): device seeding of query
 * positions [start, end) on the forward (rev=0) or reverse-complement block.
 * shape_pos lists the care positions; every care position admits a
 * transition (seed.rs:53-56). Writes the seed count to *n_out. */
int hspz_generate_seeds(hspz_engine *h, int slot, int rev, uint32_t start, uint32_t end,
                        const uint32_t *shape_pos, uint32_t kmer_size, int transitions,
                        uint32_t *n_out);

/* Engine::seed_and_filter (gpu/mod.rs:1499-2178): count, scan, chunk by
 * max_hits, expand hits, score gate, compact survivors, materialize, compact,
 * dedup_and_order per chunk. Returns with the device drained. */
int hspz_seed_and_filter(hspz_engine *h, int slot, int rev, hspz_filter_output *out);
void hspz_filter_output_free(hspz_filter_output *out);

/* hspz_seed_and_filter without any allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same lists as hspz_seed_and_filter on the same slot and strand, written to
 * the caller's out[0..cap); *n_out receives the kept count, *num_hits and
 * *raw_hsps (either may be NULL) what FilterOutput carries. All device work
 * runs in a file-scope device array of 256 MiB per device (held from module
 * load to process exit), never in memory the call allocates: it calls no
 * hipMalloc, hipFree or allocating library routine. Limits, refused with
 * HSPZ_E_OOM and a message: the seeds of the slot must leave at least 64 MiB of
 * the array (up to about 47 M seeds), a max_hits chunk may have 1.3 M gate
 * survivors (fewer, down to 0.3 M, as the slot holds more seeds), a call up to
 * 256 max_hits chunks, and out must hold each
 * chunk's pre-dedup list after the lists kept before it. Any number of hits:
 * they are expanded and gated in tiles. Every call on a device shares that
 * device's array, so calls that overlap on one device race: serialise them
 * (one thread, or one stream's worth of calls at a time). */
int hspz_seed_and_filter_noalloc(hspz_engine *h, int slot, int rev, hspz_segment_pair *out, size_t cap,
                                 size_t *n_out, uint32_t *num_hits, uint32_t *raw_hsps);

/* Accessors (gpu/mod.rs:1201-1213, 2292-2306). */
#define HSPZ_STAT_MAX_HITS 0
#define HSPZ_STAT_LAUNCHES 1
#define HSPZ_STAT_PIPELINE_SYNCS 2
#define HSPZ_STAT_REFERENCE_UPLOADS 3
#define HSPZ_STAT_QUERY_SWAPS 4
#define HSPZ_STAT_PEAK_DEVICE_BYTES 5
#define HSPZ_STAT_GATE_NS 6        /* cumulative score-gate kernel time (HIP events) */
#define HSPZ_STAT_GATE_HITS 7      /* hits the score gate has scored */
#define HSPZ_STAT_GATE_LAUNCHES 8
#define HSPZ_STAT_GATE_DEFERRED 9  /* hits the packed gate handed to the exact wave gate */
#define HSPZ_STAT_GRAPH_CAPTURES 10 /* batched device passes captured into a hipGraph */
#define HSPZ_STAT_GRAPH_REPLAYS 11  /* batched device passes replayed from the cached graph */
/* Cumulative wall time of hspz_filter_batch, in ns, by part: */
#define HSPZ_STAT_BATCH_PREP_NS 12   /* host layout of the arena and batches, uploads */
#define HSPZ_STAT_BATCH_DEVICE_NS 13 /* device pass, launch to drained stream */
#define HSPZ_STAT_BATCH_READ_NS 14   /* read-back of counters and kept HSPs */
#define HSPZ_STAT_BATCH_POST_NS 15   /* host dedup_and_order per scope and assembly */
#define HSPZ_STAT_INDEX_BUILD_NS 16  /* device index build of hspz_engine_new_device_index */
#define HSPZ_STAT_POS_LEN 17         /* indexed positions held by the engine */
#define HSPZ_STAT_INDEX_LEN 18       /* index_table entries (4^k) held by the engine */
#define HSPZ_STAT_BATCH_RERUNS 19    /* batched passes run again because a survivor or deferred-hit
                                        list overflowed (each time both lists are doubled) */
#define HSPZ_STAT_BATCH_PASSES 20     /* device passes the last hspz_filter_batch call ran (see
                                        HSPZ_FLAG_BATCH_MAX_POSITIONS) */
#define HSPZ_STAT_DEVICE_BYTES 21     /* device bytes the engine holds now (resident reference side and
                                        grow-only work buffers; PEAK_DEVICE_BYTES is the most it has held) */
#define HSPZ_STAT_BATCH_CANDIDATES 22 /* hits of the last batched pass the coarse bound passed on */
#define HSPZ_STAT_SEED_UPLOADS 23     /* seed uploads (hspz_upload_seeds and _async) */
#define HSPZ_STAT_PINNED_ACTIVE 24    /* 1 once hspz_pinned_take handed out a pinned buffer */
int hspz_get_stat(const hspz_engine *h, int which, uint64_t *value);

/* Engine switches (the reference's pub flags, gpu/mod.rs:836-949, have no
 * meaning here; these are this engine's own). */
#define HSPZ_FLAG_BATCH_GRAPH 0        /* 1 (default): capture/replay the batched pass as a hipGraph */
#define HSPZ_FLAG_BATCH_GATE_BLOCKS 1  /* persistent grid of the batched gate in blocks; 0 (default) = as many
                                          as are resident at once */
#define HSPZ_FLAG_BATCH_STEP_CELLS 2   /* cells per gate step: 16, 32, 64, 96 (default) or 128. 128: 64
                                          right + 64 left cells of a hit in one step; 96: 32 right + 64
                                          left, the right side continuing with its next 32 cells.
                                          Seed-major order runs both as 64 */
#define HSPZ_FLAG_BATCH_BLOCK_THREADS 3 /* threads per gate block: 512 or 1024 (default) */
#define HSPZ_FLAG_BATCH_CLAIM_SEEDS 4  /* seeds a wave claims from the work counter at a time (1..4096,
                                          default 256) */
#define HSPZ_FLAG_BATCH_ORDER 5        /* hit order of the batched gate: 0 seed-major, 1 k-mer-major (a
                                          reference window is reused by every query seed of its k-mer),
                                          2 (default) k-mer-major when the batch holds at least one seed
                                          per two k-mers */
#define HSPZ_FLAG_BATCH_ORDER_USED 6   /* read-only: the order of the last batched pass */
#define HSPZ_FLAG_BATCH_DEVICE_DEDUP 7 /* 1 (default): dedup_and_order of the batched path on the device
                                          (stable radix sorts of all scopes at once); 0: on host threads */
#define HSPZ_FLAG_REF_RECORDS 8        /* reference records: 48 B per indexed position in pos_table order
                                          (the gate's reference words, aligned), built on first use when
                                          device memory allows. The merged-step gates (96- and 128-cell
                                          steps) build and read them whenever they run; 1: the narrower
                                          steps want them too; 0 (default): only the merged steps do;
                                          -1: none does (the gates read windows from the plane). Setting 0
                                          or -1 releases built ones. hspz_get_flag reports 2 once they are
                                          built. Any k-mer-major gate reads them when they exist */
#define HSPZ_FLAG_BATCH_LOOKAHEAD 9    /* 1 (default): with 96- or 128-cell steps and 512-thread blocks, the
                                          k-mer-major gate issues each hit's and each item's loads one
                                          iteration ahead of their use; it reads the reference records
                                          (built on first use when device memory allows, as
                                          HSPZ_FLAG_REF_RECORDS builds them). 0: the plain merged gate */
#define HSPZ_FLAG_BATCH_LIST_CAP 10    /* starting size, in records, of the batched pass's survivor and
                                          deferred-hit lists. 0 (default): sized from the batch, and kept at
                                          what earlier passes grew them to. 64..2^28: every pass starts at
                                          this size. A list that overflows is doubled and the pass run again
                                          (HSPZ_STAT_BATCH_RERUNS), so the lists only change how often that
                                          happens, never the results */
#define HSPZ_FLAG_NOALLOC_TILE_HITS 11 /* hits hspz_seed_and_filter_noalloc expands and gates at a time: 0
                                          (default) = as many as its scratch array holds; 64 or more caps it
                                          (smaller tiles only add launches; results do not change) */
#define HSPZ_FLAG_BATCH_MAX_POSITIONS 12 /* query positions (both strands) one device pass of
                                          hspz_filter_batch holds: a call runs its batches, in order, in
                                          as many passes as this needs (a batch is never split; each is its
                                          own dedup scope, so the lists do not change). 0 (default): half of
                                          the free device memory's worth, between 2^20 and 2^26 */
#define HSPZ_FLAG_BATCH_COARSE 13      /* 1 (default): in k-mer-major order with reference records, every hit
                                          first meets a cheaper upper bound of its X-drop total (each cell
                                          scored by the best score of its class: match, transition,
                                          transversion; counted per 16-cell word) and only hits the bound
                                          does not reject run the two-cell maps; two hits per lane in packed
                                          16-bit sums when every class score is within +-127 (two reference
                                          positions sharing each query record), else one hit per lane in
                                          32-bit sums. 3: the packed form with two query records sharing one
                                          reference position. 2: always the one-hit form. 0: every hit runs
                                          the two-cell maps. The lists do not change */
#define HSPZ_FLAG_BATCH_COARSE_USED 14 /* read-only: the coarse bound form the last batched pass ran:
                                          0 none, 1 one hit per lane, 2 two query records per lane, 3 two
                                          reference positions per lane */
#define HSPZ_FLAG_DUMP_RAW 15          /* Engine::dump_raw (gpu/mod.rs:55-67): 1 = hspz_seed_and_filter keeps
                                          the pre-dedup records in hspz_filter_output.raw; 0 (default) */
#define HSPZ_FLAG_BATCH_CAND_PLANES 16 /* where the two-cell maps behind the coarse bound read a candidate's
                                          windows: 1 (default) the 2-bit planes at R and Q, 0 its reference
                                          and query records. The lists do not change */
#define HSPZ_FLAG_BATCH_LANE_MATERIALIZE 17 /* 1 (default): the batched materializer extends one survivor per
                                          lane (sides longer than 2048 cells go to one wave per survivor);
                                          0: one wave per survivor throughout. The lists do not change */
#define HSPZ_FLAG_BATCH_CAND_BINS 18   /* when the two-cell maps read the planes, bin the coarse bound's
                                          candidates by reference position first (1024 bins over the
                                          target; each XCD's waves work through one eighth of the bins
                                          together, so that the plane lines they read stay in its L2).
                                          1 (default): in passes of at most 12 Mi query bases (a query
                                          plane that stays in L2 beside them); 2: every pass; 0: never.
                                          Needs a second candidate list in device memory and is skipped
                                          when that does not fit. The lists do not change */
#define HSPZ_FLAG_BATCH_CAND_BINNED 19 /* read-only: 1 if the last batched pass binned its candidates */
#define HSPZ_FLAG_BATCH_COARSE_LONG 20 /* 1 (default): in the paired coarse forms, a candidate with a side the
                                          bound did not see drop and a side whose class-score sum reached
                                          ten best matches goes straight to the deferred-hit gate instead
                                          of the two-cell maps (a homologous extension that would outlive
                                          their windows); 0: every candidate runs the two-cell maps. The
                                          lists do not change */
int hspz_set_flag(hspz_engine *h, int which, int value);
/* Current value of a flag (the resolved grid for HSPZ_FLAG_BATCH_GATE_BLOCKS
 * once a batched pass has run). */
int hspz_get_flag(const hspz_engine *h, int which, int *value);

/* resolve_max_hits / device_count / device_memory (gpu/mod.rs:2436-2463). */
uint32_t hspz_default_max_hits(uint64_t total_device_bytes);
int hspz_device_count(void);
int hspz_device_memory(int device, uint64_t *free_bytes, uint64_t *total_bytes);

const char *hspz_last_error(void);

/* ---- Host-side path functions (seed.rs, hsp.rs, run.rs) ---------------- */

/* SeedTable::build_parallel (seed.rs:207-361), byte-identical to the serial
 * build (seed.rs:135-180). raw is the raw reference block (seeding alphabet:
 * uppercase ACGT only). pattern is "12of19", "14of22" or a 1/0/T string.
 * index_table must hold 4^kmer_size entries; *pos_table is library-allocated
 * (free with hspz_free). */
int hspz_seed_table_build(const uint8_t *raw, size_t n, const char *pattern, uint32_t step,
                          int threads, uint32_t *index_table, size_t index_len,
                          uint32_t **pos_table, size_t *pos_len);

/* Shape::parse (seed.rs:47-84): care positions into pos (capacity 32). */
int hspz_shape_parse(const char *pattern, uint32_t *size, uint32_t *kmer_size, uint32_t *pos);

/* chunk_seeds_parts + concat_parts (seed.rs:428-477): seeds for query
 * positions [lo, hi) of the raw block; out holds (hi-lo)*(1+kmer_size). */
int hspz_chunk_seeds(const uint8_t *raw, size_t n, const char *pattern, int transitions,
                     uint32_t lo, uint32_t hi, int threads, uint64_t *out, size_t *n_out);

/* build_sub_mat (scoring.rs:41-120): the 8x8 matrix over the scoring alphabet.
 * acgt is the 4x4 ACGT block row-major, or NULL for HOXD70 (scoring.rs:26-31);
 * ambiguous is "" / "n" / "iupac" / "<n|iupac>,<reward>,<penalty>". */
int hspz_build_sub_mat(const char *ambiguous, int32_t xdrop, const int32_t *acgt, int32_t *out);

/* dedup_and_order (hsp.rs:94-126), in place; writes the kept count. */
int hspz_dedup_and_order(hspz_segment_pair *hsps, size_t n, size_t *n_kept);

/* seed_and_filter_all (run.rs:266-495) over one query pass on the engine's
 * current query: batches per interval (plus chunks then minus chunks), device
 * seeding, results concatenated per interval in batch order. counts receives
 * 2 entries per interval (fw, rc); both outputs are library-allocated
 * (hspz_free). */
typedef struct hspz_pass_params {
    const char *pattern;
    int32_t transitions;
    int32_t plus;
    int32_t minus;
    uint32_t wga_chunk;     /* 250000 (cli.rs:81-82) */
    uint32_t interval_size; /* 10000000 (cli.rs:83-84) */
} hspz_pass_params;
typedef struct hspz_pass_stats {
    uint64_t seeds;
    uint64_t hits;
    uint64_t raw_hsps;
    uint64_t hsps;
    uint64_t batches;
} hspz_pass_stats;
int hspz_run_pass(hspz_engine *h, const hspz_pass_params *p, hspz_segment_pair **hsps,
                  size_t **counts, size_t *n_intervals, hspz_pass_stats *stats);

/* ---- Batched path ------------------------------------------------------- */

/* One query block: encoded forward and reverse-complement bytes (records
 * joined by '&', as Engine::swap_query takes them), or with rc = NULL the raw
 * forward block, whose two strands are then encoded on the device as
 * hspz_swap_query_raw does. */
typedef struct hspz_query_chunk {
    const uint8_t *fwd;
    const uint8_t *rc;
    size_t len;
} hspz_query_chunk;

/* Results per chunk, in chunk order: for chunk c and interval i, the fw list
 * then the rc list; counts[2*(b_c + i)] and counts[2*(b_c + i) + 1] are their
 * lengths, b_c = sum of chunk_intervals[0..c). Library-allocated; release with
 * hspz_batch_output_free. */
typedef struct hspz_batch_output {
    hspz_segment_pair *hsps;
    size_t n_hsps;
    size_t *counts;
    size_t *chunk_intervals;
    size_t n_chunks;
} hspz_batch_output;

/* Batched seed-and-filter: the output for each chunk equals hspz_run_pass on
 * that chunk alone (same batches, same max_hits dedup scopes, same order), but
 * all batches of all chunks run in one device pass: device seeding, hit
 * counting, per-batch max_hits chunking, a fused hit-expansion + score-gate
 * kernel, exact materialization, then per-scope dedup_and_order on the host.
 * No counterpart in the reference, which runs batches one by one
 * (run.rs:266-495); its callers are the batched and serving loops. */
int hspz_filter_batch(hspz_engine *h, const hspz_pass_params *p, const hspz_query_chunk *chunks,
                      size_t n_chunks, hspz_batch_output *out, hspz_pass_stats *stats);
void hspz_batch_output_free(hspz_batch_output *out);

/* One measured configuration of the batched gate. */
typedef struct hspz_tune_cell {
    uint32_t step_cells;    /* HSPZ_FLAG_BATCH_STEP_CELLS */
    uint32_t block_threads; /* HSPZ_FLAG_BATCH_BLOCK_THREADS */
    uint32_t grid_blocks;   /* HSPZ_FLAG_BATCH_GATE_BLOCKS as launched */
    uint32_t claim_seeds;   /* HSPZ_FLAG_BATCH_CLAIM_SEEDS */
    uint32_t order;         /* HSPZ_FLAG_BATCH_ORDER (0 or 1) */
    double gate_ms;         /* best of the repetitions: fused gate + exact gate for deferred hits */
    double after_ms;        /* best of the repetitions: materializer and output compaction */
    uint64_t hits, survivors, deferred;
    int status;             /* hspz_filter_batch's status for this cell */
} hspz_tune_cell;

/* Autotuner of the batched gate (no counterpart in the reference). Runs
 * hspz_filter_batch over `chunks` for both hit orders at the engine's knobs,
 * then, in the faster order, for every cell of a grid over step width
 * {16, 32, 64, 96, 128} cells x block size {512, 1024} x wave claim {16, 64, 256} seeds at
 * the resident grid (in k-mer-major order behind the coarse bound, whose
 * candidates' gate has its own step: block size {512, 1024} x wave claim
 * {64, 256, 1024} at the engine's step), then over grid sizes {1/2, 3/4, 1, 3/2, 2} x resident
 * for the best of those, `reps` times each, and records every cell it measured
 * in cells[0..*n_cells) (up to cap; *n_cells counts all of them). The engine
 * keeps the cell with the least gate_ms + after_ms, which is also returned in
 * *best. Outputs do not depend on the configuration. */
int hspz_autotune_batch(hspz_engine *h, const hspz_pass_params *p, const hspz_query_chunk *chunks,
                        size_t n_chunks, uint32_t reps, hspz_tune_cell *cells, size_t cap, size_t *n_cells,
                        hspz_tune_cell *best);

/* ---- Serving loop --------------------------------------------------------- */

/* A continuous-batching request loop over resident engines (no counterpart in
 * the reference; its callers are serving front ends). Each worker thread owns
 * one engine (same reference, same device, its own stream) and repeatedly
 * admits the longest FIFO prefix of waiting requests that fits the remaining
 * HBM budget, runs one hspz_filter_batch over them and publishes each
 * request's lists, which equal hspz_run_pass on that request alone. */
typedef struct hspz_server hspz_server;
typedef struct hspz_server_config {
    hspz_pass_params params;
    uint64_t hbm_budget;         /* device bytes admitted at once over all workers; 0 = 1/4 of free */
    uint32_t max_batch_requests; /* 0 = 256 */
    uint64_t max_batch_bases;    /* query bases per pass; 0 = 64 Mi */
} hspz_server_config;

typedef struct hspz_request_result {
    int32_t status;      /* 0 or the failing pass's status */
    hspz_segment_pair *hsps;
    size_t n_hsps;
    size_t *counts;      /* 2 per interval: fw, rc */
    size_t n_intervals;
    double t_submit, t_start, t_done; /* seconds since the server started */
    uint32_t batch_id;
    uint32_t batch_requests;
} hspz_request_result;

typedef struct hspz_server_stats {
    uint64_t batches, requests, bases, hits, rejected, failed_batches, max_batch_requests;
    uint64_t queued, in_flight_bytes;
} hspz_server_stats;

int hspz_server_new(hspz_engine **engines, size_t n_engines, const hspz_server_config *cfg,
                    hspz_server **out);
/* Queues one request (encoded fwd/rc block, or with rc = NULL the raw block;
 * copied). Refuses with HSPZ_E_OOM a
 * request whose estimated footprint exceeds the whole budget. */
int hspz_server_submit(hspz_server *s, const uint8_t *fwd, const uint8_t *rc, size_t len,
                       uint64_t *request_id);
/* Blocks until the request is done; release with hspz_request_result_free.
 * A result is handed out once: waiting again on a collected id (or a second
 * concurrent wait on it) returns HSPZ_E_ARG. */
int hspz_server_wait(hspz_server *s, uint64_t request_id, hspz_request_result *out);
void hspz_request_result_free(hspz_request_result *r);
int hspz_server_get_stats(hspz_server *s, hspz_server_stats *out);
/* Stops admitting new requests, finishes the queued ones, joins the workers. */
void hspz_server_free(hspz_server *s);

void hspz_free(void *p);

#ifdef __cplusplus
}
#endif
#endif

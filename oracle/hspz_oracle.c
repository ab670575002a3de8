/*
 * hspz_oracle.c — CPU oracle for the seed-and-filter path (test infrastructure
 * only; see hspz_oracle.h for what pins each function).
 *
 * Every function restates one piece of the reference (hspZ) and cites it.
 * Integer arithmetic that the reference performs in wrapping u32/i32 is done
 * here in uint32_t and converted, so C's signed-overflow rules never apply.
 * Build with -ffp-contract=off: the entropy branch must not fuse multiplies.
 */
#include "hspz_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

void orc_free(void *p) { free(p); }

/* ---------------------------------------------------------------- sequence */

/* sequence.rs:66-79: A0 C1 G2 T3, acgt -> L4, nN -> N5, '&' -> E7, rest X6. */
void orc_encode(const uint8_t *raw, size_t n, uint8_t *out) {
    for (size_t i = 0; i < n; i++) {
        uint8_t c = raw[i], v;
        switch (c) {
        case 'A': v = 0; break;
        case 'C': v = 1; break;
        case 'G': v = 2; break;
        case 'T': v = 3; break;
        case 'a': case 'c': case 'g': case 't': v = 4; break;
        case 'n': case 'N': v = 5; break;
        case '&': v = 7; break;
        default: v = 6; break;
        }
        out[i] = v;
    }
}

/* sequence.rs:83-86 complement table, rebuilt from its IUPAC pairs: uppercase
 * letters default to 'N', lowercase letters to 'n', '&' maps to itself and
 * every other byte (including >= 128) to 'N'. */
static uint8_t rc_table[256];
static pthread_once_t rc_once = PTHREAD_ONCE_INIT;
static void rc_init(void) {
    static const char pairs[] = "ATBVCGDHGCHDKMMKRYSSTAUAVBWWYR";
    for (int i = 0; i < 256; i++) rc_table[i] = 'N';
    for (int c = 'a'; c <= 'z'; c++) rc_table[c] = 'n';
    rc_table['&'] = '&';
    for (size_t i = 0; pairs[i]; i += 2) {
        uint8_t from = (uint8_t)pairs[i], to = (uint8_t)pairs[i + 1];
        rc_table[from] = to;
        rc_table[from + 32] = (uint8_t)(to + 32);
    }
}

/* sequence.rs:212-230 (bases only; the chr table remap lives with the caller) */
void orc_revcomp(const uint8_t *raw, size_t n, uint8_t *out) {
    pthread_once(&rc_once, rc_init);
    for (size_t i = 0; i < n; i++) out[i] = rc_table[raw[n - 1 - i]];
}

/* ---------------------------------------------------------------- scoring */

static const int32_t HOXD70[16] = {91,  -114, -31,  -123, -114, 100, -125, -31,
                                   -31, -125, 100,  -114, -123, -31, -114, 91};

/* scoring.rs:41-120, with the 4x4 block supplied directly (the LASTZ file
 * reader at scoring.rs:130-248 is input parsing and not restated). */
int orc_build_sub_mat(const char *ambiguous, int32_t xdrop, const int32_t *acgt,
                      int32_t out[64]) {
    const int32_t BAD = -1000, FILL = -100;
    const int L = 4, N = 5, X = 6, E = 7;
    if (!acgt) acgt = HOXD70;
    if (!ambiguous) ambiguous = "";

    /* fields = ambiguous.split(',') */
    char field[64] = {0};
    const char *c1 = strchr(ambiguous, ',');
    size_t flen = c1 ? (size_t)(c1 - ambiguous) : strlen(ambiguous);
    if (flen >= sizeof field) return -1;
    memcpy(field, ambiguous, flen);
    int nfields = 1;
    for (const char *p = ambiguous; *p; p++)
        if (*p == ',') nfields++;
    int32_t reward, penalty;
    if (nfields == 3) {
        const char *c2 = strchr(c1 + 1, ',');
        char *end;
        long r = strtol(c1 + 1, &end, 10);
        if (end != c2 || c2 == c1 + 1) return -1;
        long pn = strtol(c2 + 1, &end, 10);
        if (*end != 0 || *(c2 + 1) == 0) return -1;
        reward = (int32_t)r;
        penalty = (int32_t)-pn;
    } else if (!strcmp(ambiguous, "n") || !strcmp(ambiguous, "iupac")) {
        reward = 0;
        penalty = 0;
    } else {
        reward = -100;
        penalty = -100;
    }
    int is_n = !strcmp(field, "n") || !strcmp(field, "iupac");
    int is_iupac = !strcmp(field, "iupac");

    memset(out, 0, 64 * sizeof(int32_t));
    for (int i = 0; i < L; i++)
        for (int j = 0; j < L; j++) out[i * 8 + j] = acgt[i * 4 + j];
    for (int i = 0; i < L; i++) {
        out[i * 8 + L] = BAD;
        out[L * 8 + i] = BAD;
    }
    out[L * 8 + L] = BAD;
    int32_t n0 = is_n ? penalty : BAD, n1 = is_n ? reward : BAD;
    for (int i = 0; i < N; i++) {
        out[i * 8 + N] = n0;
        out[N * 8 + i] = n0;
    }
    out[N * 8 + N] = n1;
    if (is_iupac) {
        for (int i = 0; i < X; i++) {
            out[i * 8 + X] = penalty;
            out[X * 8 + i] = penalty;
        }
        out[X * 8 + X] = reward;
    } else {
        for (int i = 0; i < L; i++) {
            out[i * 8 + X] = FILL;
            out[X * 8 + i] = FILL;
        }
        for (int i = L; i < X; i++) {
            out[i * 8 + X] = BAD;
            out[X * 8 + i] = BAD;
        }
        out[X * 8 + X] = FILL;
    }
    for (int i = 0; i < E; i++) {
        out[i * 8 + E] = -10 * xdrop;
        out[E * 8 + i] = -10 * xdrop;
    }
    out[E * 8 + E] = -10 * xdrop;
    return 0;
}

/* ---------------------------------------------------------------- seeds */

/* seed.rs:47-84 */
int orc_shape_parse(const char *spec, orc_shape *s) {
    const char *pattern = spec;
    if (!strcmp(spec, "12of19")) pattern = "TTT0T00TT00T0T0TTTT";
    else if (!strcmp(spec, "14of22")) pattern = "TTT0T0TT00TT00T0T0TTTT";
    size_t n = strlen(pattern);
    if (n > 32) return -1;
    memset(s, 0, sizeof *s);
    for (size_t i = 0; i < n; i++) {
        char c = pattern[i];
        int care = (c == '1' || c == 'T');
        if (care) {
            s->pos[s->kmer_size] = (uint32_t)i;
            /* custom patterns are rewritten to all-T (seed.rs:53-56) */
            s->transition[s->kmer_size] = 1;
            s->kmer_size++;
        }
    }
    s->size = (uint32_t)n;
    if (s->kmer_size < 4 || s->kmer_size > 15) return -1;
    return 0;
}

static inline int seed_code(uint8_t c) {
    switch (c) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
    default: return -1;
    }
}

/* seed.rs:91-103: whole window must be uppercase ACGT */
uint32_t orc_kmer_at(const orc_shape *s, const uint8_t *raw, size_t pos) {
    for (uint32_t i = 0; i < s->size; i++)
        if (seed_code(raw[pos + i]) < 0) return ORC_INVALID_KMER;
    uint32_t k = 0;
    for (uint32_t t = 0; t < s->kmer_size; t++) k = (k << 2) + (uint32_t)seed_code(raw[pos + s->pos[t]]);
    return k;
}

/* seed.rs:135-180: counting sort of positions start_offset + i*step. */
int orc_seed_table_build(const uint8_t *raw, size_t n, const orc_shape *s, uint32_t step,
                         uint32_t **index_table, uint32_t **pos_table, size_t *n_pos) {
    size_t st = step ? step : 1;
    size_t offset = (s->size + 1) % st;
    size_t start_offset = st - offset;
    size_t nk = (size_t)1 << (2 * s->kmer_size);
    size_t num_steps = (n + offset) > s->size ? (n + offset - s->size) / st : 0;
    uint32_t *counts = calloc(nk + 1, sizeof(uint32_t));
    uint32_t *idxs = malloc((num_steps ? num_steps : 1) * sizeof(uint32_t));
    if (!counts || !idxs) return -1;
    for (size_t i = 0; i < num_steps; i++) {
        uint32_t k = orc_kmer_at(s, raw, start_offset + i * st);
        idxs[i] = k;
        if (k != ORC_INVALID_KMER) counts[k + 1]++;
    }
    for (size_t i = 1; i <= nk; i++) counts[i] += counts[i - 1];
    size_t total = counts[nk];
    uint32_t *pos = malloc((total ? total : 1) * sizeof(uint32_t));
    uint32_t *cursor = malloc((nk + 1) * sizeof(uint32_t));
    if (!pos || !cursor) return -1;
    memcpy(cursor, counts, (nk + 1) * sizeof(uint32_t));
    for (size_t i = 0; i < num_steps; i++) {
        uint32_t k = idxs[i];
        if (k != ORC_INVALID_KMER) pos[cursor[k]++] = (uint32_t)(start_offset + i * st);
    }
    uint32_t *index = malloc(nk * sizeof(uint32_t));
    if (!index) return -1;
    memcpy(index, counts + 1, nk * sizeof(uint32_t));
    free(counts);
    free(cursor);
    free(idxs);
    *index_table = index;
    *pos_table = pos;
    *n_pos = total;
    return 0;
}

/* seed.rs:385-398 + push_seeds seed.rs:107-122 */
size_t orc_chunk_seeds(const uint8_t *raw, const orc_shape *s, int transitions, uint32_t lo,
                       uint32_t hi, uint64_t *out) {
    size_t n = 0;
    for (uint32_t j = lo; j < hi; j++) {
        uint32_t k = orc_kmer_at(s, raw, j);
        if (k == ORC_INVALID_KMER) continue;
        out[n++] = ((uint64_t)k << 32) | j;
        if (!transitions) continue;
        for (uint32_t t = 0; t < s->kmer_size; t++) {
            if (!s->transition[t]) continue;
            uint64_t v = (uint64_t)k ^ ((uint64_t)2 << (2 * t));
            out[n++] = (v << 32) | j;
        }
    }
    return n;
}

/* sequence.rs:322-334 */
size_t orc_intervals(size_t block_len, uint32_t seed_size, uint32_t interval_size,
                     uint32_t *out, size_t cap) {
    size_t n = 0;
    if (block_len <= seed_size) return 0;
    uint32_t end_pos = (uint32_t)(block_len - seed_size);
    for (uint64_t curr = 0; curr < end_pos; curr += interval_size) {
        uint64_t e = curr + interval_size;
        if (e > end_pos) e = end_pos;
        if (n < cap) {
            out[2 * n] = (uint32_t)curr;
            out[2 * n + 1] = (uint32_t)e;
        }
        n++;
    }
    return n;
}

/* seed.rs:488-496 */
size_t orc_chunks(uint32_t start, uint32_t end, uint32_t wga_chunk, uint32_t *out, size_t cap) {
    size_t n = 0;
    for (uint64_t i = start; i < end; i += wga_chunk) {
        uint64_t e = i + wga_chunk;
        if (e > (uint64_t)end + 1) e = (uint64_t)end + 1;
        if (n < cap) {
            out[2 * n] = (uint32_t)i;
            out[2 * n + 1] = (uint32_t)e;
        }
        n++;
    }
    return n;
}

/* ---------------------------------------------------------------- extension */

/* Scalar form of the warp scans (kernels.rs:1279-1407): strictly-greater
 * earliest argmax, stop at the first cell with max - score > xdrop or at the
 * sequence end. Sums wrap like the device's i32 adds. */
void orc_extend_right(const uint8_t *ref, uint32_t ref_len, const uint8_t *qry,
                      uint32_t qry_len, const int32_t *sub, int32_t xdrop, uint32_t r0,
                      uint32_t q0, int32_t *score, int32_t *ext, uint64_t *cells) {
    int32_t c = 0, m = 0, pos = -1;
    uint64_t k = 0;
    for (;; k++) {
        if ((uint64_t)r0 + k >= ref_len || (uint64_t)q0 + k >= qry_len) break;
        c = (int32_t)((uint32_t)c + (uint32_t)sub[ref[r0 + k] * 8 + qry[q0 + k]]);
        if (c > m) {
            m = c;
            pos = (int32_t)k;
        }
        if ((int32_t)((uint32_t)m - (uint32_t)c) > xdrop) {
            k++;
            break;
        }
    }
    if (cells) *cells += k;
    *score = m;
    *ext = pos;
}

/* kernels.rs:1417-1543 */
void orc_extend_left(const uint8_t *ref, uint32_t ref_len, const uint8_t *qry,
                     uint32_t qry_len, const int32_t *sub, int32_t xdrop, uint32_t r0,
                     uint32_t q0, int32_t *score, int32_t *ext, uint64_t *cells) {
    (void)ref_len;
    (void)qry_len;
    int32_t c = 0, m = 0, pos = 0;
    uint64_t k = 1, n = 0;
    for (;; k++) {
        if (r0 < k || q0 < k) break;
        n++;
        c = (int32_t)((uint32_t)c + (uint32_t)sub[ref[r0 - k] * 8 + qry[q0 - k]]);
        if (c > m) {
            m = c;
            pos = (int32_t)k;
        }
        if ((int32_t)((uint32_t)m - (uint32_t)c) > xdrop) break;
    }
    if (cells) *cells += n;
    *score = m;
    *ext = pos;
}

/* Rust's saturating `f64 as i32` (NaN -> 0). */
static int32_t trunc_sat(double v) {
    if (v != v) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (int32_t)-2147483647 - 1;
    return (int32_t)v;
}

/* kernels.rs:1048-1051 */
int orc_gate_keep(int32_t total, int32_t hspthresh) {
    return trunc_sat((double)(float)total) >= hspthresh;
}

/* kernels.rs:1885-1908: atanh series on the mantissa folded into [1/sqrt2, sqrt2].
 * Evaluated in the reference's operation order; no contraction. */
double orc_ln(double x) {
    const double LN2 = 0.6931471805599453;
    const double SQRT2 = 1.4142135623730951;
    uint64_t bits;
    memcpy(&bits, &x, 8);
    int32_t e = (int32_t)((bits >> 52) & 0x7FF) - 1023;
    uint64_t mb = (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    double m;
    memcpy(&m, &mb, 8);
    if (m > SQRT2) {
        m *= 0.5;
        e += 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double z = s * s;
    double poly =
        1.0 + z * (1.0 / 3.0 +
                   z * (1.0 / 5.0 +
                        z * (1.0 / 7.0 + z * (1.0 / 9.0 + z * (1.0 / 11.0 +
                                                               z * (1.0 / 13.0 + z * (1.0 / 15.0)))))));
    double a = 2.0 * s;
    double b = a * poly;
    double c = (double)e * LN2;
    return b + c;
}

/* kernels.rs:1555-1702 */
int orc_materialize(const uint8_t *ref, uint32_t ref_len, const uint8_t *qry, uint32_t qry_len,
                    const int32_t *sub, int32_t xdrop, int32_t hspthresh, int noentropy,
                    uint32_t r0, uint32_t q0, orc_sp *out) {
    int32_t rs, re, ls, le;
    orc_extend_right(ref, ref_len, qry, qry_len, sub, xdrop, r0, q0, &rs, &re, NULL);
    orc_extend_left(ref, ref_len, qry, qry_len, sub, xdrop, r0, q0, &ls, &le, NULL);
    int32_t total = (int32_t)((uint32_t)rs + (uint32_t)ls);
    int32_t extent = (int32_t)((uint32_t)re + (uint32_t)le);
    uint32_t left = (uint32_t)le;
    double entropy = 1.0;
    if (total >= hspthresh && (int64_t)total <= (int64_t)3 * hspthresh && !noentropy) {
        uint32_t rstart = r0 - left, qstart = q0 - left;
        int32_t positions = extent + 1;
        uint64_t packed = 0;
        for (int32_t i = 0; i < positions; i++) {
            uint8_t r = ref[rstart + (uint32_t)i], q = qry[qstart + (uint32_t)i];
            if (r == q && r < 4) packed += (uint64_t)1 << (16 * r);
        }
        int32_t cnt[4] = {(int32_t)(packed & 0xFFFF), (int32_t)((packed >> 16) & 0xFFFF),
                          (int32_t)((packed >> 32) & 0xFFFF), (int32_t)((packed >> 48) & 0xFFFF)};
        int32_t tc = cnt[0] + cnt[1] + cnt[2] + cnt[3];
        if (tc >= 20) {
            double denom = (double)(extent + 1);
            double acc = 0.0;
            for (int i = 0; i < 4; i++) {
                double p = (double)cnt[i] / denom;
                double l = cnt[i] != 0 ? orc_ln(p) : 0.0;
                double t = p * l;
                acc = acc + t;
            }
            entropy = -acc / ORC_LOG4;
        }
    }
    double gated = (double)(float)total * entropy;
    if (trunc_sat(gated) >= hspthresh) {
        out->ref_start = r0 - left;
        out->query_start = q0 - left;
        out->len = (uint32_t)extent;
        out->score = entropy > 0.0 ? trunc_sat((double)total * entropy) : 0;
        return 1;
    }
    return 0;
}

/* ---------------------------------------------------------------- chunking */

/* gpu/mod.rs:2397-2420: the lower_bound walk. */
long orc_chunk_limits(const uint32_t *hit_num, uint32_t n, uint32_t max_hits, uint32_t *out,
                      size_t cap) {
    if (n == 0 || max_hits == 0) return -1;
    uint32_t num_hits = hit_num[n - 1];
    uint32_t num_iter = num_hits / max_hits + 1;
    uint32_t *limits = malloc((size_t)num_iter * sizeof(uint32_t));
    if (!limits) return -1;
    uint32_t lim = max_hits;
    for (uint32_t it = 0; it + 1 < num_iter; it++) {
        /* partition_point(v < lim) */
        uint32_t lo = 0, hi = n;
        while (lo < hi) {
            uint32_t mid = lo + (hi - lo) / 2;
            if (hit_num[mid] < lim) lo = mid + 1;
            else hi = mid;
        }
        if (lo == 0) {
            free(limits);
            return -1;
        }
        uint32_t pos = lo - 1;
        lim = hit_num[pos] + max_hits;
        limits[it] = pos;
    }
    limits[num_iter - 1] = n - 1;
    uint32_t start_seed = 0, start_hit = 0;
    for (uint32_t i = 0; i < num_iter; i++) {
        uint32_t pos = limits[i];
        if (i < cap) {
            out[4 * i] = start_seed;
            out[4 * i + 1] = pos;
            out[4 * i + 2] = start_hit;
            out[4 * i + 3] = hit_num[pos];
        }
        start_seed = pos + 1;
        start_hit = hit_num[pos];
    }
    free(limits);
    return num_iter;
}

/* ---------------------------------------------------------------- dedup */

static int cmp_diag(const void *a, const void *b) {
    const orc_sp *x = a, *y = b;
    uint32_t dx = x->ref_start - x->query_start, dy = y->ref_start - y->query_start;
    if (dx != dy) return dx < dy ? -1 : 1;
    if (x->ref_start != y->ref_start) return x->ref_start < y->ref_start ? -1 : 1;
    if (x->len != y->len) return x->len < y->len ? -1 : 1;
    if (x->score != y->score) return x->score > y->score ? -1 : 1;
    return 0;
}

static int cmp_lastz(const void *a, const void *b) {
    const orc_sp *x = a, *y = b;
    if (x->query_start != y->query_start) return x->query_start < y->query_start ? -1 : 1;
    if (x->ref_start != y->ref_start) return x->ref_start < y->ref_start ? -1 : 1;
    if (x->len != y->len) return x->len < y->len ? -1 : 1;
    if (x->score != y->score) return x->score > y->score ? -1 : 1;
    return 0;
}

/* hsp.rs:83-87 with u32 wrapping ends */
static int contained(const orc_sp *x, const orc_sp *y) {
    if (x->ref_start - x->query_start != y->ref_start - y->query_start) return 0;
    uint32_t xe = x->ref_start + x->len, ye = y->ref_start + y->len;
    return (x->ref_start >= y->ref_start && xe <= ye) || (y->ref_start >= x->ref_start && ye <= xe);
}

/* hsp.rs:94-126. Equal sort keys imply identical records, so qsort's lack of
 * stability cannot change the result. */
size_t orc_dedup_and_order(orc_sp *h, size_t n) {
    if (n == 0) return 0;
    qsort(h, n, sizeof(orc_sp), cmp_diag);
    orc_sp *kept = malloc(n * sizeof(orc_sp));
    size_t k = 0;
    for (size_t i = 0; i < n; i++)
        if (i == 0 || !contained(&h[i - 1], &h[i])) kept[k++] = h[i];
    qsort(kept, k, sizeof(orc_sp), cmp_lastz);
    memcpy(h, kept, k * sizeof(orc_sp));
    free(kept);
    return k;
}

/* ---------------------------------------------------------------- per batch */

static inline uint32_t bucket_start(const uint32_t *idx, uint32_t k) { return k ? idx[k - 1] : 0; }

/* gpu/mod.rs:1499-2178 (plus A.3/A.6): count, scan, chunk, extend every hit,
 * gate, materialize, dedup per chunk, append in chunk order. */
int orc_seed_and_filter(const orc_engine *e, const uint64_t *seeds, uint32_t nseeds, int rev,
                        orc_sp **out, size_t *nout, orc_stats *st) {
    *out = NULL;
    *nout = 0;
    if (nseeds == 0) return 0;
    const uint8_t *q = rev ? e->q_rc_enc : e->q_enc;
    uint32_t *hit_num = malloc((size_t)nseeds * sizeof(uint32_t));
    if (!hit_num) return -1;
    uint32_t acc = 0;
    for (uint32_t i = 0; i < nseeds; i++) {
        uint32_t k = (uint32_t)(seeds[i] >> 32);
        acc += e->index_table[k] - bucket_start(e->index_table, k);
        hit_num[i] = acc;
    }
    if (st) {
        st->seeds += nseeds;
        st->hits += acc;
    }
    if (acc == 0) {
        free(hit_num);
        return 0;
    }
    uint32_t *lim;
    long nchunks;
    uint32_t single[4];
    if (acc <= e->max_hits) {
        single[0] = 0;
        single[1] = nseeds - 1;
        single[2] = 0;
        single[3] = acc;
        lim = single;
        nchunks = 1;
    } else {
        nchunks = orc_chunk_limits(hit_num, nseeds, e->max_hits, NULL, 0);
        if (nchunks < 0) {
            free(hit_num);
            return -2;
        }
        lim = malloc((size_t)nchunks * 4 * sizeof(uint32_t));
        orc_chunk_limits(hit_num, nseeds, e->max_hits, lim, (size_t)nchunks);
    }
    size_t cap = 0, n = 0;
    orc_sp *res = NULL;
    orc_sp *tmp = NULL;
    size_t tcap = 0;
    for (long c = 0; c < nchunks; c++) {
        uint32_t s0 = lim[4 * c], s1 = lim[4 * c + 1];
        if (lim[4 * c + 3] - lim[4 * c + 2] == 0) continue;
        size_t tn = 0;
        for (uint32_t s = s0; s <= s1; s++) {
            uint32_t k = (uint32_t)(seeds[s] >> 32);
            uint32_t qloc = (uint32_t)(seeds[s] & 0xFFFFFFFFu) + e->seed_size;
            uint32_t b0 = bucket_start(e->index_table, k), b1 = e->index_table[k];
            for (uint32_t j = b0; j < b1; j++) {
                uint32_t rloc = e->pos_table[j] + e->seed_size;
                int32_t rs, re, ls, le;
                uint64_t cells = 0;
                orc_extend_right(e->ref_enc, e->ref_len, q, e->query_len, e->sub, e->xdrop, rloc,
                                 qloc, &rs, &re, &cells);
                orc_extend_left(e->ref_enc, e->ref_len, q, e->query_len, e->sub, e->xdrop, rloc,
                                qloc, &ls, &le, &cells);
                if (st) st->cells += cells;
                int32_t total = (int32_t)((uint32_t)rs + (uint32_t)ls);
                if (!orc_gate_keep(total, e->hspthresh)) continue;
                if (st) st->gate_pass++;
                orc_sp sp;
                if (!orc_materialize(e->ref_enc, e->ref_len, q, e->query_len, e->sub, e->xdrop,
                                     e->hspthresh, e->noentropy, rloc, qloc, &sp))
                    continue;
                if (tn == tcap) {
                    tcap = tcap ? 2 * tcap : 256;
                    tmp = realloc(tmp, tcap * sizeof(orc_sp));
                }
                tmp[tn++] = sp;
            }
        }
        if (st) st->raw_hsps += tn;
        size_t kept = orc_dedup_and_order(tmp, tn);
        if (n + kept > cap) {
            cap = (n + kept) * 2 + 16;
            res = realloc(res, cap * sizeof(orc_sp));
        }
        memcpy(res + n, tmp, kept * sizeof(orc_sp));
        n += kept;
    }
    if (st) st->hsps += n;
    free(tmp);
    if (lim != single) free(lim);
    free(hit_num);
    *out = res;
    *nout = n;
    return 0;
}

/* ---------------------------------------------------------------- one pass */

typedef struct {
    uint32_t interval;
    int rev;
    uint32_t lo, hi;
    orc_sp *res;
    size_t n;
    orc_stats st;
    int status;
} batch_t;

typedef struct {
    const orc_engine *e;
    const uint8_t *fwd, *rc;
    const orc_shape *shape;
    int transitions;
    batch_t *batches;
    size_t nb;
    size_t next;
    pthread_mutex_t mu;
} pass_t;

static void *pass_worker(void *arg) {
    pass_t *p = arg;
    size_t per = 1 + p->shape->kmer_size;
    uint64_t *seeds = NULL;
    size_t scap = 0;
    for (;;) {
        pthread_mutex_lock(&p->mu);
        size_t i = p->next++;
        pthread_mutex_unlock(&p->mu);
        if (i >= p->nb) break;
        batch_t *b = &p->batches[i];
        size_t need = (size_t)(b->hi - b->lo) * per;
        if (need > scap) {
            free(seeds);
            scap = need;
            seeds = malloc(scap * sizeof(uint64_t));
        }
        const uint8_t *seq = b->rev ? p->rc : p->fwd;
        size_t ns = orc_chunk_seeds(seq, p->shape, p->transitions, b->lo, b->hi, seeds);
        memset(&b->st, 0, sizeof b->st);
        b->status = orc_seed_and_filter(p->e, seeds, (uint32_t)ns, b->rev, &b->res, &b->n, &b->st);
    }
    free(seeds);
    return NULL;
}

/* run.rs:278-299 batch list; results concatenated per interval in batch
 * order (run.rs:443-451). */
int orc_run_pass(const orc_engine *e, const uint8_t *raw_fwd, const uint8_t *raw_rc,
                 uint32_t block_len, const orc_shape *shape, int transitions, int plus,
                 int minus, uint32_t wga_chunk, uint32_t interval_size, int threads,
                 orc_sp **out, size_t **counts, size_t *n_intervals, orc_stats *st) {
    size_t ni = orc_intervals(block_len, shape->size, interval_size, NULL, 0);
    uint32_t *iv = malloc((ni ? ni : 1) * 2 * sizeof(uint32_t));
    orc_intervals(block_len, shape->size, interval_size, iv, ni);
    uint32_t qbl = block_len - shape->size;
    size_t nb = 0, bcap = 64;
    batch_t *bs = malloc(bcap * sizeof(batch_t));
    for (size_t i = 0; i < ni; i++) {
        for (int strand = 0; strand < 2; strand++) {
            if (strand == 0 && !plus) continue;
            if (strand == 1 && !minus) continue;
            uint32_t a = strand ? qbl - iv[2 * i + 1] : iv[2 * i];
            uint32_t b = strand ? qbl - iv[2 * i] : iv[2 * i + 1];
            size_t nc = orc_chunks(a, b, wga_chunk, NULL, 0);
            uint32_t *ch = malloc((nc ? nc : 1) * 2 * sizeof(uint32_t));
            orc_chunks(a, b, wga_chunk, ch, nc);
            for (size_t c = 0; c < nc; c++) {
                if (nb == bcap) {
                    bcap *= 2;
                    bs = realloc(bs, bcap * sizeof(batch_t));
                }
                memset(&bs[nb], 0, sizeof(batch_t));
                bs[nb].interval = (uint32_t)i;
                bs[nb].rev = strand;
                bs[nb].lo = ch[2 * c];
                bs[nb].hi = ch[2 * c + 1];
                nb++;
            }
            free(ch);
        }
    }
    pass_t p = {e, raw_fwd, raw_rc, shape, transitions, bs, nb, 0, PTHREAD_MUTEX_INITIALIZER};
    if (threads < 1) threads = 1;
    pthread_t *th = malloc((size_t)threads * sizeof(pthread_t));
    for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, pass_worker, &p);
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);

    int status = 0;
    size_t total = 0;
    for (size_t b = 0; b < nb; b++) {
        total += bs[b].n;
        if (bs[b].status) status = bs[b].status;
        if (st) {
            st->seeds += bs[b].st.seeds;
            st->hits += bs[b].st.hits;
            st->gate_pass += bs[b].st.gate_pass;
            st->raw_hsps += bs[b].st.raw_hsps;
            st->hsps += bs[b].st.hsps;
            st->cells += bs[b].st.cells;
        }
    }
    orc_sp *all = malloc((total ? total : 1) * sizeof(orc_sp));
    size_t *cnt = calloc((ni ? ni : 1) * 2, sizeof(size_t));
    size_t at = 0;
    for (size_t i = 0; i < ni; i++)
        for (int strand = 0; strand < 2; strand++)
            for (size_t b = 0; b < nb; b++) {
                if (bs[b].interval != i || bs[b].rev != strand) continue;
                memcpy(all + at, bs[b].res, bs[b].n * sizeof(orc_sp));
                at += bs[b].n;
                cnt[2 * i + strand] += bs[b].n;
            }
    for (size_t b = 0; b < nb; b++) free(bs[b].res);
    free(bs);
    free(iv);
    *out = all;
    *counts = cnt;
    *n_intervals = ni;
    return status;
}

/* benchmark.rs:401-421 */
uint64_t orc_hash_pass(const orc_sp *all, const size_t *counts, size_t n_intervals) {
    uint64_t h = 0xcbf29ce484222325ull;
    size_t at = 0;
    for (size_t i = 0; i < 2 * n_intervals; i++) {
        for (size_t j = 0; j < counts[i]; j++, at++) {
            uint32_t v[4] = {all[at].ref_start, all[at].query_start, all[at].len,
                             (uint32_t)all[at].score};
            for (int f = 0; f < 4; f++)
                for (int b = 0; b < 4; b++) {
                    h ^= (v[f] >> (8 * b)) & 0xFF;
                    h *= 0x100000001b3ull;
                }
        }
    }
    return h;
}

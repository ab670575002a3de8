// host.cpp — host-side path functions of the C ABI (include/hspz.h):
// seed shapes, the reference index, query seeding and dedup_and_order.
//
// These restate hspZ's src/seed.rs and src/hsp.rs on the host in C++ (the
// reference's host is Rust; no Rust toolchain exists here). Parallel versions
// are byte-identical to the serial definitions by construction: contiguous
// per-thread ranges, concatenated (or cursor-offset) in range order.
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hspz.h"
#include "common.hpp"

namespace hspz {

// Shape::parse (seed.rs:47-84). Named patterns or a 1/0/T string whose care
// positions all admit transitions.
int parse_shape(const char *spec, Shape *s) {
    std::string pattern = spec ? spec : "";
    if (pattern == "12of19") pattern = "TTT0T00TT00T0T0TTTT";
    else if (pattern == "14of22") pattern = "TTT0T0TT00TT00T0T0TTTT";
    if (pattern.size() > 32) return set_error(HSPZ_E_SHAPE, "seed pattern longer than 32");
    s->size = (uint32_t)pattern.size();
    s->kmer_size = 0;
    for (size_t i = 0; i < pattern.size(); i++)
        if (pattern[i] == '1' || pattern[i] == 'T') s->pos[s->kmer_size++] = (uint32_t)i;
    if (s->kmer_size < 4 || s->kmer_size > 15)
        return set_error(HSPZ_E_SHAPE, "seed has %u care positions; 4..=15 required", s->kmer_size);
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

// Shape::kmer_at (seed.rs:91-103): the whole window must be uppercase ACGT.
static inline uint32_t kmer_at(const Shape &s, const uint8_t *raw, size_t pos) {
    for (uint32_t i = 0; i < s.size; i++)
        if (seed_code(raw[pos + i]) < 0) return INVALID_KMER;
    uint32_t k = 0;
    for (uint32_t t = 0; t < s.kmer_size; t++) k = (k << 2) + (uint32_t)seed_code(raw[pos + s.pos[t]]);
    return k;
}

static size_t chunk_seeds_serial(const uint8_t *raw, const Shape &s, bool transitions, uint32_t lo,
                                 uint32_t hi, uint64_t *out) {
    size_t n = 0;
    for (uint32_t j = lo; j < hi; j++) {
        uint32_t k = kmer_at(s, raw, j);
        if (k == INVALID_KMER) continue;
        out[n++] = ((uint64_t)k << 32) | j;
        if (!transitions) continue;
        for (uint32_t t = 0; t < s.kmer_size; t++)
            out[n++] = ((uint64_t)(k ^ (2u << (2 * t))) << 32) | j;
    }
    return n;
}

// Rolling k-mer extraction for a contiguous range of window starts: the last
// `size` bases live in a 2-bit window (first base most significant) and the
// care positions are gathered with three-or-more 13-bit lookup tables (a
// software bit-extract), while a running count of non-ACGT bytes decides
// validity. Produces exactly kmer_at(s, raw, p) for every p.
struct KmerRoller {
    const Shape &s;
    uint64_t mask;
    int nchunk;
    std::vector<uint32_t> tab;  // nchunk * 8192
    explicit KmerRoller(const Shape &sh) : s(sh) {
        mask = s.size >= 32 ? ~0ull : ((1ull << (2 * s.size)) - 1);
        uint64_t care = 0;
        for (uint32_t t = 0; t < s.kmer_size; t++) care |= 3ull << (2 * (s.size - 1 - s.pos[t]));
        nchunk = (int)((2 * s.size + 12) / 13);
        tab.assign((size_t)nchunk * 8192, 0);
        int shift = 0;
        for (int c = 0; c < nchunk; c++) {
            uint64_t m = (care >> (13 * c)) & 0x1FFF;
            for (uint32_t v = 0; v < 8192; v++) {
                uint32_t out = 0;
                int o = 0;
                for (int b = 0; b < 13; b++)
                    if ((m >> b) & 1) out |= ((v >> b) & 1u) << o++;
                tab[(size_t)c * 8192 + v] = out << shift;
            }
            shift += __builtin_popcountll(m);
        }
    }
    // Calls f(p, kmer) for every window start p = first + i*step, i in [0, count).
    template <class F>
    void run(const uint8_t *raw, size_t first, size_t step, size_t count, F &&f) const {
        if (count == 0) return;
        size_t last = first + (count - 1) * step;
        uint64_t w = 0;
        uint32_t bad = 0;
        size_t begin = first;
        // prime the window with the first size-1 bases
        for (size_t q = begin; q + 1 < begin + s.size; q++) {
            int c = seed_code(raw[q]);
            w = ((w << 2) | (uint64_t)(c < 0 ? 0 : c)) & mask;
            bad += c < 0;
        }
        for (size_t p = begin; p <= last; p++) {
            size_t q = p + s.size - 1;
            int c = seed_code(raw[q]);
            w = ((w << 2) | (uint64_t)(c < 0 ? 0 : c)) & mask;
            bad += c < 0;
            if (((p - first) % step) == 0) {
                uint32_t k = INVALID_KMER;
                if (!bad) {
                    k = 0;
                    for (int ch = 0; ch < nchunk; ch++) k |= tab[(size_t)ch * 8192 + ((w >> (13 * ch)) & 0x1FFF)];
                }
                f(p, k);
            }
            bad -= seed_code(raw[p]) < 0;
        }
    }
};

}  // namespace hspz

using namespace hspz;

extern "C" int hspz_shape_parse(const char *pattern, uint32_t *size, uint32_t *kmer_size,
                                uint32_t *pos) {
    Shape s;
    int rc = parse_shape(pattern, &s);
    if (rc) return rc;
    if (size) *size = s.size;
    if (kmer_size) *kmer_size = s.kmer_size;
    if (pos) memcpy(pos, s.pos, sizeof(uint32_t) * s.kmer_size);
    return 0;
}

// SeedTable::build_parallel (seed.rs:207-361): pass 1 counts per contiguous
// range of step indices, the per-thread counts become disjoint cursors (thread
// t's region of a bucket precedes thread t+1's), pass 2 rescans and scatters.
// Positions land in ascending order inside each bucket, as in seed.rs:135-180.
extern "C" int hspz_seed_table_build(const uint8_t *raw, size_t n, const char *pattern,
                                     uint32_t step, int threads, uint32_t *index_table,
                                     size_t index_len, uint32_t **pos_table, size_t *pos_len) {
    clear_error();
    Shape s;
    if (int rc = parse_shape(pattern, &s)) return rc;
    size_t nk = (size_t)1 << (2 * s.kmer_size);
    if (!index_table || index_len != nk || !pos_table || !pos_len)
        return set_error(HSPZ_E_ARG, "index_table must hold 4^%u entries", s.kmer_size);
    size_t st = step ? step : 1;
    size_t offset = (s.size + 1) % st;
    size_t start_offset = st - offset;
    size_t num_steps = (n + offset) > s.size ? (n + offset - s.size) / st : 0;
    if (threads < 1) threads = 1;
    size_t T = std::min<size_t>((size_t)threads, std::max<size_t>(num_steps, 1));
    if (num_steps < (1u << 16)) T = 1;
    size_t per = (num_steps + T - 1) / std::max<size_t>(T, 1);

    KmerRoller roller(s);
    std::vector<std::vector<uint32_t>> counts(T, std::vector<uint32_t>(nk, 0));
    auto pass1 = [&](size_t t) {
        size_t lo = t * per, hi = std::min(num_steps, lo + per);
        uint32_t *c = counts[t].data();
        if (lo >= hi) return;
        roller.run(raw, start_offset + lo * st, st, hi - lo, [&](size_t, uint32_t k) {
            if (k != INVALID_KMER) c[k]++;
        });
    };
    {
        std::vector<std::thread> th;
        for (size_t t = 1; t < T; t++) th.emplace_back(pass1, t);
        pass1(0);
        for (auto &x : th) x.join();
    }
    // cursors: for each k in order, thread regions in thread order
    uint64_t running = 0;
    for (size_t k = 0; k < nk; k++) {
        for (size_t t = 0; t < T; t++) {
            uint32_t c = counts[t][k];
            counts[t][k] = (uint32_t)running;
            running += c;
        }
        index_table[k] = (uint32_t)running;
    }
    if (running > 0xFFFFFFFFull) return set_error(HSPZ_E_ARG, "more than 2^32 indexed positions");
    uint32_t *pos = (uint32_t *)malloc(std::max<size_t>(running, 1) * sizeof(uint32_t));
    if (!pos) return set_error(HSPZ_E_OOM, "pos_table allocation failed");
    auto pass2 = [&](size_t t) {
        size_t lo = t * per, hi = std::min(num_steps, lo + per);
        uint32_t *cur = counts[t].data();
        if (lo >= hi) return;
        roller.run(raw, start_offset + lo * st, st, hi - lo, [&](size_t p, uint32_t k) {
            if (k != INVALID_KMER) pos[cur[k]++] = (uint32_t)p;
        });
    };
    {
        std::vector<std::thread> th;
        for (size_t t = 1; t < T; t++) th.emplace_back(pass2, t);
        pass2(0);
        for (auto &x : th) x.join();
    }
    *pos_table = pos;
    *pos_len = running;
    return 0;
}

// chunk_seeds_parts + concat_parts (seed.rs:428-477): contiguous position
// ranges seeded independently and concatenated in range order.
extern "C" int hspz_chunk_seeds(const uint8_t *raw, size_t n, const char *pattern,
                                int transitions, uint32_t lo, uint32_t hi, int threads,
                                uint64_t *out, size_t *n_out) {
    clear_error();
    Shape s;
    if (int rc = parse_shape(pattern, &s)) return rc;
    if (hi > lo && (size_t)hi - 1 + s.size > n)
        return set_error(HSPZ_E_RANGE, "seed range [%u, %u) exceeds the block", lo, hi);
    uint32_t span = hi > lo ? hi - lo : 0;
    size_t per_pos = transitions ? 1 + s.kmer_size : 1;
    if (threads <= 1 || span < 4096) {
        *n_out = chunk_seeds_serial(raw, s, transitions != 0, lo, hi, out);
        return 0;
    }
    uint32_t T = std::min<uint32_t>((uint32_t)threads, span);
    uint32_t per = (span + T - 1) / T;
    std::vector<std::vector<uint64_t>> parts(T);
    std::vector<std::thread> th;
    for (uint32_t t = 0; t < T; t++)
        th.emplace_back([&, t] {
            uint32_t a = lo + t * per, b = std::min(hi, a + per);
            if (a >= b) return;
            parts[t].resize((size_t)(b - a) * per_pos);
            parts[t].resize(chunk_seeds_serial(raw, s, transitions != 0, a, b, parts[t].data()));
        });
    for (auto &x : th) x.join();
    size_t at = 0;
    for (auto &p : parts) {
        memcpy(out + at, p.data(), p.size() * sizeof(uint64_t));
        at += p.size();
    }
    *n_out = at;
    return 0;
}

// dedup_and_order (hsp.rs:94-126). Equal sort keys imply identical records,
// so an unstable sort gives the same result as the reference's stable one.
namespace hspz {
size_t dedup_and_order(hspz_segment_pair *h, size_t n) {
    if (n == 0) return 0;
    std::sort(h, h + n, [](const hspz_segment_pair &x, const hspz_segment_pair &y) {
        uint32_t dx = x.ref_start - x.query_start, dy = y.ref_start - y.query_start;
        if (dx != dy) return dx < dy;
        if (x.ref_start != y.ref_start) return x.ref_start < y.ref_start;
        if (x.len != y.len) return x.len < y.len;
        return x.score > y.score;
    });
    auto contained = [](const hspz_segment_pair &x, const hspz_segment_pair &y) {
        if (x.ref_start - x.query_start != y.ref_start - y.query_start) return false;
        uint32_t xe = x.ref_start + x.len, ye = y.ref_start + y.len;
        return (x.ref_start >= y.ref_start && xe <= ye) || (y.ref_start >= x.ref_start && ye <= xe);
    };
    // compare against the previous INPUT element (hsp.rs:79-99)
    size_t k = 1;
    hspz_segment_pair prev = h[0];
    for (size_t i = 1; i < n; i++) {
        hspz_segment_pair cur = h[i];
        if (!contained(prev, cur)) h[k++] = cur;
        prev = cur;
    }
    std::sort(h, h + k, [](const hspz_segment_pair &x, const hspz_segment_pair &y) {
        if (x.query_start != y.query_start) return x.query_start < y.query_start;
        if (x.ref_start != y.ref_start) return x.ref_start < y.ref_start;
        if (x.len != y.len) return x.len < y.len;
        return x.score > y.score;
    });
    return k;
}
}  // namespace hspz

extern "C" int hspz_dedup_and_order(hspz_segment_pair *hsps, size_t n, size_t *n_kept) {
    clear_error();
    if (n && !hsps) return set_error(HSPZ_E_ARG, "null hsps");
    *n_kept = dedup_and_order(hsps, n);
    return 0;
}

// InitializeProcessor's MAX_HITS_PER_GB * (totalGlobalMem / GiB) in f32
// (gpu/mod.rs:2423-2432).
extern "C" uint32_t hspz_default_max_hits(uint64_t total_device_bytes) {
    float gib = total_device_bytes > 0 ? (float)total_device_bytes / 1073741824.0f : 1.0f;
    float v = 4194304.0f * gib;
    uint64_t u = v >= 18446744073709551615.0f ? ~0ull : (uint64_t)v;
    if (u < 1) u = 1;
    if (u > 0xFFFFFFFFull) u = 0xFFFFFFFFull;
    return (uint32_t)u;
}

// build_sub_mat (scoring.rs:41-120) with the ACGT block given directly (NULL =
// HOXD70, scoring.rs:26-31); the LASTZ score-file reader is input parsing and
// stays with the caller. Rows/columns outside ACGT come from the fixed BAD
// (-1000) and FILL (-100) scores, the --ambiguous spec and -10 * xdrop for the
// separator, applied in the reference's order.
extern "C" int hspz_build_sub_mat(const char *ambiguous, int32_t xdrop, const int32_t *acgt,
                                  int32_t *out) {
    clear_error();
    static const int32_t HOXD70[16] = {91,  -114, -31,  -123, -114, 100, -125, -31,
                                       -31, -125, 100,  -114, -123, -31, -114, 91};
    const int32_t BAD = -1000, FILL = -100;
    const int L = 4, N = 5, X = 6, E = 7;
    if (!out) return set_error(HSPZ_E_ARG, "null output");
    if (!acgt) acgt = HOXD70;
    std::string amb = ambiguous ? ambiguous : "";
    std::vector<std::string> fields;
    size_t at = 0;
    for (;;) {
        size_t c = amb.find(',', at);
        fields.push_back(amb.substr(at, c == std::string::npos ? std::string::npos : c - at));
        if (c == std::string::npos) break;
        at = c + 1;
    }
    const std::string &field = fields[0];
    int32_t reward, penalty;
    if (fields.size() == 3) {
        char *end = nullptr;
        long r = strtol(fields[1].c_str(), &end, 10);
        if (fields[1].empty() || *end) return set_error(HSPZ_E_ARG, "--ambiguous reward: %s", fields[1].c_str());
        long pn = strtol(fields[2].c_str(), &end, 10);
        if (fields[2].empty() || *end) return set_error(HSPZ_E_ARG, "--ambiguous penalty: %s", fields[2].c_str());
        reward = (int32_t)r;
        penalty = (int32_t)-pn;
    } else if (amb == "n" || amb == "iupac") {
        reward = penalty = 0;
    } else {
        reward = penalty = -100;
    }
    bool is_n = field == "n" || field == "iupac", is_iupac = field == "iupac";
    memset(out, 0, 64 * sizeof(int32_t));
    for (int i = 0; i < L; i++)
        for (int j = 0; j < L; j++) out[i * 8 + j] = acgt[i * 4 + j];
    for (int i = 0; i < L; i++) out[i * 8 + L] = out[L * 8 + i] = BAD;
    out[L * 8 + L] = BAD;
    for (int i = 0; i < N; i++) out[i * 8 + N] = out[N * 8 + i] = is_n ? penalty : BAD;
    out[N * 8 + N] = is_n ? reward : BAD;
    if (is_iupac) {
        for (int i = 0; i < X; i++) out[i * 8 + X] = out[X * 8 + i] = penalty;
        out[X * 8 + X] = reward;
    } else {
        for (int i = 0; i < L; i++) out[i * 8 + X] = out[X * 8 + i] = FILL;
        for (int i = L; i < X; i++) out[i * 8 + X] = out[X * 8 + i] = BAD;
        out[X * 8 + X] = FILL;
    }
    for (int i = 0; i < E; i++) out[i * 8 + E] = out[E * 8 + i] = -10 * xdrop;
    out[E * 8 + E] = -10 * xdrop;
    return 0;
}

extern "C" void hspz_free(void *p) { free(p); }

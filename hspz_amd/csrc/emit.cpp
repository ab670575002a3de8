// emit.cpp — output side of the host path (include/hspz_io.h): the numeric
// `.segments` records, their text, the diagonal partitioner (-D), the
// directory and .tar.gz sinks (-Z), the per-work-unit emitter and the pass /
// segments hashes.
//
// Restates hspZ src/hsp.rs (records, render_records, render_segments),
// src/partition.rs, src/sink.rs, run.rs's Emitter and benchmark.rs's hashes
// in C++. The archive is written straight from the formatted bytes: a GNU tar
// stream (512-byte headers, mode 0644, mtime 0, uid/gid 0) deflated by zlib
// into a gzip member, so two runs over the same lists give identical files.
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <zlib.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/hspz_io.h"
#include "common.hpp"

namespace hspz {

constexpr uint64_t FNV_OFFSET_E = 0xcbf29ce484222325ull;
inline uint64_t fnv(const void *p, size_t n, uint64_t h) {
    const uint8_t *b = (const uint8_t *)p;
    for (size_t i = 0; i < n; i++) {
        h ^= b[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

// chr_at (sequence.rs:316-318): upper_bound over starts, minus one.
inline long chr_at(const hspz_chr *chrs, size_t n, uint64_t off) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (chrs[mid].start <= off) lo = mid + 1;
        else hi = mid;
    }
    return (long)lo - 1;
}

int make_records(const hspz_segment_pair *hsps, size_t n, const hspz_chr *r, size_t nr, const hspz_chr *q, size_t nq,
                 char strand, hspz_record *out) {
    for (size_t i = 0; i < n; i++) {
        const hspz_segment_pair &e = hsps[strand == '-' ? n - 1 - i : i];
        long ri = chr_at(r, nr, e.ref_start), qi = chr_at(q, nq, e.query_start);
        if (ri < 0 || qi < 0)
            return set_error(HSPZ_E_ARG, "HSP at ref %u / query %u lies before the first record", e.ref_start,
                             e.query_start);
        out[i] = {(uint32_t)ri,
                  (uint32_t)qi,
                  (uint64_t)e.ref_start + 1 - r[ri].start,
                  (uint64_t)e.ref_start + e.len + 1 - r[ri].start,
                  (uint64_t)e.query_start + 1 - q[qi].start,
                  (uint64_t)e.query_start + e.len + 1 - q[qi].start,
                  e.score};
    }
    return 0;
}

inline char *put_u64(char *o, uint64_t v) {
    char t[24];
    int k = 0;
    do {
        t[k++] = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    while (k) *o++ = t[--k];
    return o;
}

inline char *put_i32(char *o, int32_t v) {
    if (v < 0) {
        *o++ = '-';
        return put_u64(o, (uint64_t)(-(int64_t)v));
    }
    return put_u64(o, (uint64_t)v);
}

// render_records (hsp.rs:237-254): name, start, end, name, start, end,
// strand, score, tab-separated, one line per record.
int render(const hspz_record *recs, size_t n, const hspz_chr *r, size_t nr, const hspz_chr *q, size_t nq, char strand,
           std::string &out) {
    std::vector<size_t> rl(nr), ql(nq);
    for (size_t i = 0; i < nr; i++) rl[i] = strlen(r[i].name);
    for (size_t i = 0; i < nq; i++) ql[i] = strlen(q[i].name);
    size_t at = out.size();
    for (size_t i = 0; i < n; i++) {
        const hspz_record &x = recs[i];
        if (x.r_chr >= nr || x.q_chr >= nq) return set_error(HSPZ_E_ARG, "record %zu names a missing chromosome", i);
        out.resize(at + rl[x.r_chr] + ql[x.q_chr] + 4 * 21 + 12 + 8);
        char *o = &out[at];
        memcpy(o, r[x.r_chr].name, rl[x.r_chr]);
        o += rl[x.r_chr];
        *o++ = '\t';
        o = put_u64(o, x.r_start);
        *o++ = '\t';
        o = put_u64(o, x.r_end);
        *o++ = '\t';
        memcpy(o, q[x.q_chr].name, ql[x.q_chr]);
        o += ql[x.q_chr];
        *o++ = '\t';
        o = put_u64(o, x.q_start);
        *o++ = '\t';
        o = put_u64(o, x.q_end);
        *o++ = '\t';
        *o++ = strand;
        *o++ = '\t';
        o = put_i32(o, x.score);
        *o++ = '\n';
        at = (size_t)(o - &out[0]);
    }
    out.resize(at);
    return 0;
}

int render_pairs(const hspz_segment_pair *hsps, size_t n, const hspz_chr *r, size_t nr, const hspz_chr *q, size_t nq,
                 char strand, std::string &out) {
    std::vector<hspz_record> recs(n);
    if (int rc = make_records(hsps, n, r, nr, q, nq, strand, recs.data())) return rc;
    return render(recs.data(), n, r, nr, q, nq, strand, out);
}

// ---- partition.rs ---------------------------------------------------------

constexpr uint64_t MIN_CHUNK = 5000;   // partition.rs:31
constexpr uint64_t MAX_CHUNK = 50000;  // partition.rs:33

// statistics.quantiles(data) with n=4 and the exclusive method over a sorted
// sample (partition.rs:58-80): linear interpolation at k(n+1)/4, clamped.
double quantile_exclusive(const std::vector<uint64_t> &s, int k) {
    const size_t n = s.size();
    double pos = k * (double)(n + 1) / 4.0;
    if (pos < 1.0) return (double)s[0];
    if (pos >= (double)n) return (double)s[n - 1];
    size_t lo = (size_t)pos;
    double frac = pos - (double)lo;
    double a = (double)s[lo - 1], b = (double)s[lo];
    return a + (b - a) * frac;
}

// Partitioner::chunk_size_for (partition.rs:84-102); 0 = emit whole.
int chunk_size_for(const uint64_t *hist, size_t nh, uint64_t count, uint64_t *chunk) {
    *chunk = 0;
    if (count < MIN_CHUNK) return 0;
    uint64_t c = MAX_CHUNK;
    if (nh >= 2) {
        std::vector<uint64_t> s(hist, hist + nh);
        std::sort(s.begin(), s.end());
        // under 7 sizes the median, otherwise the upper quartile
        double pick = quantile_exclusive(s, s.size() < 7 ? 2 : 3);
        c = std::min<uint64_t>(pick > 0 ? (uint64_t)pick : 0, MAX_CHUNK);
    }
    if (count <= c) return 0;
    if (c == 0) return set_error(HSPZ_E_ARG, "observed sizes give a chunk size of 0");
    *chunk = c;
    return 0;
}

// Record::diagonal_key (hsp.rs:194-209): midpoints from the reference span's
// half length; plus strand (sum, r_mid), minus strand (difference, r_mid).
struct DiagKey {
    int64_t a, b;
    bool operator<(const DiagKey &o) const { return a < o.a || (a == o.a && b < o.b); }
};
inline DiagKey diagonal_key(const hspz_record &r, char strand) {
    int64_t half = (int64_t)((r.r_end - r.r_start) / 2);
    int64_t rm = (int64_t)r.r_start + half, qm = (int64_t)r.q_start + half;
    return strand == '-' ? DiagKey{qm - rm, rm} : DiagKey{qm + rm, rm};
}

// split (partition.rs:122-192). Pairs (r_chr, q_chr) in first-appearance
// order; with more than one pair, those of at most `chunk` records are held
// back and packed first-fit in ascending (size, pair) order, each pack then
// ordered by the first appearance of its query records; every other pair is
// sorted (stably) by the diagonal key and cut into pieces of `chunk`.
int split(const hspz_record *recs, size_t n, char strand, uint64_t chunk, std::vector<std::vector<hspz_record>> &parts) {
    if (!chunk) return set_error(HSPZ_E_ARG, "chunk must be positive");
    struct Pair {
        uint32_t r, q;
        std::vector<hspz_record> v;
    };
    std::vector<Pair> pairs;
    std::unordered_map<uint64_t, size_t> where;
    for (size_t i = 0; i < n; i++) {
        uint64_t key = (uint64_t)recs[i].r_chr << 32 | recs[i].q_chr;
        auto it = where.find(key);
        if (it == where.end()) {
            it = where.emplace(key, pairs.size()).first;
            pairs.push_back({recs[i].r_chr, recs[i].q_chr, {}});
        }
        pairs[it->second].v.push_back(recs[i]);
    }
    std::unordered_map<uint32_t, size_t> qrank;
    for (const Pair &p : pairs) qrank.emplace(p.q, qrank.size());
    const bool multi = pairs.size() > 1;
    std::vector<Pair *> skip;
    for (Pair &p : pairs) {
        if (multi && p.v.size() <= chunk) {
            skip.push_back(&p);
            continue;
        }
        std::vector<DiagKey> keys(p.v.size());
        std::vector<size_t> idx(p.v.size());
        for (size_t i = 0; i < p.v.size(); i++) keys[i] = diagonal_key(p.v[i], strand), idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return keys[a] < keys[b]; });
        for (size_t s = 0; s < idx.size(); s += chunk) {
            std::vector<hspz_record> piece;
            for (size_t i = s; i < std::min<size_t>(idx.size(), s + chunk); i++) piece.push_back(p.v[idx[i]]);
            parts.push_back(std::move(piece));
        }
    }
    if (!skip.empty()) {
        std::stable_sort(skip.begin(), skip.end(), [](const Pair *a, const Pair *b) {
            if (a->v.size() != b->v.size()) return a->v.size() < b->v.size();
            return a->r != b->r ? a->r < b->r : a->q < b->q;
        });
        std::vector<std::vector<Pair *>> bins(1);
        uint64_t cur = 0;
        for (Pair *p : skip) {
            if (cur + p->v.size() <= chunk) {
                cur += p->v.size();
            } else {
                bins.emplace_back();
                cur = p->v.size();
            }
            bins.back().push_back(p);
        }
        for (auto &bin : bins) {
            if (bin.empty()) continue;
            std::stable_sort(bin.begin(), bin.end(),
                             [&](const Pair *a, const Pair *b) { return qrank[a->q] < qrank[b->q]; });
            std::vector<hspz_record> agg;
            for (Pair *p : bin) agg.insert(agg.end(), p->v.begin(), p->v.end());
            parts.push_back(std::move(agg));
        }
    }
    return 0;
}

// ---- sink.rs --------------------------------------------------------------

int mkdirs(const std::string &dir) {
    if (dir.empty()) return 0;
    std::string cur;
    for (size_t i = 0; i <= dir.size(); i++) {
        if (i == dir.size() || dir[i] == '/') {
            if (!cur.empty() && mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST)
                return set_error(HSPZ_E_IO, "%s: cannot create directory", cur.c_str());
        }
        if (i < dir.size()) cur.push_back(dir[i]);
    }
    struct stat st;
    if (stat(dir.c_str(), &st) != 0 || !S_ISDIR(st.st_mode)) return set_error(HSPZ_E_IO, "%s: not a directory", dir.c_str());
    return 0;
}

struct Sink {
    virtual ~Sink() {}
    virtual int write_entry(const std::string &name, const std::string &bytes) = 0;
    virtual int finish() = 0;
    virtual uint64_t bytes_out() = 0;
    uint64_t bytes_in = 0;
};

// DirectorySink (sink.rs:33-71): one file per entry.
struct DirSink : Sink {
    std::string dir;
    int write_entry(const std::string &name, const std::string &bytes) override {
        std::string path = dir + "/" + name;
        FILE *f = fopen(path.c_str(), "wb");
        if (!f) return set_error(HSPZ_E_IO, "%s: cannot create", path.c_str());
        size_t w = bytes.empty() ? 0 : fwrite(bytes.data(), 1, bytes.size(), f);
        int e = fclose(f);
        if (w != bytes.size() || e) return set_error(HSPZ_E_IO, "%s: write failed", path.c_str());
        bytes_in += bytes.size();
        return 0;
    }
    int finish() override { return 0; }
    uint64_t bytes_out() override { return bytes_in; }
};

// TarGzSink (sink.rs:73-124): a GNU tar stream through one gzip member at
// zlib's default level; finish() writes the two zero blocks that end the
// archive, then the gzip trailer.
struct TarGzSink : Sink {
    std::string path;
    FILE *f = nullptr;
    z_stream z;
    bool zinit = false;
    std::vector<uint8_t> obuf = std::vector<uint8_t>(1u << 20);
    uint64_t written = 0;

    int open() {
        size_t slash = path.rfind('/');
        if (slash != std::string::npos && slash > 0)
            if (int rc = mkdirs(path.substr(0, slash))) return rc;
        f = fopen(path.c_str(), "wb");
        if (!f) return set_error(HSPZ_E_IO, "%s: cannot create", path.c_str());
        memset(&z, 0, sizeof z);
        if (deflateInit2(&z, Z_DEFAULT_COMPRESSION, Z_DEFLATED, 15 + 16, 8, Z_DEFAULT_STRATEGY) != Z_OK)
            return set_error(HSPZ_E_IO, "zlib init failed");
        zinit = true;
        return 0;
    }
    ~TarGzSink() override {
        if (zinit) deflateEnd(&z);
        if (f) fclose(f);
    }
    int feed(const void *p, size_t n, int flush) {
        z.next_in = (Bytef *)p;
        z.avail_in = (uInt)n;
        do {
            z.next_out = obuf.data();
            z.avail_out = (uInt)obuf.size();
            int r = deflate(&z, flush);
            if (r == Z_STREAM_ERROR) return set_error(HSPZ_E_IO, "%s: deflate failed", path.c_str());
            size_t have = obuf.size() - z.avail_out;
            if (have && fwrite(obuf.data(), 1, have, f) != have) return set_error(HSPZ_E_IO, "%s: write failed", path.c_str());
            written += have;
        } while (z.avail_out == 0 || (flush == Z_FINISH && z.avail_in));
        return 0;
    }
    int feed_all(const void *p, size_t n) {
        const uint8_t *b = (const uint8_t *)p;
        while (n) {
            size_t k = std::min<size_t>(n, 1u << 30);
            if (int rc = feed(b, k, Z_NO_FLUSH)) return rc;
            b += k, n -= k;
        }
        return 0;
    }
    static void octal(char *field, size_t width, uint64_t v) {  // width-1 digits, then NUL
        for (size_t i = width - 1; i-- > 0;) {
            field[i] = (char)('0' + (v & 7));
            v >>= 3;
        }
        field[width - 1] = 0;
    }
    int header(const std::string &name, uint64_t size, char type) {
        char h[512];
        memset(h, 0, sizeof h);
        memcpy(h, name.data(), std::min<size_t>(name.size(), 100));
        octal(h + 100, 8, 0644);
        octal(h + 108, 8, 0);
        octal(h + 116, 8, 0);
        octal(h + 124, 12, size);
        octal(h + 136, 12, 0);
        h[156] = type;
        memcpy(h + 257, "ustar  ", 8);  // GNU magic "ustar " and version " \0"
        memset(h + 148, ' ', 8);
        unsigned sum = 0;
        for (int i = 0; i < 512; i++) sum += (uint8_t)h[i];
        octal(h + 148, 7, sum);  // six digits, NUL, then the space already there
        return feed_all(h, 512);
    }
    int pad(uint64_t n) {
        static const char zeros[512] = {0};
        return n % 512 ? feed_all(zeros, 512 - n % 512) : 0;
    }
    int write_entry(const std::string &name, const std::string &bytes) override {
        // emitted names are at most ~60 bytes; the header's name field holds 100
        if (name.size() > 100) return set_error(HSPZ_E_ARG, "%s: entry name longer than 100 bytes", name.c_str());
        if (int rc = header(name, bytes.size(), '0')) return rc;
        if (int rc = feed_all(bytes.data(), bytes.size())) return rc;
        if (int rc = pad(bytes.size())) return rc;
        bytes_in += bytes.size();
        return 0;
    }
    int finish() override {
        static const char zeros[1024] = {0};
        if (int rc = feed(zeros, sizeof zeros, Z_FINISH)) return rc;
        deflateEnd(&z);
        zinit = false;
        int e = fclose(f);
        f = nullptr;
        return e ? set_error(HSPZ_E_IO, "%s: close failed", path.c_str()) : 0;
    }
    uint64_t bytes_out() override { return written; }
};

double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

}  // namespace hspz

using namespace hspz;

struct hspz_emitter {
    Sink *sink = nullptr;
    bool diagonal = false;
    std::vector<uint64_t> history;  // Partitioner's observed sizes (partition.rs:37-40)
    double partition_ms = 0, format_ms = 0, archive_ms = 0;
    uint64_t files = 0;
    ~hspz_emitter() { delete sink; }
};

extern "C" int hspz_records(const hspz_segment_pair *hsps, size_t n, const hspz_chr *r_chrs, size_t n_r,
                            const hspz_chr *q_chrs, size_t n_q, char strand, hspz_record *out) {
    clear_error();
    if (n && (!hsps || !out || !r_chrs || !q_chrs)) return set_error(HSPZ_E_ARG, "null argument");
    return make_records(hsps, n, r_chrs, n_r, q_chrs, n_q, strand, out);
}

static int text_out(const std::string &s, char **text, size_t *len) {
    char *p = (char *)malloc(s.size() + 1);
    if (!p) return set_error(HSPZ_E_OOM, "text allocation failed");
    memcpy(p, s.data(), s.size());
    p[s.size()] = 0;
    *text = p;
    *len = s.size();
    return 0;
}

extern "C" int hspz_render_records(const hspz_record *recs, size_t n, const hspz_chr *r_chrs, size_t n_r,
                                   const hspz_chr *q_chrs, size_t n_q, char strand, char **text, size_t *len) {
    clear_error();
    if (!text || !len || (n && !recs)) return set_error(HSPZ_E_ARG, "null argument");
    std::string s;
    if (int rc = render(recs, n, r_chrs, n_r, q_chrs, n_q, strand, s)) return rc;
    return text_out(s, text, len);
}

extern "C" int hspz_render_segments(const hspz_segment_pair *hsps, size_t n, const hspz_chr *r_chrs, size_t n_r,
                                    const hspz_chr *q_chrs, size_t n_q, char strand, char **text, size_t *len) {
    clear_error();
    if (!text || !len || (n && !hsps)) return set_error(HSPZ_E_ARG, "null argument");
    std::string s;
    if (int rc = render_pairs(hsps, n, r_chrs, n_r, q_chrs, n_q, strand, s)) return rc;
    return text_out(s, text, len);
}

extern "C" int hspz_partition_chunk_size(const uint64_t *history, size_t n_history, uint64_t count, uint64_t *chunk) {
    clear_error();
    if (!chunk || (n_history && !history)) return set_error(HSPZ_E_ARG, "null argument");
    return chunk_size_for(history, n_history, count, chunk);
}

extern "C" int hspz_partition_split(const hspz_record *recs, size_t n, char strand, uint64_t chunk, hspz_record *out,
                                    uint64_t *sizes, size_t *n_parts) {
    clear_error();
    if (!n_parts || (n && (!recs || !out || !sizes))) return set_error(HSPZ_E_ARG, "null argument");
    std::vector<std::vector<hspz_record>> parts;
    if (int rc = split(recs, n, strand, chunk, parts)) return rc;
    size_t at = 0;
    for (size_t i = 0; i < parts.size(); i++) {
        memcpy(out + at, parts[i].data(), parts[i].size() * sizeof(hspz_record));
        at += parts[i].size();
        sizes[i] = parts[i].size();
    }
    *n_parts = parts.size();
    return 0;
}

extern "C" uint64_t hspz_hash_pass(const hspz_segment_pair *hsps, const size_t *counts, size_t n_intervals) {
    uint64_t h = FNV_OFFSET_E;
    size_t total = 0;
    for (size_t i = 0; i < 2 * n_intervals; i++) total += counts[i];
    for (size_t i = 0; i < total; i++) {
        uint32_t v[4] = {hsps[i].ref_start, hsps[i].query_start, hsps[i].len, (uint32_t)hsps[i].score};
        h = fnv(v, 16, h);
    }
    return h;
}

extern "C" int hspz_hash_segments(const hspz_segment_pair *hsps, const size_t *counts, size_t n_intervals,
                                  const hspz_chr *r_chrs, size_t n_r, const hspz_chr *q_chrs, size_t n_q,
                                  const hspz_chr *rc_chrs, size_t n_rc, uint64_t *hash) {
    clear_error();
    if (!hash || (n_intervals && !counts)) return set_error(HSPZ_E_ARG, "null argument");
    uint64_t h = FNV_OFFSET_E;
    size_t at = 0;
    std::string s;
    for (size_t i = 0; i < 2 * n_intervals; i++) {
        bool rc = i & 1;
        s.clear();
        if (int e = render_pairs(hsps + at, counts[i], r_chrs, n_r, rc ? rc_chrs : q_chrs, rc ? n_rc : n_q,
                                 rc ? '-' : '+', s))
            return e;
        h = fnv(s.data(), s.size(), h);
        at += counts[i];
    }
    *hash = h;
    return 0;
}

extern "C" int hspz_emitter_new(const char *output_dir, const char *tarball, int diagonal, hspz_emitter **out) {
    clear_error();
    if (!out || (!output_dir && (!tarball || !*tarball || !strcmp(tarball, "-"))))
        return set_error(HSPZ_E_ARG, "null argument");
    *out = nullptr;
    hspz_emitter *em = new hspz_emitter;
    em->diagonal = diagonal != 0;
    if (tarball) {
        TarGzSink *t = new TarGzSink;
        t->path = (!*tarball || !strcmp(tarball, "-")) ? std::string(output_dir) + ".tar.gz" : std::string(tarball);
        em->sink = t;
        if (int rc = t->open()) {
            delete em;
            return rc;
        }
    } else {
        DirSink *d = new DirSink;
        d->dir = output_dir;
        em->sink = d;
        if (int rc = mkdirs(d->dir)) {
            delete em;
            return rc;
        }
    }
    *out = em;
    return 0;
}

extern "C" int hspz_emit_unit(hspz_emitter *em, uint32_t ref_bin, uint32_t query_bin, const hspz_chr *ref_chrs,
                              size_t n_ref, const hspz_chr *query_chrs, size_t n_query, const hspz_chr *rc_chrs,
                              size_t n_rc, const hspz_segment_pair *hsps, const size_t *counts, size_t n_intervals) {
    clear_error();
    if (!em || (n_intervals && !counts)) return set_error(HSPZ_E_ARG, "null argument");
    using clk = std::chrono::steady_clock;
    size_t at = 0;
    char base[96];
    for (size_t n = 0; n < n_intervals; n++) {
        snprintf(base, sizeof base, "tmp%zu.block%u.r%u", n + 1, query_bin, ref_bin);
        for (int side = 0; side < 2; side++) {
            size_t cnt = counts[2 * n + side];
            const hspz_segment_pair *list = hsps + at;
            at += cnt;
            if (!cnt) continue;
            const char strand = side ? '-' : '+';
            const hspz_chr *q = side ? rc_chrs : query_chrs;
            size_t nq = side ? n_rc : n_query;
            auto t = clk::now();
            std::vector<hspz_record> recs(cnt);
            if (int rc = make_records(list, cnt, ref_chrs, n_ref, q, nq, strand, recs.data())) return rc;
            std::vector<std::vector<hspz_record>> parts;
            bool whole = true;
            if (em->diagonal) {  // Partitioner::plan (partition.rs:110-118)
                em->history.push_back(cnt);
                uint64_t chunk = 0;
                if (int rc = chunk_size_for(em->history.data(), em->history.size(), cnt, &chunk)) return rc;
                if (chunk) {
                    whole = false;
                    if (int rc = split(recs.data(), cnt, strand, chunk, parts)) return rc;
                }
            }
            em->partition_ms += ms_since(t);
            std::string stem = std::string(base) + (side ? ".minus" : ".plus");
            auto emit_file = [&](const hspz_record *r, size_t k, const std::string &name) -> int {
                auto t0 = clk::now();
                std::string text;
                text.reserve(k * 64);
                if (int rc = render(r, k, ref_chrs, n_ref, q, nq, strand, text)) return rc;
                em->format_ms += ms_since(t0);
                auto t1 = clk::now();
                int rc = em->sink->write_entry(name, text);
                em->archive_ms += ms_since(t1);
                em->files++;
                return rc;
            };
            if (whole) {
                if (int rc = emit_file(recs.data(), cnt, stem + ".segments")) return rc;
            } else {
                for (size_t k = 0; k < parts.size(); k++)
                    if (int rc = emit_file(parts[k].data(), parts[k].size(),
                                           stem + ".split" + std::to_string(k + 1) + ".segments"))
                        return rc;
            }
        }
    }
    return 0;
}

extern "C" int hspz_emitter_finish(hspz_emitter *em, hspz_output_report *report) {
    clear_error();
    if (!em) return set_error(HSPZ_E_ARG, "null argument");
    auto t = std::chrono::steady_clock::now();
    int rc = em->sink->finish();
    em->archive_ms += ms_since(t);
    if (report) {
        report->partition_ms = em->partition_ms;
        report->format_ms = em->format_ms;
        report->archive_ms = em->archive_ms;
        report->files = em->files;
        report->bytes_in = em->sink->bytes_in;
        report->bytes_out = em->sink->bytes_out();
    }
    delete em;
    return rc;
}

// io.cpp — input side of the host path (include/hspz_io.h): format detection,
// the FASTA / gzipped FASTA / UCSC 2bit readers, pack, reverse complement, the
// set digest, encode and intervals.
//
// Restates hspZ src/sequence.rs in C++ (the reference's host is Rust; no Rust
// toolchain exists here). The readers write straight into the packed block
// (records joined by '&', trailing separator kept in the buffer and left out of
// block_len), which is what pack() makes of the reader's record list
// (sequence.rs:185-205), without a second copy of a multi-gigabase genome.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include <algorithm>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hspz_io.h"
#include "common.hpp"

// Growable byte buffer that leaves new bytes uninitialised (the readers write
// every byte they append).
struct ByteBuf {
    uint8_t *p = nullptr;
    size_t n = 0, cap = 0;
    ByteBuf() = default;
    ByteBuf(const ByteBuf &) = delete;
    ByteBuf &operator=(const ByteBuf &) = delete;
    ~ByteBuf() { free(p); }
    bool reserve(size_t want) {
        if (want <= cap) return true;
        uint8_t *q = (uint8_t *)realloc(p, want);
        if (!q) return false;
        p = q;
        cap = want;
        return true;
    }
    // room for k more bytes at p + n (size grows only via commit)
    uint8_t *room(size_t k) {
        if (n + k > cap && !reserve(std::max(n + k, cap + cap / 2 + 4096))) throw std::bad_alloc();
        return p + n;
    }
    void commit(size_t k) { n += k; }
    void push(uint8_t b) {
        *room(1) = b;
        n++;
    }
    size_t size() const { return n; }
    bool empty() const { return n == 0; }
    uint8_t *data() const { return p; }
};

struct hspz_seqset {
    ByteBuf buf;
    std::vector<std::string> names;
    std::vector<uint64_t> starts;
    std::vector<uint32_t> lens;
    std::vector<hspz_chr> chrs;  // views of names/starts/lens, rebuilt by finish()
    size_t block_len = 0;
    int32_t format = HSPZ_FORMAT_FASTA;
    uint64_t bytes_read = 0;
};

namespace hspz {

constexpr uint8_t SEP = '&';  // sequence.rs:29

// FNV-1a, 64-bit (sequence.rs:289-306, benchmark.rs:401-409)
constexpr uint64_t FNV_OFFSET = 0xcbf29ce484222325ull;
inline uint64_t fnv1a(const void *p, size_t n, uint64_t h) {
    const uint8_t *b = (const uint8_t *)p;
    for (size_t i = 0; i < n; i++) {
        h ^= b[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

// Rust's u8::is_ascii_whitespace: space, \t, \n, \x0C, \r (not \x0B).
inline bool ascii_ws(uint8_t c) { return c == ' ' || c == '\t' || c == '\n' || c == '\x0C' || c == '\r'; }

// String::from_utf8_lossy: each maximal invalid subsequence becomes U+FFFD.
std::string utf8_lossy(const uint8_t *s, size_t n) {
    std::string out;
    out.reserve(n);
    size_t i = 0;
    while (i < n) {
        uint8_t c = s[i];
        if (c < 0x80) {
            out.push_back((char)c);
            i++;
            continue;
        }
        size_t need;
        uint8_t lo = 0x80, hi = 0xBF;  // range of the first continuation byte
        if (c >= 0xC2 && c <= 0xDF) need = 1;
        else if (c >= 0xE0 && c <= 0xEF) {
            need = 2;
            if (c == 0xE0) lo = 0xA0;
            if (c == 0xED) hi = 0x9F;
        } else if (c >= 0xF0 && c <= 0xF4) {
            need = 3;
            if (c == 0xF0) lo = 0x90;
            if (c == 0xF4) hi = 0x8F;
        } else {
            out += "\xEF\xBF\xBD";
            i++;
            continue;
        }
        size_t k = 1;
        bool ok = true;
        for (; k <= need; k++) {
            if (i + k >= n) { ok = false; break; }
            uint8_t d = s[i + k];
            uint8_t l = k == 1 ? lo : 0x80, h = k == 1 ? hi : 0xBF;
            if (d < l || d > h) { ok = false; break; }
        }
        if (ok) {
            out.append((const char *)s + i, need + 1);
            i += need + 1;
        } else {
            out += "\xEF\xBF\xBD";
            i += k;  // the valid prefix is consumed; the failing byte starts over
        }
    }
    return out;
}

// Builds a packed set record by record.
struct Packer {
    hspz_seqset *s;
    std::string prefix;
    bool open = false;
    void begin(const std::string &name) {
        end();
        s->names.push_back(prefix + name);
        s->starts.push_back(s->buf.size());
        open = true;
    }
    void end() {
        if (!open) return;
        s->lens.push_back((uint32_t)(s->buf.size() - s->starts.back()));  // `seq.len() as u32`
        s->buf.push(SEP);
        open = false;
    }
};

void finish(hspz_seqset *s, bool packed) {
    s->chrs.resize(s->names.size());
    for (size_t i = 0; i < s->names.size(); i++) s->chrs[i] = {s->names[i].c_str(), s->starts[i], s->lens[i]};
    if (packed) s->block_len = s->buf.empty() ? 0 : s->buf.size() - 1;
}

// Genome::digest (sequence.rs:289-306): FNV-1a over each record's name, start
// (usize, little-endian) and len (u32), then the buffer, then block_len.
uint64_t set_digest(const hspz_seqset *s) {
    uint64_t h = FNV_OFFSET;
    for (size_t i = 0; i < s->names.size(); i++) {
        h = fnv1a(s->names[i].data(), s->names[i].size(), h);
        uint64_t st = s->starts[i];
        uint32_t ln = s->lens[i];
        h = fnv1a(&st, 8, h);  // usize, little-endian
        h = fnv1a(&ln, 4, h);
    }
    h = fnv1a(s->buf.data(), s->buf.size(), h);
    uint64_t bl = s->block_len;
    return fnv1a(&bl, 8, h);
}

// One FASTA line (sequence.rs:395-414): '>' starts a record named by the first
// whitespace-separated token; any other line appends its non-whitespace bytes
// to the open record (lines before the first header are dropped).
inline void fasta_line(Packer &pk, const uint8_t *line, size_t n) {
    if (n && line[0] == '>') {
        size_t a = 1;
        while (a < n && ascii_ws(line[a])) a++;
        size_t b = a;
        while (b < n && !ascii_ws(line[b])) b++;
        pk.begin(utf8_lossy(line + a, b - a));
        return;
    }
    if (!pk.open || !n) return;
    uint8_t *o = pk.s->buf.room(n);
    // a sequence line rarely holds whitespace: copy it whole while testing,
    // and compact it only when the test fires
    uint8_t ws = 0;
    for (size_t i = 0; i < n; i++) {
        uint8_t c = line[i];
        o[i] = c;
        ws |= (uint8_t)((c == ' ') | (c == '\t') | (c == '\n') | (c == '\x0C') | (c == '\r'));
    }
    size_t k = n;
    if (ws) {
        k = 0;
        for (size_t i = 0; i < n; i++) {
            o[k] = line[i];
            k += !ascii_ws(line[i]);
        }
    }
    pk.s->buf.commit(k);
}

// Streams a FASTA body through fasta_line; `read` fills a block and returns
// its size (0 at the end, < 0 on error).
template <class Read>
int fasta_stream(Packer &pk, Read read) {
    std::vector<uint8_t> block(4u << 20);
    std::vector<uint8_t> carry;
    for (;;) {
        long n = read(block.data(), block.size());
        if (n < 0) return -1;
        if (n == 0) break;
        const uint8_t *p = block.data(), *e = p + n;
        while (p < e) {
            const uint8_t *nl = (const uint8_t *)memchr(p, '\n', (size_t)(e - p));
            if (!nl) {
                carry.insert(carry.end(), p, e);
                break;
            }
            if (carry.empty()) {
                fasta_line(pk, p, (size_t)(nl - p));
            } else {
                carry.insert(carry.end(), p, nl);
                fasta_line(pk, carry.data(), carry.size());
                carry.clear();
            }
            p = nl + 1;
        }
    }
    if (!carry.empty()) fasta_line(pk, carry.data(), carry.size());
    pk.end();
    return 0;
}

int read_fasta_plain(const char *path, Packer &pk) {
    FILE *f = fopen(path, "rb");
    if (!f) return set_error(HSPZ_E_IO, "%s: cannot open", path);
    std::unique_ptr<FILE, int (*)(FILE *)> closer(f, fclose);
    if (fseek(f, 0, SEEK_END) == 0) {
        long sz = ftell(f);
        if (sz > 0 && !pk.s->buf.reserve((size_t)sz + 1))
            return set_error(HSPZ_E_OOM, "%s: no memory for %ld bytes", path, sz);
        fseek(f, 0, SEEK_SET);
    }
    int rc = fasta_stream(pk, [&](uint8_t *b, size_t cap) -> long {
        size_t n = fread(b, 1, cap, f);
        return ferror(f) ? -1 : (long)n;
    });
    return rc ? set_error(HSPZ_E_IO, "%s: read error", path) : 0;
}

// gzread decodes concatenated members in sequence, as a multi-member decoder.
int read_fasta_gz(const char *path, Packer &pk) {
    gzFile g = gzopen(path, "rb");
    if (!g) return set_error(HSPZ_E_IO, "%s: cannot open", path);
    std::unique_ptr<gzFile_s, int (*)(gzFile)> closer(g, gzclose);
    gzbuffer(g, 1u << 20);
    int rc = fasta_stream(pk, [&](uint8_t *b, size_t cap) -> long {
        int n = gzread(g, b, (unsigned)std::min<size_t>(cap, 1u << 30));
        return n < 0 ? -1 : (long)n;
    });
    int err = 0;
    const char *msg = gzerror(g, &err);
    std::string m = msg ? msg : "";
    if (rc || (err != Z_OK && err != Z_BUF_ERROR)) return set_error(HSPZ_E_IO, "%s: gzip: %s", path, m.c_str());
    return 0;
}

// UCSC 2bit (twoBit.c / the published format): header {magic 0x1A412743 in the
// writer's byte order, version (0: 32-bit offsets, 1: 64-bit), count,
// reserved}; index {u8 name length, name, offset}; at each offset {dnaSize,
// nBlockCount, starts[], sizes[], maskBlockCount, starts[], sizes[], reserved,
// 2-bit bases, T0 C1 A2 G3, first base in the top bits}. N blocks read as 'N',
// then mask blocks are lowercased (so a masked N reads 'n'), in index order
// (sequence.rs:161-174 reads with soft masks kept).
struct Cursor {
    const uint8_t *p;
    size_t n, at;
    bool swap, bad = false;
    uint32_t u32() {
        if (at + 4 > n) { bad = true; return 0; }
        uint32_t v;
        memcpy(&v, p + at, 4);
        at += 4;
        return swap ? __builtin_bswap32(v) : v;
    }
    uint64_t u64() {
        if (at + 8 > n) { bad = true; return 0; }
        uint64_t v;
        memcpy(&v, p + at, 8);
        at += 8;
        return swap ? __builtin_bswap64(v) : v;
    }
};

// packed 2bit byte -> its four bases as chars, first base in the low byte
struct TwoBitQuad {
    uint32_t q[256];
    TwoBitQuad() {
        const char tcag[4] = {'T', 'C', 'A', 'G'};
        for (int b = 0; b < 256; b++) {
            uint32_t v = 0;
            for (int k = 0; k < 4; k++) v |= (uint32_t)(uint8_t)tcag[(b >> (6 - 2 * k)) & 3] << (8 * k);
            q[b] = v;
        }
    }
};
const TwoBitQuad TWOBIT_QUAD;

int read_2bit(const char *path, const std::vector<uint8_t> &file, Packer &pk) {
    const size_t n = file.size();
    const uint8_t *p = file.data();
    uint32_t m;
    memcpy(&m, p, 4);
    Cursor c{p, n, 4, m != 0x1A412743u};
    uint32_t version = c.u32(), count = c.u32();
    c.u32();
    if (c.bad || version > 1) return set_error(HSPZ_E_IO, "%s: unsupported 2bit version %u", path, version);
    std::vector<std::string> names(count);
    std::vector<uint64_t> offs(count);
    for (uint32_t i = 0; i < count; i++) {
        if (c.at >= n) return set_error(HSPZ_E_IO, "%s: 2bit index truncated", path);
        size_t len = p[c.at++];
        if (c.at + len > n) return set_error(HSPZ_E_IO, "%s: 2bit index truncated", path);
        names[i] = utf8_lossy(p + c.at, len);
        c.at += len;
        offs[i] = version ? c.u64() : c.u32();
        if (c.bad) return set_error(HSPZ_E_IO, "%s: 2bit index truncated", path);
    }
    const uint32_t *quad = TWOBIT_QUAD.q;
    std::vector<uint32_t> st, sz;
    for (uint32_t i = 0; i < count; i++) {
        Cursor r{p, n, (size_t)offs[i], c.swap};
        if (offs[i] > n) return set_error(HSPZ_E_IO, "%s: %s: record offset beyond the file", path, names[i].c_str());
        uint32_t dna = r.u32();
        auto blocks = [&](std::vector<uint32_t> &a, std::vector<uint32_t> &b) {
            uint32_t k = r.u32();
            if (r.bad || (uint64_t)k * 8 > n - std::min(n, r.at)) { r.bad = true; return; }
            a.resize(k);
            b.resize(k);
            for (uint32_t j = 0; j < k; j++) a[j] = r.u32();
            for (uint32_t j = 0; j < k; j++) b[j] = r.u32();
        };
        std::vector<uint32_t> ns, nl, ms, ml;
        blocks(ns, nl);
        blocks(ms, ml);
        r.u32();
        size_t bytes = ((size_t)dna + 3) / 4;
        if (r.bad || r.at + bytes > n) return set_error(HSPZ_E_IO, "%s: %s: 2bit record truncated", path, names[i].c_str());
        pk.begin(names[i]);
        uint8_t *o = pk.s->buf.room(bytes * 4);
        const uint8_t *dp = p + r.at;
        for (size_t j = 0; j < bytes; j++) memcpy(o + 4 * j, &quad[dp[j]], 4);
        pk.s->buf.commit(dna);
        for (size_t j = 0; j < ns.size(); j++) {
            uint64_t a = std::min<uint64_t>(ns[j], dna), b = std::min<uint64_t>((uint64_t)ns[j] + nl[j], dna);
            memset(o + a, 'N', b - a);
        }
        for (size_t j = 0; j < ms.size(); j++) {
            uint64_t a = std::min<uint64_t>(ms[j], dna), b = std::min<uint64_t>((uint64_t)ms[j] + ml[j], dna);
            for (uint64_t t = a; t < b; t++) o[t] |= 0x20;  // 'A'..'Z' -> 'a'..'z'
        }
        pk.end();
    }
    return 0;
}

int read_file(const char *path, std::vector<uint8_t> &out) {
    FILE *f = fopen(path, "rb");
    if (!f) return set_error(HSPZ_E_IO, "%s: cannot open", path);
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (sz < 0) {
        fclose(f);
        return set_error(HSPZ_E_IO, "%s: cannot size", path);
    }
    out.resize((size_t)sz);
    size_t got = sz ? fread(out.data(), 1, (size_t)sz, f) : 0;
    fclose(f);
    if (got != (size_t)sz) return set_error(HSPZ_E_IO, "%s: short read", path);
    return 0;
}

// The IUPAC complement of one byte (sequence.rs:80-86): A<->T, C<->G,
// B<->V, D<->H, K<->M, R<->Y, S, W to themselves, U->A, in both cases; other
// lowercase letters 'n', '&' itself, every other byte (>= 128 included) 'N'.
struct Complement {
    uint8_t t[256];
    Complement() {
        memset(t, 'N', sizeof t);
        for (int c = 'a'; c <= 'z'; c++) t[c] = 'n';
        t[SEP] = SEP;
        const char *pairs = "ATTACGGCBVVBDHHDKMMKRYYRSSWWUA";
        for (int i = 0; pairs[i]; i += 2) {
            t[(uint8_t)pairs[i]] = (uint8_t)pairs[i + 1];
            t[(uint8_t)pairs[i] + 32] = (uint8_t)pairs[i + 1] + 32;
        }
    }
};
const Complement COMPLEMENT;

// The scoring alphabet (sequence.rs:66-79).
struct Encoding {
    uint8_t t[256];
    Encoding() {
        memset(t, 6, sizeof t);
        t['A'] = 0, t['C'] = 1, t['G'] = 2, t['T'] = 3;
        t['a'] = t['c'] = t['g'] = t['t'] = 4;
        t['n'] = t['N'] = 5;
        t[SEP] = 7;
    }
};
const Encoding ENCODING;

const uint8_t *encoding_table() { return ENCODING.t; }
const uint8_t *complement_table() { return COMPLEMENT.t; }

template <class F>
void parallel_ranges(size_t n, int threads, F f) {
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    size_t t = threads > 0 ? (size_t)threads : hw;
    t = std::max<size_t>(1, std::min<size_t>(t, n / (1u << 20) + 1));
    if (t == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> pool;
    for (size_t i = 0; i < t; i++) pool.emplace_back([&, i] { f(n * i / t, n * (i + 1) / t); });
    for (auto &th : pool) th.join();
}

}  // namespace hspz

using namespace hspz;

static int seqset_read(const char *path, const char *prefix, hspz_seqset **out) {
    if (!path || !out) return set_error(HSPZ_E_ARG, "null argument");
    *out = nullptr;
    uint8_t head[4] = {0, 0, 0, 0};
    size_t nh = 0;
    {
        FILE *f = fopen(path, "rb");
        if (!f) return set_error(HSPZ_E_IO, "%s: cannot open", path);
        nh = fread(head, 1, 4, f);
        fclose(f);
    }
    std::unique_ptr<hspz_seqset> s(new hspz_seqset);
    Packer pk{s.get(), prefix ? prefix : ""};
    uint32_t le = (uint32_t)head[0] | (uint32_t)head[1] << 8 | (uint32_t)head[2] << 16 | (uint32_t)head[3] << 24;
    int rc;
    if (nh >= 4 && (le == 0x1A412743u || __builtin_bswap32(le) == 0x1A412743u)) {
        s->format = HSPZ_FORMAT_TWOBIT;
        std::vector<uint8_t> file;
        rc = read_file(path, file);
        if (!rc) rc = read_2bit(path, file, pk);
    } else if (nh >= 2 && head[0] == 0x1f && head[1] == 0x8b) {
        s->format = HSPZ_FORMAT_FASTA_GZ;
        rc = read_fasta_gz(path, pk);
    } else {
        s->format = HSPZ_FORMAT_FASTA;
        rc = read_fasta_plain(path, pk);
    }
    if (rc) return rc;
    FILE *f = fopen(path, "rb");
    if (f) {
        fseek(f, 0, SEEK_END);
        long sz = ftell(f);
        s->bytes_read = sz > 0 ? (uint64_t)sz : 0;
        fclose(f);
    }
    finish(s.get(), true);
    *out = s.release();
    return 0;
}

static int seqset_pack(size_t n, const char *const *names, const uint8_t *const *seqs, const size_t *lens,
                       const char *prefix, hspz_seqset **out) {
    if (!out || (n && (!names || !seqs || !lens))) return set_error(HSPZ_E_ARG, "null argument");
    std::unique_ptr<hspz_seqset> s(new hspz_seqset);
    Packer pk{s.get(), prefix ? prefix : ""};
    size_t total = 0;
    for (size_t i = 0; i < n; i++) total += lens[i] + 1;
    if (!s->buf.reserve(total)) return set_error(HSPZ_E_OOM, "no memory for %zu bytes", total);
    for (size_t i = 0; i < n; i++) {
        pk.begin(names[i] ? names[i] : "");
        if (lens[i]) {
            memcpy(s->buf.room(lens[i]), seqs[i], lens[i]);
            s->buf.commit(lens[i]);
        }
        pk.end();
    }
    finish(s.get(), true);
    *out = s.release();
    return 0;
}

static int seqset_rc(const hspz_seqset *s, hspz_seqset **out) {
    if (!s || !out) return set_error(HSPZ_E_ARG, "null argument");
    std::unique_ptr<hspz_seqset> r(new hspz_seqset);
    const size_t n = s->block_len;
    if (!r->buf.reserve(std::max<size_t>(n, 1))) return set_error(HSPZ_E_OOM, "no memory for %zu bytes", n);
    r->buf.commit(n);
    const uint8_t *src = s->buf.data();
    uint8_t *dst = r->buf.data();
    parallel_ranges(n, 0, [&](size_t a, size_t b) {
        for (size_t i = a; i < b; i++) dst[i] = COMPLEMENT.t[src[n - 1 - i]];
    });
    for (size_t i = s->names.size(); i-- > 0;) {
        r->names.push_back(s->names[i]);
        r->starts.push_back(n - s->starts[i] - s->lens[i]);
        r->lens.push_back(s->lens[i]);
    }
    r->block_len = n;
    r->format = s->format;
    r->bytes_read = s->bytes_read;
    finish(r.get(), false);
    *out = r.release();
    return 0;
}

// No exception crosses the ABI: an allocation failure becomes HSPZ_E_OOM.
template <class F>
static int guarded(F f) {
    clear_error();
    try {
        return f();
    } catch (const std::bad_alloc &) {
        return set_error(HSPZ_E_OOM, "host allocation failed");
    }
}

extern "C" int hspz_seqset_read(const char *path, const char *prefix, hspz_seqset **out) {
    return guarded([&] { return seqset_read(path, prefix, out); });
}

extern "C" int hspz_seqset_pack(size_t n, const char *const *names, const uint8_t *const *seqs, const size_t *lens,
                                const char *prefix, hspz_seqset **out) {
    return guarded([&] { return seqset_pack(n, names, seqs, lens, prefix, out); });
}

extern "C" int hspz_seqset_reverse_complement(const hspz_seqset *s, hspz_seqset **out) {
    return guarded([&] { return seqset_rc(s, out); });
}

extern "C" int hspz_seqset_get(const hspz_seqset *s, hspz_seqset_info *info) {
    clear_error();
    if (!s || !info) return set_error(HSPZ_E_ARG, "null argument");
    info->buf = s->buf.data();
    info->buf_len = s->buf.size();
    info->block_len = s->block_len;
    info->chrs = s->chrs.data();
    info->n_chrs = s->chrs.size();
    info->format = s->format;
    info->bytes_read = s->bytes_read;
    return 0;
}

extern "C" uint64_t hspz_seqset_digest(const hspz_seqset *s) { return s ? set_digest(s) : 0; }

extern "C" void hspz_seqset_free(hspz_seqset *s) { delete s; }

extern "C" int hspz_encode(const uint8_t *raw, size_t n, int threads, uint8_t *codes) {
    clear_error();
    if (n && (!raw || !codes)) return set_error(HSPZ_E_ARG, "null argument");
    parallel_ranges(n, threads, [&](size_t a, size_t b) {
        for (size_t i = a; i < b; i++) codes[i] = ENCODING.t[raw[i]];
    });
    return 0;
}

extern "C" int hspz_intervals(size_t block_len, uint32_t seed_size, uint32_t interval_size, uint32_t *pairs,
                              size_t cap, size_t *n_out) {
    clear_error();
    if (!n_out || (cap && !pairs)) return set_error(HSPZ_E_ARG, "null argument");
    if (!interval_size) return set_error(HSPZ_E_ARG, "interval_size must be positive");
    size_t k = 0;
    if (block_len > seed_size) {
        uint32_t end_pos = (uint32_t)(block_len - seed_size);  // `as u32`
        for (uint64_t cur = 0; cur < end_pos; cur += interval_size, k++)
            if (k < cap) {
                pairs[2 * k] = (uint32_t)cur;
                pairs[2 * k + 1] = (uint32_t)std::min<uint64_t>(end_pos, cur + interval_size);
            }
    }
    *n_out = k;
    return 0;
}

// synth.cpp — deterministic synthetic genomes for tests and benchmarks
// (include/hspz_synth.h). xoshiro256** seeded through SplitMix64; every draw
// order is fixed here, so a (seed, parameters) pair always yields the same
// bytes on every machine.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../include/hspz_synth.h"

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        for (int i = 0; i < 4; i++) {
            seed += 0x9E3779B97F4A7C15ull;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
    // number of trials to the first success, p in (0, 1]: {1, 2, ...}
    uint32_t geometric(double p) {
        uint32_t k = 1;
        while (uniform() >= p && k < 1000000) k++;
        return k;
    }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
};

inline uint8_t base_from(double u, double gc) {
    double at = (1.0 - gc) / 2.0, cg = gc / 2.0;
    if (u < at) return 'A';
    if (u < at + cg) return 'C';
    if (u < at + 2 * cg) return 'G';
    return 'T';
}

inline bool is_lower(uint8_t c) { return c >= 'a' && c <= 'z'; }
inline uint8_t upper(uint8_t c) { return is_lower(c) ? c - 32 : c; }

inline uint8_t transition(uint8_t c) {
    switch (c) {
    case 'A': return 'G';
    case 'G': return 'A';
    case 'C': return 'T';
    case 'T': return 'C';
    default: return c;
    }
}

inline uint8_t transversion(uint8_t c, int which) {
    switch (c) {
    case 'A': return which ? 'T' : 'C';
    case 'G': return which ? 'T' : 'C';
    case 'C': return which ? 'G' : 'A';
    case 'T': return which ? 'G' : 'A';
    default: return c;
    }
}

inline uint8_t complement(uint8_t c) {
    switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
    case 'a': return 't';
    case 'c': return 'g';
    case 'g': return 'c';
    case 't': return 'a';
    default: return c;
    }
}

}  // namespace

// Four bases per 64-bit draw: each 16-bit field is compared against the
// cumulative A/C/G thresholds scaled to 65536.
extern "C" void hspz_synth_random(uint64_t seed, size_t n, double gc, uint8_t *out) {
    Rng r(seed);
    uint32_t ta = (uint32_t)((1.0 - gc) / 2.0 * 65536.0);
    uint32_t tc = (uint32_t)((1.0 - gc) / 2.0 * 65536.0 + gc / 2.0 * 65536.0);
    uint32_t tg = (uint32_t)((1.0 - gc) / 2.0 * 65536.0 + gc * 65536.0);
    for (size_t i = 0; i < n; i += 4) {
        uint64_t x = r.next();
        for (int k = 0; k < 4 && i + k < n; k++) {
            uint32_t u = (uint32_t)(x >> (16 * k)) & 0xFFFF;
            out[i + k] = u < ta ? 'A' : (u < tc ? 'C' : (u < tg ? 'G' : 'T'));
        }
    }
}

// Per source base, in this order: with probability `indel` an indel event
// (insertion or deletion with equal odds, length ~ Geometric(1/3)): an
// insertion emits that many random bases before the source base, a deletion
// skips that many source bases. A base that is emitted is substituted with
// probability `sub`: a transition with probability 2/3, otherwise one of the
// two transversions with equal odds. Soft-mask case is kept.
extern "C" size_t hspz_synth_mutate(const uint8_t *src, size_t n, uint64_t seed, double sub,
                                    double indel, double gc, uint8_t *out, size_t cap) {
    Rng r(seed);
    size_t o = 0;
    for (size_t i = 0; i < n;) {
        if (r.uniform() < indel) {
            uint32_t len = r.geometric(1.0 / 3.0);
            if (r.uniform() < 0.5) {
                for (uint32_t k = 0; k < len && o < cap; k++) {
                    uint8_t b = base_from(r.uniform(), gc);
                    out[o++] = is_lower(src[i]) ? b + 32 : b;
                }
            } else {
                i += len;
                continue;
            }
        }
        uint8_t c = src[i++];
        if (r.uniform() < sub) {
            bool low = is_lower(c);
            uint8_t u = upper(c);
            double v = r.uniform();
            uint8_t m = v < 2.0 / 3.0 ? transition(u) : transversion(u, v < 5.0 / 6.0 ? 0 : 1);
            c = low && m >= 'A' && m <= 'Z' ? m + 32 : m;
        }
        if (o < cap) out[o++] = c;
    }
    return o;
}

extern "C" void hspz_synth_revcomp(uint8_t *seq, size_t n) {
    for (size_t i = 0, j = n ? n - 1 : 0; n && i < j; i++, j--) {
        uint8_t a = complement(seq[i]), b = complement(seq[j]);
        seq[i] = b;
        seq[j] = a;
    }
    if (n & 1) seq[n / 2] = complement(seq[n / 2]);
}

extern "C" void hspz_synth_revcomp_blocks(uint8_t *seq, size_t n, size_t block) {
    for (size_t b = 1; block && b * block < n; b += 2) {
        size_t lo = b * block, len = (lo + block <= n) ? block : n - lo;
        hspz_synth_revcomp(seq + lo, len);
    }
}

// Interspersed repeats: `families` random consensus sequences of
// `consensus_len` bases; copies, each mutated at its own divergence
// d ~ U(div_lo, div_hi) (substitutions d, indels d/15), overwrite the sequence
// at uniform positions until `fraction` of it is repeat-derived. Copies are
// written lowercase when `lowercase` is set (soft-masked).
extern "C" size_t hspz_synth_repeats(uint8_t *seq, size_t n, uint64_t seed, uint32_t families,
                                     uint32_t consensus_len, double div_lo, double div_hi,
                                     double fraction, int lowercase, double gc) {
    Rng r(seed);
    std::vector<std::vector<uint8_t>> fam(families, std::vector<uint8_t>(consensus_len));
    for (auto &f : fam)
        for (auto &b : f) b = base_from(r.uniform(), gc);
    std::vector<uint8_t> copy(consensus_len * 2 + 64);
    size_t covered = 0, target = (size_t)(fraction * (double)n);
    while (covered < target && n > 2 * (size_t)consensus_len) {
        const auto &f = fam[r.below(families)];
        double d = div_lo + (div_hi - div_lo) * r.uniform();
        size_t len = hspz_synth_mutate(f.data(), f.size(), r.next(), d, d / 15.0, gc, copy.data(),
                                       copy.size());
        size_t at = r.below(n - len);
        for (size_t k = 0; k < len; k++) seq[at + k] = lowercase ? copy[k] + 32 : copy[k];
        covered += len;
    }
    return covered;
}

// A lognormal length clipped to [lo, hi]: exp(N(ln median, sigma)).
extern "C" size_t hspz_synth_lognormal(uint64_t seed, size_t count, double median, double sigma,
                                       double lo, double hi, uint64_t *out) {
    Rng r(seed);
    size_t total = 0;
    for (size_t i = 0; i < count; i++) {
        double v = exp(log(median) + sigma * r.normal());
        v = v < lo ? lo : (v > hi ? hi : v);
        out[i] = (uint64_t)v;
        total += out[i];
    }
    return total;
}

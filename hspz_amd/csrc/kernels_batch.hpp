// kernels_batch.hpp — device code of the batched path (hspz_filter_batch).
//
// Every query chunk keeps the reference's per-block decomposition (intervals,
// plus then minus 250 kbp batches, run.rs:278-299) and every batch keeps its
// own dedup scope (gpu/mod.rs:2144-2151), but all batches of all chunks run in
// one device pass: seeding, hit counting, max_hits chunking, a fused
// hit-expansion + score-gate kernel that never materializes anchors, the exact
// materializer, and a scope-tagged HSP list. No host round trip until the end.
//
// Query layout: an arena of blocks (chunk 0 forward, chunk 0 reverse
// complement, chunk 1 forward, ...), each preceded by QPAD bytes of code 7, so
// the packed gate's windows never cross a block edge without touching an
// exception. Seeds carry arena positions; batches map back to blocks.
#pragma once

#include "kernels.hpp"

namespace hspz {

constexpr uint32_t QPAD = 256;       // code-7 bytes before every block and after the last
constexpr uint32_t MAX_CHUNKS = 64;  // max_hits chunks one batch may split into

struct BatchDesc {
    uint32_t block;  // arena block (2 * chunk + strand)
    uint32_t lo, hi; // seeded positions [lo, hi) of the block
};

struct SurvRec {
    uint32_t R, Q, seed, pad;
};

// upper_bound(a[0..n), v) - 1 for sorted a with a[0] <= v.
__device__ __forceinline__ uint32_t find_le(const uint32_t *__restrict__ a, uint32_t n, uint32_t v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1;
        else hi = mid;
    }
    return lo - 1;
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// Seedable windows of every batch position (seed_kmers validity,
// kernels.rs:273-313): flattened position p -> batch b -> block position j.
__global__ void k_bseed_valid(const uint8_t *__restrict__ arena, const uint32_t *__restrict__ blk_off,
                              const BatchDesc *__restrict__ batches, const uint32_t *__restrict__ poff,
                              uint32_t nb, uint32_t npos, uint32_t seed_size, uint8_t *__restrict__ valid) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npos; p += gridDim.x * blockDim.x) {
        uint32_t b = find_le(poff, nb, p);
        BatchDesc d = batches[b];
        const uint8_t *w = arena + blk_off[d.block] + d.lo + (p - poff[b]);
        uint8_t ok = 1;
        for (uint32_t j = 0; j < seed_size; j++) ok &= w[j] < 4;
        valid[p] = ok;
    }
}

// voff[b] = number of valid positions before batch b (lower_bound over the
// compacted, ascending position list); voff[nb] = total.
__global__ void k_bvoff(const uint32_t *__restrict__ ids, const uint32_t *__restrict__ nvalid,
                        const uint32_t *__restrict__ poff, uint32_t nb, uint32_t *__restrict__ voff) {
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nb) return;
    uint32_t n = *nvalid;
    uint32_t v = poff[b], lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (ids[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    voff[b] = b == nb ? n : lo;
}

// scatter_seeds (kernels.rs:319-355) for every batch at once; the seed's low
// word is the arena position of the window.
__global__ void k_bseed_scatter(const uint8_t *__restrict__ arena, const uint32_t *__restrict__ blk_off,
                                const BatchDesc *__restrict__ batches, const uint32_t *__restrict__ poff,
                                uint32_t nb, const uint32_t *__restrict__ ids,
                                const uint32_t *__restrict__ nvalid, ShapeArg shape, int transitions,
                                uint64_t *__restrict__ seeds) {
    const uint32_t n = *nvalid;
    const uint32_t per = transitions ? 1 + shape.kmer_size : 1;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        uint32_t p = ids[r];
        uint32_t b = find_le(poff, nb, p);
        BatchDesc d = batches[b];
        uint32_t apos = blk_off[d.block] + d.lo + (p - poff[b]);
        uint32_t kmer = 0;
        for (uint32_t t = 0; t < shape.kmer_size; t++) kmer = (kmer << 2) + arena[apos + shape.pos[t]];
        uint64_t *o = seeds + (uint64_t)r * per;
        o[0] = ((uint64_t)kmer << 32) | apos;
        if (transitions)
            for (uint32_t t = 0; t < shape.kmer_size; t++)
                o[1 + t] = ((uint64_t)(kmer ^ (2u << (2 * t))) << 32) | apos;
    }
}

// Bucket sizes of all seeds (count read from the device).
__global__ void k_bnum_hits(const uint64_t *__restrict__ seeds, const uint32_t *__restrict__ nvalid,
                            uint32_t per, const uint32_t *__restrict__ index, uint32_t *__restrict__ out) {
    const uint32_t n = *nvalid * per;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t k = (uint32_t)(seeds[i] >> 32);
        out[i] = index[k] - (k ? index[k - 1] : 0);
    }
}

// u64 inclusive scan of u32 counts: tile sums, one-block tile scan, apply.
__global__ __launch_bounds__(TILE_THREADS) void k_tile_sum_u64(const uint32_t *__restrict__ in,
                                                               const uint32_t *__restrict__ nvalid,
                                                               uint32_t per,
                                                               uint64_t *__restrict__ tile_sums) {
    const uint32_t n = *nvalid * per;
    uint32_t ntiles = (n + TILE - 1) / TILE;
    if (blockIdx.x >= ntiles) return;
    uint32_t base = blockIdx.x * TILE + threadIdx.x * TILE_ITEMS;
    uint64_t s = 0;
    for (int i = 0; i < TILE_ITEMS; i++)
        if (base + i < n) s += in[base + i];
    __shared__ uint64_t red[TILE_THREADS / WAVE];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, WAVE);
    if (lane_id() == 0) red[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < TILE_THREADS / WAVE; w++) t += red[w];
        tile_sums[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles_u64(uint64_t *__restrict__ sums,
                                                                 const uint32_t *__restrict__ nvalid,
                                                                 uint32_t per) {
    const uint32_t n = *nvalid * per;
    uint32_t ntiles = (n + TILE - 1) / TILE;
    uint32_t chunk = (ntiles + SCAN_THREADS - 1) / SCAN_THREADS;
    uint32_t lo = threadIdx.x * chunk, hi = min(lo + chunk, ntiles);
    uint64_t s = 0;
    for (uint32_t i = lo; i < hi; i++) s += sums[i];
    __shared__ uint64_t part[SCAN_THREADS];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        uint64_t v = threadIdx.x >= (uint32_t)off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t run = part[threadIdx.x] - s;
    for (uint32_t i = lo; i < hi; i++) {
        uint64_t v = sums[i];
        sums[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(TILE_THREADS) void k_tile_apply_u64(const uint32_t *__restrict__ in,
                                                                 const uint32_t *__restrict__ nvalid,
                                                                 uint32_t per,
                                                                 const uint64_t *__restrict__ offs,
                                                                 uint64_t *__restrict__ out) {
    const uint32_t n = *nvalid * per;
    uint32_t ntiles = (n + TILE - 1) / TILE;
    if (blockIdx.x >= ntiles) return;
    uint32_t base = blockIdx.x * TILE + threadIdx.x * TILE_ITEMS;
    uint32_t v[TILE_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < TILE_ITEMS; i++) {
        v[i] = base + i < n ? in[base + i] : 0;
        s += v[i];
    }
    uint32_t tot;
    uint64_t run = (uint64_t)block_excl_scan<TILE_THREADS>(s, &tot) + offs[blockIdx.x];
#pragma unroll
    for (int i = 0; i < TILE_ITEMS; i++) {
        run += v[i];
        if (base + i < n) out[base + i] = run;
    }
}

// Per batch: total hits and the max_hits chunk starts (gpu/mod.rs:2397-2420's
// lower_bound walk over the batch's cumulative counts). starts[b*MAX_CHUNKS+i]
// is the first seed of chunk i; err gets 1 on the reference's underflow and 2
// when a batch needs more than MAX_CHUNKS chunks.
__global__ void k_bchunks(const uint64_t *__restrict__ P, const uint32_t *__restrict__ voff, uint32_t nb,
                          uint32_t per, uint32_t max_hits, uint32_t *__restrict__ nchunks,
                          uint32_t *__restrict__ starts, uint64_t *__restrict__ bhits,
                          uint32_t *__restrict__ err) {
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    uint32_t s0 = voff[b] * per, s1 = voff[b + 1] * per;
    uint32_t *st = starts + (uint64_t)b * MAX_CHUNKS;
    st[0] = s0;
    if (s1 <= s0) {
        nchunks[b] = 0;
        bhits[b] = 0;
        return;
    }
    uint64_t base = s0 ? P[s0 - 1] : 0;
    uint64_t total = P[s1 - 1] - base;
    bhits[b] = total;
    if (total > 0xFFFFFFFFull) {
        atomicOr(err, 4u);
        nchunks[b] = 0;
        return;
    }
    if (total <= max_hits) {
        nchunks[b] = total ? 1 : 0;
        return;
    }
    uint32_t num_iter = (uint32_t)(total / max_hits) + 1;
    if (num_iter > MAX_CHUNKS) {
        atomicOr(err, 2u);
        nchunks[b] = 0;
        return;
    }
    uint64_t lim = max_hits;
    for (uint32_t it = 0; it + 1 < num_iter; it++) {
        uint32_t lo = s0, hi = s1;  // first s with P[s] - base >= lim
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (P[mid] - base < lim) lo = mid + 1;
            else hi = mid;
        }
        if (lo == s0) {
            atomicOr(err, 1u);
            nchunks[b] = 0;
            return;
        }
        uint32_t pos = lo - 1;  // inclusive last seed of chunk it
        st[it + 1] = pos + 1;
        lim = (P[pos] - base) + max_hits;
    }
    nchunks[b] = num_iter;
}

// Hit source over seeds (the batched path): a wave claims `claim`
// consecutive seeds at a time from a grid-wide counter (one atomic per claim,
// so waves that draw short buckets simply claim more often) and hands their
// hits (pos_table order) to idle lanes; hits are never written out. Survivors
// and deferred hits are appended as (R, Q, seed) records.
struct SeedSource {
    const uint64_t *seeds;
    const uint32_t *index, *pos;
    unsigned long long *work;  // seeds claimed so far; zero before the launch
    uint32_t seed_size, nseeds, claim;
    uint32_t s_cur, s_end, j_cur, cnt_cur, b0_cur, q_cur;
    bool dry;
    SurvRec *surv, *slow;
    uint32_t *surv_n, *slow_n, cap_surv, cap_slow;
    uint32_t *err;

    // Claims the wave's next seed range [s_cur, s_end); false once every
    // seed has been claimed.
    __device__ __forceinline__ bool take() {
        uint32_t leader = (uint32_t)(__ffsll((unsigned long long)__ballot(true)) - 1);
        unsigned long long v = 0;
        if (lane_id() == leader) v = atomicAdd(work, (unsigned long long)claim);
        uint64_t start = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) |
                         (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
        bool ok = start < (uint64_t)nseeds;
        uint64_t end = start + claim;
        s_cur = ok ? (uint32_t)start : nseeds;
        s_end = ok && end < (uint64_t)nseeds ? (uint32_t)end : nseeds;
        dry = !ok;
        return ok;
    }
    __device__ __forceinline__ void init() {
        dry = false;
        take();
        load_seed();
    }
    // Loads seed s_cur (or an empty stand-in past the end) without branching,
    // so every field stays a plain wave-uniform register.
    __device__ __forceinline__ void load_seed() {
        bool in = s_cur < s_end;
        uint64_t sd = seeds[in ? s_cur : 0];
        uint32_t k = in ? (uint32_t)(sd >> 32) : 0u;
        uint32_t b1 = index[k], b0 = k ? index[k - 1] : 0u;
        b0_cur = __builtin_amdgcn_readfirstlane(b0);
        cnt_cur = __builtin_amdgcn_readfirstlane(in ? b1 - b0 : 0u);
        q_cur = __builtin_amdgcn_readfirstlane((uint32_t)sd + seed_size);
        j_cur = 0;
    }
    // Next seed with hits; false when the grid's seeds are done.
    __device__ __forceinline__ bool advance() {
        for (;;) {
            if (dry) return false;
            s_cur++;
            if (s_cur >= s_end && !take()) return false;
            load_seed();
            if (cnt_cur) return true;
        }
    }
    __device__ __forceinline__ void refill(bool &has, bool &fresh, uint32_t &R, uint32_t &Q, uint32_t &tag) {
        uint64_t need = __ballot(!has);
        while (need) {
            if (j_cur >= cnt_cur) {
                if (!advance()) break;
                continue;
            }
            uint32_t avail = cnt_cur - j_cur;
            uint32_t rank = mbcnt(need);
            if (!has && rank < avail) {
                R = pos[b0_cur + j_cur + rank] + seed_size;
                Q = q_cur;
                tag = s_cur;
                has = true;
                fresh = true;
            }
            j_cur += min((uint32_t)__popcll(need), avail);
            need = __ballot(!has);
        }
    }
    __device__ __forceinline__ static void append(bool want, SurvRec *list, uint32_t *count, uint32_t cap,
                                                  uint32_t *err, uint32_t R, uint32_t Q, uint32_t tag) {
        uint64_t m = __ballot(want);
        if (!m) return;
        uint32_t leader = (uint32_t)(__ffsll((unsigned long long)m) - 1);
        uint32_t base = 0;
        if (lane_id() == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
        base = __shfl(base, leader, WAVE);
        uint32_t at = base + mbcnt(m);
        if (want) {
            if (at < cap) {
                *(uint4 *)(list + at) = make_uint4(R, Q, tag, 0u);
            } else {
                atomicOr(err, 8u);
            }
        }
    }
    __device__ __forceinline__ void emit(bool done, bool pass, uint32_t tag, uint32_t R, uint32_t Q) {
        append(done && pass, surv, surv_n, cap_surv, err, R, Q, tag);
    }
    __device__ __forceinline__ void defer(bool d, uint32_t tag, uint32_t R, uint32_t Q) {
        append(d, slow, slow_n, cap_slow, err, R, Q, tag);
    }
};

// The batched gate: fused hit expansion + packed gate over a persistent grid
// of BT-thread blocks, 16 * SW cells per step, `claim` seeds per wave claim.
template <int SW, int BT>
__global__ __launch_bounds__(BT) void k_bgate_packed(
    const uint32_t *__restrict__ rplane, const uint32_t *__restrict__ rcoarse, uint32_t ref_len,
    const uint32_t *__restrict__ qplane, const uint32_t *__restrict__ qcoarse, uint32_t qlen,
    const uint32_t *__restrict__ pairtab, int32_t xdrop, int32_t thresh,
    const uint64_t *__restrict__ seeds, const uint32_t *__restrict__ nvalid, uint32_t per,
    const uint32_t *__restrict__ index, const uint32_t *__restrict__ pos, uint32_t seed_size,
    uint32_t claim, unsigned long long *__restrict__ work,
    SurvRec *__restrict__ surv, uint32_t *__restrict__ surv_n, uint32_t cap_surv,
    SurvRec *__restrict__ slow, uint32_t *__restrict__ slow_n, uint32_t cap_slow,
    uint32_t *__restrict__ err) {
    __shared__ uint32_t s_tab[2 * 256 * G2_TAB_COPIES];
    for (uint32_t i = threadIdx.x; i < 2 * 256 * G2_TAB_COPIES; i += BT) s_tab[i] = pairtab[i / G2_TAB_COPIES];
    __syncthreads();
    const uint32_t *tab_fwd = s_tab + (lane_id() % G2_TAB_COPIES);
    SeedSource src;
    src.seeds = seeds;
    src.index = index;
    src.pos = pos;
    src.work = work;
    src.seed_size = seed_size;
    src.nseeds = __builtin_amdgcn_readfirstlane(*nvalid * per);
    src.claim = claim;
    src.surv = surv;
    src.surv_n = surv_n;
    src.cap_surv = cap_surv;
    src.slow = slow;
    src.slow_n = slow_n;
    src.cap_slow = cap_slow;
    src.err = err;
    src.init();
    packed_gate_loop<SW>(src, rplane, rcoarse, ref_len, qplane, qcoarse, qlen, tab_fwd,
                         tab_fwd + 256 * G2_TAB_COPIES, xdrop, thresh);
}

// Block bounds of a seed's batch: arena offset and length of its block.
struct BlockOf {
    const uint64_t *seeds;
    const uint32_t *voff, *blk_off, *blk_len;
    const BatchDesc *batches;
    uint32_t nb, per;
    __device__ __forceinline__ uint32_t batch_of(uint32_t seed) const {
        return find_le(voff, nb, seed / per);
    }
};

// Exact gate for deferred hits (one wave per record), chunk-local bounds.
__global__ __launch_bounds__(256) void k_bgate_wave(const uint8_t *__restrict__ ref, uint32_t ref_len,
                                                    const uint8_t *__restrict__ arena, BlockOf bo,
                                                    const int32_t *__restrict__ sub, int32_t xdrop,
                                                    int32_t thresh, const SurvRec *__restrict__ slow,
                                                    const uint32_t *__restrict__ slow_n, uint32_t cap_slow,
                                                    SurvRec *__restrict__ surv, uint32_t *__restrict__ surv_n,
                                                    uint32_t cap_surv, uint32_t *__restrict__ err) {
    __shared__ int32_t s_sub[64];
    if (threadIdx.x < 64) s_sub[threadIdx.x] = sub[threadIdx.x];
    __syncthreads();
    const uint32_t n = min(*slow_n, cap_slow);
    const uint32_t nwaves = gridDim.x * (blockDim.x / WAVE);
    for (uint32_t i = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; i < n; i += nwaves) {
        SurvRec rec = slow[i];
        uint32_t b = bo.batch_of(rec.seed);
        uint32_t blk = bo.batches[b].block;
        uint32_t q0 = bo.blk_off[blk], qlen = bo.blk_len[blk];
        int32_t mr, ml, pr, pl;
        wave_extend<false>(ref, ref_len, arena + q0, qlen, s_sub, xdrop, rec.R, rec.Q - q0, false, &mr, &pr);
        wave_extend<false>(ref, ref_len, arena + q0, qlen, s_sub, xdrop, rec.R, rec.Q - q0, true, &ml, &pl);
        bool pass = gate_keep((int32_t)((uint32_t)mr + (uint32_t)ml), thresh);
        if (pass && lane_id() == 0) {
            uint32_t at = atomicAdd(surv_n, 1u);
            if (at < cap_surv) surv[at] = rec;
            else atomicOr(err, 8u);
        }
    }
}

// find_hsps for the batched survivors (one wave per record): exact extension
// with positions in the chunk-local block, entropy, keep test. Output record
// i: HSP in block coordinates, its scope key (batch * MAX_CHUNKS + chunk) and
// the keep flag.
__global__ __launch_bounds__(256) void k_bmaterialize(
    const uint8_t *__restrict__ ref, uint32_t ref_len, const uint8_t *__restrict__ arena, BlockOf bo,
    const uint32_t *__restrict__ starts, const uint32_t *__restrict__ nchunks,
    const int32_t *__restrict__ sub, int32_t xdrop, int32_t thresh, int32_t noentropy, double log4,
    const SurvRec *__restrict__ surv, const uint32_t *__restrict__ surv_n, uint32_t cap_surv,
    SegmentPair *__restrict__ hsp, uint32_t *__restrict__ scope, uint8_t *__restrict__ keep) {
    __shared__ int32_t s_sub[64];
    if (threadIdx.x < 64) s_sub[threadIdx.x] = sub[threadIdx.x];
    __syncthreads();
    const uint32_t lane = lane_id();
    const uint32_t n = min(*surv_n, cap_surv);
    const uint32_t nwaves = gridDim.x * (blockDim.x / WAVE);
    for (uint32_t i = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; i < n; i += nwaves) {
        SurvRec rec = surv[i];
        uint32_t b = bo.batch_of(rec.seed);
        uint32_t blk = bo.batches[b].block;
        uint32_t q0 = bo.blk_off[blk], qlen = bo.blk_len[blk];
        const uint8_t *qry = arena + q0;
        uint32_t R = rec.R, Q = rec.Q - q0;
        int32_t Mr, rpos, Ml, lpos;
        wave_extend<true>(ref, ref_len, qry, qlen, s_sub, xdrop, R, Q, false, &Mr, &rpos);
        wave_extend<true>(ref, ref_len, qry, qlen, s_sub, xdrop, R, Q, true, &Ml, &lpos);
        int32_t total = (int32_t)((uint32_t)Mr + (uint32_t)Ml);
        int32_t extent = (int32_t)((uint32_t)rpos + (uint32_t)lpos);
        uint32_t leftx = (uint32_t)lpos;
        double entropy = 1.0;
        if (total >= thresh && (int64_t)total <= 3 * (int64_t)thresh && noentropy == 0) {
            uint32_t rs = R - leftx, qs = Q - leftx;
            int32_t positions = extent + 1;
            uint64_t packed = 0;
            for (int32_t p = (int32_t)lane; p < positions; p += WAVE) {
                uint32_t r = ref[rs + (uint32_t)p], q = qry[qs + (uint32_t)p];
                if (r == q && r < 4) packed += (uint64_t)1 << (16 * r);
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) packed += __shfl_xor(packed, off, WAVE);
            int32_t cnt[4] = {(int32_t)(packed & 0xFFFF), (int32_t)((packed >> 16) & 0xFFFF),
                              (int32_t)((packed >> 32) & 0xFFFF), (int32_t)((packed >> 48) & 0xFFFF)};
            if (cnt[0] + cnt[1] + cnt[2] + cnt[3] >= 20) {
                double denom = (double)(extent + 1);
                double acc = 0.0;
                for (int k = 0; k < 4; k++) {
                    double pb = (double)cnt[k] / denom;
                    double l = cnt[k] != 0 ? ref_ln(pb) : 0.0;
                    double t = pb * l;
                    acc = acc + t;
                }
                entropy = -acc / log4;
            }
        }
        if (lane == 0) {
            double gated = (double)(float)total * entropy;
            bool kept = trunc_sat(gated) >= thresh;
            if (kept) {
                SegmentPair o;
                o.ref_start = R - leftx;
                o.query_start = Q - leftx;
                o.len = (uint32_t)extent;
                o.score = entropy > 0.0 ? trunc_sat((double)total * entropy) : 0;
                hsp[i] = o;
                uint32_t nc = nchunks[b];
                uint32_t ch = nc > 1 ? find_le(starts + (uint64_t)b * MAX_CHUNKS, nc, rec.seed) : 0;
                scope[i] = b * MAX_CHUNKS + ch;
            }
            keep[i] = kept ? 1 : 0;
        }
    }
}

__global__ void k_gather_scoped(const SegmentPair *__restrict__ in, const uint32_t *__restrict__ sc_in,
                                const uint32_t *__restrict__ ids, const uint32_t *__restrict__ n_ptr,
                                SegmentPair *__restrict__ out, uint32_t *__restrict__ sc_out) {
    const uint32_t n = *n_ptr;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        out[i] = in[ids[i]];
        sc_out[i] = sc_in[ids[i]];
    }
}

}  // namespace hspz

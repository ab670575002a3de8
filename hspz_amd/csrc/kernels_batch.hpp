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
// Sum of one u64 per thread over the block (every thread gets it).
template <int THREADS>
__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v) {
    __shared__ uint64_t part[THREADS / WAVE];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, WAVE);
    if (lane_id() == 0) part[threadIdx.x / WAVE] = v;
    __syncthreads();
    uint64_t t = 0;
#pragma unroll
    for (int w = 0; w < THREADS / WAVE; w++) t += part[w];
    __syncthreads();
    return t;
}

// Exclusive block scan of one u64 per thread.
template <int THREADS>
__device__ __forceinline__ uint64_t block_excl_scan_u64(uint64_t v) {
    constexpr int NW = THREADS / WAVE;
    __shared__ uint64_t wsum[NW];
    const uint32_t lane = lane_id(), w = threadIdx.x / WAVE;
    uint64_t inc = v;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const uint64_t up = __shfl_up(inc, off, WAVE);
        if (lane >= (uint32_t)off) inc += up;
    }
    if (lane == WAVE - 1) wsum[w] = inc;
    __syncthreads();
    uint64_t base = 0;
#pragma unroll
    for (int k = 0; k < NW; k++) base += k < (int)w ? wsum[k] : 0;
    __syncthreads();
    return base + inc - v;
}

// Hit counts of the valid seed slots (bucket sizes of their k-mers) and the
// per-tile sums of the u64 scan: a block per tile of TILE slots, whose items
// lie at tile start + i * TILE_THREADS + thread, so every load and store is
// coalesced and each thread has all its seed and index loads in flight at
// once.
__global__ __launch_bounds__(TILE_THREADS) void k_bnum_tiles(const uint64_t *__restrict__ seeds,
                                                             const uint32_t *__restrict__ nvalid, uint32_t per,
                                                             const uint32_t *__restrict__ index,
                                                             uint32_t *__restrict__ out,
                                                             uint64_t *__restrict__ tile_sums) {
    const uint32_t n = *nvalid * per;
    if (blockIdx.x >= (n + TILE - 1) / TILE) return;
    const uint32_t base = blockIdx.x * TILE + threadIdx.x;
    uint32_t k[TILE_ITEMS];
#pragma unroll
    for (int i = 0; i < TILE_ITEMS; i++) {
        const uint32_t at = base + i * TILE_THREADS;
        k[i] = at < n ? (uint32_t)(seeds[at] >> 32) : 0u;
    }
    uint32_t hi[TILE_ITEMS], lo[TILE_ITEMS];
#pragma unroll
    for (int i = 0; i < TILE_ITEMS; i++) {
        hi[i] = index[k[i]];
        lo[i] = index[k[i] ? k[i] - 1 : 0];
    }
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < TILE_ITEMS; i++) {
        const uint32_t at = base + i * TILE_THREADS;
        const uint32_t c = hi[i] - (k[i] ? lo[i] : 0u);
        if (at < n) {
            out[at] = c;
            s += c;
        }
    }
    const uint64_t t = block_sum_u64<TILE_THREADS>(s);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = t;
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

// Phase 3 of the u64 scan: each tile's inclusive prefix of the counts, plus
// the tile's offset. Items are read and written coalesced (item i * THREADS +
// thread) and transposed through LDS, so each thread scans 16 consecutive
// items (one spare slot per 16 items spreads a wave's transposed reads over
// the banks).
__global__ __launch_bounds__(TILE_THREADS) void k_tile_apply_u64(const uint32_t *__restrict__ in,
                                                                 const uint32_t *__restrict__ nvalid,
                                                                 uint32_t per,
                                                                 const uint64_t *__restrict__ offs,
                                                                 uint64_t *__restrict__ out) {
    constexpr uint32_t SLOTS = TILE + TILE / TILE_ITEMS;
    __shared__ uint64_t s_v[SLOTS];
    const uint32_t n = *nvalid * per;
    if (blockIdx.x >= (n + TILE - 1) / TILE) return;
    const uint32_t t0 = blockIdx.x * TILE;
    auto slot = [](uint32_t x) { return x + x / TILE_ITEMS; };
#pragma unroll
    for (int i = 0; i < TILE_ITEMS; i++) {
        const uint32_t x = i * TILE_THREADS + threadIdx.x;
        s_v[slot(x)] = t0 + x < n ? in[t0 + x] : 0u;
    }
    __syncthreads();
    uint64_t v[TILE_ITEMS], run = 0;
#pragma unroll
    for (int j = 0; j < TILE_ITEMS; j++) {
        run += s_v[slot(threadIdx.x * TILE_ITEMS + j)];
        v[j] = run;
    }
    const uint64_t off = block_excl_scan_u64<TILE_THREADS>(run) + offs[blockIdx.x];
#pragma unroll
    for (int j = 0; j < TILE_ITEMS; j++) s_v[slot(threadIdx.x * TILE_ITEMS + j)] = v[j] + off;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < TILE_ITEMS; i++) {
        const uint32_t x = i * TILE_THREADS + threadIdx.x;
        if (t0 + x < n) out[t0 + x] = s_v[slot(x)];
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
    WaveQueue<uint4> sq, dq;  // survivors, deferred hits

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
        if (want && at < cap) *(uint4 *)(list + at) = make_uint4(R, Q, tag, 0u);
        if (lane_id() == leader && (uint64_t)base + (uint64_t)__popcll(m) > cap) atomicOr(err, 8u);
    }
    __device__ __forceinline__ void emit(bool done, bool pass, uint32_t tag, uint32_t R, uint32_t Q) {
        sq.push(done && pass, make_uint4(R, Q, tag, 0u), err);
    }
    __device__ __forceinline__ void defer(bool d, uint32_t tag, uint32_t R, uint32_t Q) {
        dq.push(d, make_uint4(R, Q, tag, 0u), err);
    }
    __device__ __forceinline__ void finish() {
        sq.flush(err);
        dq.flush(err);
    }
    // The lists and queues of a gate kernel: survivors appended directly (or
    // queued when `sq_slots` is given), deferred hits queued in `dq_slots`.
    __device__ __forceinline__ void set_lists(uint4 *sq_slots, uint4 *dq_slots) {
        sq = {sq_slots, 0u, (uint4 *)surv, surv_n, cap_surv};
        dq = {dq_slots, 0u, (uint4 *)slow, slow_n, cap_slow};
    }
};

// The batched gate: fused hit expansion + packed gate over a persistent grid
// of BT-thread blocks, 16 * SW cells per step, `claim` seeds per wave claim.
// Two blocks share a CU (the LDS table is 64 KiB); at 1024 threads that is 8
// waves per SIMD, which caps the kernel at 64 VGPRs.
template <int SW, int BT>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(2 * BT / 256, 8))) void k_bgate_packed(
    const uint32_t *__restrict__ rplane, const uint32_t *__restrict__ rcoarse, uint32_t ref_len,
    const uint32_t *__restrict__ qplane, const uint32_t *__restrict__ qcoarse, uint32_t qlen,
    const uint32_t *__restrict__ pairtab, int32_t xdrop, int32_t thresh,
    const uint64_t *__restrict__ seeds, const uint32_t *__restrict__ nvalid, uint32_t per,
    const uint32_t *__restrict__ index, const uint32_t *__restrict__ pos, uint32_t seed_size,
    uint32_t claim, unsigned long long *__restrict__ work,
    SurvRec *__restrict__ surv, uint32_t *__restrict__ surv_n, uint32_t cap_surv,
    SurvRec *__restrict__ slow, uint32_t *__restrict__ slow_n, uint32_t cap_slow,
    uint32_t *__restrict__ err) {
    __shared__ uint32_t s_map[G2_MAP_WORDS];
    __shared__ uint4 s_dq[BT / WAVE][WQ_CAP];
    load_map_table(s_map, pairtab, BT);
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
    src.set_lists(nullptr, s_dq[threadIdx.x / WAVE]);
    src.init();
    packed_gate_loop<SW>(src, rplane, rcoarse, ref_len, qplane, qcoarse, qlen, (const uint8_t *)s_map, xdrop,
                         thresh);
}

// ---------------------------------------------------------------- k-mer-major order
//
// The seed-major walk above gives every hit its own random reference window
// (one or two 128-byte lines per hit, past L2). Against a 1 Gbp target a
// multi-megabase batch holds several query seeds per k-mer, so the batched
// gate can instead group seeds by k-mer (a counting sort) and give each lane
// one reference position of the k-mer's bucket: the lane keeps that window in
// registers and walks every query seed of the group, so a reference window is
// fetched once per (k-mer, position) instead of once per hit.

// Swaps the two nibbles of every byte: applied to byte-swapped left words, it
// brings a byte's nearer cell pair into the low nibble, so the left stream
// forms its keys exactly like the right one.
__device__ __forceinline__ uint32_t nib_swap(uint32_t x) {
    return ((x >> 4) & 0x0F0F0F0Fu) | ((x << 4) & 0xF0F0F0F0u);
}

// The 2-bit Gray code of every cell of a packed word (code c -> c ^ (c >> 1):
// A0 C1 G3 T2). It is its own inverse and moves no cell, so it commutes with
// the word alignments and swaps above. The reference and query records hold
// their words in this code: the xor of two cells is then 00 for a match, 11
// for a transition and 01 or 10 for a transversion, so a word's class counts
// come from one unmasked popcount (matches 0, transversions 1, transitions 2
// per cell) and one masked one (cells that differ). The two-cell maps that
// read records index the Gray-keyed maps of the pair table (PAIRTAB_GRAY).
__device__ __forceinline__ uint32_t gray2(uint32_t w) { return w ^ ((w >> 1) & 0x55555555u); }

// Offset of the Gray-keyed maps in the engine's pair table (the plane-keyed
// ones come first).
constexpr uint32_t PAIRTAB_GRAY = 512;

// Per pos_table entry: 1 when the anchor R = pos + seed_size has the packed
// gate's margins and no exception code in its window (the reference-side
// test of packed_gate_loop, done once per engine).
__global__ void k_clean_bits(const uint32_t *__restrict__ pos, uint32_t n, const uint32_t *__restrict__ rcoarse,
                             uint32_t ref_len, uint32_t seed_size, uint32_t *__restrict__ bits) {
    const uint32_t lane = lane_id();
    const uint64_t nwaves = (uint64_t)gridDim.x * (blockDim.x / WAVE);
    const uint32_t nwords = (n + 31) / 32;
    for (uint64_t base = ((uint64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE) * WAVE; base < n;
         base += nwaves * WAVE) {
        uint64_t i = base + lane;
        bool ok = false;
        if (i < n) {
            uint32_t R = pos[i] + seed_size;
            ok = R >= G2_MARGIN && (uint64_t)R + G2_MARGIN <= ref_len &&
                 !coarse_hit(rcoarse, R - 16 * G2_LEFT_WORDS, R + 16 * G2_RIGHT_WORDS - 1);
        }
        uint64_t m = __ballot(ok);
        uint32_t w = (uint32_t)(base / 32);
        if (lane == 0) bits[w] = (uint32_t)m;
        if (lane == 1 && w + 1 < nwords) bits[w + 1] = (uint32_t)(m >> 32);
    }
}

// Reference records in pos_table order (one per indexed position, 48 bytes):
// the k-mer-major gate's right and left reference streams of the anchor
// R = pos + seed_size, aligned to R, left words byte- and nibble-swapped
// (zeros where the clean bit is 0), so that taking an item costs one
// coalesced load instead of a position load followed by a random window load.
__global__ void k_ref_records(const uint32_t *__restrict__ pos, uint32_t n, const uint32_t *__restrict__ clean,
                              const uint32_t *__restrict__ rplane, uint32_t seed_size, uint4 *__restrict__ out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t A[G2_RIGHT_WORDS + G2_LEFT_WORDS];
#pragma unroll
        for (int k = 0; k < (int)(G2_RIGHT_WORDS + G2_LEFT_WORDS); k++) A[k] = 0;
        if ((clean[i >> 5] >> (i & 31)) & 1u) {
            uint32_t R = pos[i] + seed_size;
            const uint32_t *rp = rplane + (R >> 4) - G2_LEFT_WORDS;
            uint32_t W[13];
#pragma unroll
            for (int k = 0; k < 13; k++) W[k] = rp[k];  // word 8 holds base R
            uint32_t sr = 2 * (R & 15);
#pragma unroll
            for (int k = 0; k < (int)G2_RIGHT_WORDS; k++) A[k] = gray2(__builtin_amdgcn_alignbit(W[9 + k], W[8 + k], sr));
#pragma unroll
            for (int k = 0; k < (int)G2_LEFT_WORDS; k++)
                A[G2_RIGHT_WORDS + k] =
                    gray2(nib_swap(__builtin_bswap32(__builtin_amdgcn_alignbit(W[8 - k], W[7 - k], sr))));
        }
        uint4 *o = out + 3 * (size_t)i;
        o[0] = make_uint4(A[0], A[1], A[2], A[3]);
        o[1] = make_uint4(A[4], A[5], A[6], A[7]);
        o[2] = make_uint4(A[8], A[9], A[10], A[11]);
    }
}

// Query record of one seeded query position (shared by the exact k-mer and
// every transition variant of that position): its right and left query
// streams as the k-mer-major gate consumes them (aligned to Q, left words
// byte- and nibble-swapped), the anchor Q, the seed slot of the position's
// exact k-mer (slot of variant t = tag0 + t) and whether the query window is
// clean.
struct QRec {
    uint32_t right[G2_RIGHT_WORDS];
    uint32_t left[G2_LEFT_WORDS];
    uint32_t Q, tag0, ok, pad;
};
static_assert(sizeof(QRec) == 64, "query records are four 16-byte loads");

// A grouped seed's sort value: its position rank r (the query record index)
// and its variant t (0 = exact k-mer), as (r << GV_SHIFT) | t.
constexpr uint32_t GV_SHIFT = 4;  // per = 1 + kmer_size <= 16 variants
constexpr uint32_t GV_MAX_POSITIONS = 1u << (32 - GV_SHIFT);

// Sort keys of the k-mer grouping: the n valid seed slots get their k-mer
// (a seed without hits lands in a group whose bucket is empty, which the gate
// skips), the unused slots [n, cap) get nk - 1: a stable sort puts them after
// every valid slot, so keys stay below nk (2k bits) and the grouped seeds are
// exactly the sorted slots [0, n). Slot i = r * per + t (position rank r,
// variant t) gets the value (r << GV_SHIFT) | t.
__global__ void k_kkeys(const uint64_t *__restrict__ seeds, const uint32_t *__restrict__ nvalid, uint32_t per,
                        uint32_t cap, uint32_t nk, uint32_t *__restrict__ key, uint32_t *__restrict__ val) {
    const uint32_t n = min(*nvalid * per, cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += gridDim.x * blockDim.x) {
        key[i] = i < n ? (uint32_t)(seeds[i] >> 32) : nk - 1;
        const uint32_t r = i / per;
        val[i] = (r << GV_SHIFT) | (i - r * per);
    }
}

// Key ranges longer than this are written by k_kgaps, a block per range,
// instead of by the slot that owns them.
constexpr uint32_t KGAP_SERIAL = 64;

// Capacity of the long-range queue for nk k-mers: the ranges are disjoint
// and each holds at least KGAP_SERIAL of the nk + 1 entries.
__host__ __device__ constexpr uint32_t kgap_capacity(uint32_t nk) { return (nk + 1) / KGAP_SERIAL + 2; }

// goff[a..b] = v: short ranges in place, long ones queued for k_kgaps.
__device__ __forceinline__ void kbounds_range(uint32_t a, uint32_t b, uint32_t v, uint32_t *__restrict__ goff,
                                              uint4 *__restrict__ gaps, uint32_t *__restrict__ ngaps,
                                              uint32_t cap_gaps) {
    if (a > b) return;
    if (b - a < KGAP_SERIAL) {
        for (uint32_t k = a; k <= b; k++) goff[k] = v;
        return;
    }
    const uint32_t at = atomicAdd(ngaps, 1u);
    if (at < cap_gaps) gaps[at] = make_uint4(a, b, v, 0u);
}

// Group starts from the sorted keys of the n valid slots: goff[k] = the first
// slot whose key is at least k, for k in [0, nk] (goff[nk] = n, the number of
// grouped seeds). Slot j owns the k-mers (key[j - 1], key[j]], and the last
// slot also those after it. A batch with few distinct k-mers leaves long
// ranges; those go to k_kgaps, so that no thread writes millions of entries.
__global__ void k_kbounds(const uint32_t *__restrict__ key, const uint32_t *__restrict__ nvalid, uint32_t per,
                          uint32_t cap, uint32_t nk, uint32_t *__restrict__ goff, uint4 *__restrict__ gaps,
                          uint32_t *__restrict__ ngaps, uint32_t cap_gaps) {
    const uint32_t n = min(*nvalid * per, cap);
    const uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
    if (n == 0) {
        for (uint32_t k = t0; k <= nk; k += stride) goff[k] = 0;
        return;
    }
    for (uint32_t j = t0; j < n; j += stride) {
        const uint32_t kj = key[j];
        kbounds_range(j ? key[j - 1] + 1 : 0, kj, j, goff, gaps, ngaps, cap_gaps);
        if (j + 1 == n) kbounds_range(kj + 1, nk, n, goff, gaps, ngaps, cap_gaps);
    }
}

// The long ranges k_kbounds queued: a block per range, a thread per entry.
__global__ void k_kgaps(const uint4 *__restrict__ gaps, const uint32_t *__restrict__ ngaps, uint32_t cap_gaps,
                        uint32_t *__restrict__ goff) {
    const uint32_t ng = min(*ngaps, cap_gaps);
    for (uint32_t g = blockIdx.x; g < ng; g += gridDim.x) {
        const uint4 r = gaps[g];
        for (uint32_t k = r.x + threadIdx.x; k <= r.y; k += blockDim.x) goff[k] = r.z;
    }
}

// One record per seeded query position (rank r of the valid positions), so
// that a hit's query side costs four loads instead of a window load, alignment
// and exception test; the position's k-mer variants share it. The arena
// position of rank r is the low word of its exact seed, seeds[r * per]. A
// block builds 256 consecutive records in LDS (80-byte rows, so neither the
// row writes nor the column reads meet more than two-way bank conflicts) and
// stores them as one contiguous 16 KiB run.
__global__ __launch_bounds__(256) void k_qrecords(const uint64_t *__restrict__ seeds,
                                                  const uint32_t *__restrict__ nvalid, uint32_t per,
                                                  const uint32_t *__restrict__ qplane,
                                                  const uint32_t *__restrict__ qcoarse, uint32_t qlen,
                                                  uint32_t seed_size, QRec *__restrict__ recs) {
    constexpr uint32_t ROW = 5;  // uint4 slots per staged record
    __shared__ uint4 s_rec[256 * ROW];
    const uint32_t total = *nvalid;
    for (uint32_t base = blockIdx.x * 256; base < total; base += gridDim.x * 256) {
        const uint32_t p = base + threadIdx.x;
        if (p < total) {
            uint32_t Q = (uint32_t)seeds[(uint64_t)p * per] + seed_size;
            QRec r;
            r.Q = Q;
            r.tag0 = p * per;
            r.pad = 0;
            bool ok = Q >= G2_MARGIN && (uint64_t)Q + G2_MARGIN <= qlen;
            uint32_t W[13];
#pragma unroll
            for (int k = 0; k < 13; k++) W[k] = 0;
            if (ok) {
                const uint32_t qlo = Q - 16 * G2_LEFT_WORDS, qhi = Q + 16 * G2_RIGHT_WORDS - 1;
                const uint32_t *qc = qcoarse + (qlo >> 11);
                ok = !coarse_bits(qc[0], qc[1], qlo, qhi);
                const uint32_t *qp = qplane + (Q >> 4) - G2_LEFT_WORDS;
#pragma unroll
                for (int k = 0; k < 13; k++) W[k] = qp[k];  // word 8 holds base Q
            }
            uint32_t sq2 = 2 * (Q & 15);
#pragma unroll
            for (int k = 0; k < (int)G2_RIGHT_WORDS; k++)
                r.right[k] = gray2(__builtin_amdgcn_alignbit(W[9 + k], W[8 + k], sq2));
#pragma unroll
            for (int k = 0; k < (int)G2_LEFT_WORDS; k++)
                r.left[k] = gray2(nib_swap(__builtin_bswap32(__builtin_amdgcn_alignbit(W[8 - k], W[7 - k], sq2))));
            r.ok = ok ? 1u : 0u;
            const uint4 *s = (const uint4 *)&r;
#pragma unroll
            for (int k = 0; k < 4; k++) s_rec[threadIdx.x * ROW + k] = s[k];
        }
        __syncthreads();
        const uint32_t n = min(256u, total - base);
        uint4 *o = (uint4 *)(recs + base);
        for (uint32_t j = threadIdx.x; j < 4 * n; j += 256) o[j] = s_rec[(j >> 2) * ROW + (j & 3)];
        __syncthreads();
    }
}

// Word g * SW + h of a stream (g = step within the phase), by selects.
template <int SW, int N>
__device__ __forceinline__ uint32_t pick_word(const uint32_t (&w)[N], uint32_t g, int h) {
    uint32_t v = w[h];
#pragma unroll
    for (int gg = 1; gg < N / SW; gg++) v = g == (uint32_t)gg ? w[gg * SW + h] : v;
    return v;
}

// K-mer-major hit source: waves claim `claim` consecutive grouped seeds at a
// time and own the groups that start inside the claim; a group's bucket
// entries go to idle lanes as items (one reference position each), and a
// lane walks its item's group of query records one hit at a time.
// The lists a k-mer-major gate appends to: survivors (or, for the coarse
// gate, candidates) through `sq`, deferred hits through `dq`.
struct GateOut {
    SurvRec *surv, *slow;
    uint32_t *surv_n, *slow_n, cap_surv, cap_slow;
    uint32_t *err;
    WaveQueue<uint4> sq, dq;  // survivors, deferred hits

    __device__ __forceinline__ void emit(bool done, bool pass, uint32_t tag, uint32_t R, uint32_t Q) {
        sq.push(done && pass, make_uint4(R, Q, tag, 0u), err);
    }
    __device__ __forceinline__ void defer(bool d, uint32_t tag, uint32_t R, uint32_t Q) {
        dq.push(d, make_uint4(R, Q, tag, 0u), err);
    }
    __device__ __forceinline__ void finish() {
        sq.flush(err);
        dq.flush(err);
    }
    __device__ __forceinline__ void set_lists(uint4 *sq_slots, uint4 *dq_slots) {
        sq = {sq_slots, 0u, (uint4 *)surv, surv_n, cap_surv};
        dq = {dq_slots, 0u, (uint4 *)slow, slow_n, cap_slow};
    }
};

struct KmerSource : GateOut {
    const uint32_t *skm, *sval, *goff, *index, *pos, *clean;
    unsigned long long *work;
    uint32_t total, claim, seed_size;
    uint32_t p_cur, p_end, g0, g1, b0, n, j;
    bool dry;

    __device__ __forceinline__ uint32_t rfl(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

    // The query record of grouped seed p and that seed's slot (its tag).
    __device__ __forceinline__ const uint4 *record(const QRec *recs, uint32_t p, uint32_t &var) const {
        const uint32_t gv = sval[p];
        var = gv & ((1u << GV_SHIFT) - 1);
        return (const uint4 *)(recs + (gv >> GV_SHIFT));
    }

    // Claims the next range of grouped seeds; p_cur = first group start in it.
    __device__ __forceinline__ bool take() {
        uint32_t leader = (uint32_t)(__ffsll((unsigned long long)__ballot(true)) - 1);
        unsigned long long v = 0;
        if (lane_id() == leader) v = atomicAdd(work, (unsigned long long)claim);
        uint64_t start = ((uint64_t)rfl((uint32_t)(v >> 32)) << 32) | (uint64_t)rfl((uint32_t)v);
        if (start >= (uint64_t)total) {
            dry = true;
            return false;
        }
        uint64_t end = start + claim;
        p_end = end < (uint64_t)total ? (uint32_t)end : total;
        uint32_t st = (uint32_t)start;
        uint32_t k = rfl(skm[st]), kp = rfl(skm[st ? st - 1 : 0]);
        if (st == 0) kp = ~0u;
        p_cur = k == kp ? rfl(goff[k + 1]) : st;
        return true;
    }
    __device__ __forceinline__ void init() {
        dry = false;
        p_cur = p_end = 0;
        n = j = 0;
    }
    // Next group of the claims; false when they are exhausted.
    __device__ __forceinline__ bool next_group() {
        for (;;) {
            if (dry) return false;
            if (p_cur >= p_end) {
                if (!take()) return false;
                continue;
            }
            uint32_t k = rfl(skm[p_cur]);
            g0 = p_cur;
            g1 = rfl(goff[k + 1]);
            p_cur = g1;
            uint32_t bb1 = index[k], bb0 = index[k ? k - 1 : 0];
            if (k == 0) bb0 = 0;
            b0 = rfl(bb0);
            n = rfl(bb1 - bb0);
            j = 0;
            if (n) return true;
        }
    }
    // A lane without a hit takes its item's next query record (index s_cur),
    // or else a new item.
    __device__ __forceinline__ void refill(bool &has, bool &item, bool &fresh_r, bool &fresh_q, uint32_t &R,
                                           uint32_t &s_cur, uint32_t &s_end, bool &rok, uint32_t &ridx) {
        if (!has && item) {
            if (s_cur < s_end) {
                has = true;
                fresh_q = true;
            } else {
                item = false;
            }
        }
        uint64_t need = __ballot(!has);
        while (need) {
            if (j >= n) {
                if (!next_group()) break;
                continue;
            }
            uint32_t avail = n - j;
            uint32_t rank = mbcnt(need);
            if (!has && rank < avail) {
                uint32_t e = b0 + j + rank;
                R = pos[e] + seed_size;
                rok = ((clean[e >> 5] >> (e & 31)) & 1u) != 0;
                ridx = e;
                s_cur = g0;
                s_end = g1;
                has = true;
                item = true;
                fresh_r = true;
                fresh_q = true;
            }
            j += min((uint32_t)__popcll(need), avail);
            need = __ballot(!has);
        }
    }
    // Hands each lane with `want` the next item of the current group(s): its
    // pos_table entry e and its group's query-record range [s0, s1). Returns
    // whether this lane got one (false for every lane once the work is drained).
    __device__ __forceinline__ bool claim_item(bool want, uint32_t &e, uint32_t &s0, uint32_t &s1) {
        bool got = false;
        uint64_t need = __ballot(want);
        while (need) {
            if (j >= n) {
                if (!next_group()) break;
                continue;
            }
            uint32_t avail = n - j;
            uint32_t rank = mbcnt(need);
            if (want && !got && rank < avail) {
                e = b0 + j + rank;
                s0 = g0;
                s1 = g1;
                got = true;
            }
            j += min((uint32_t)__popcll(need), avail);
            need = __ballot(want && !got);
        }
        return got;
    }
};

// Score-sum forms of the merged gates (the a halves of the map entries):
// SUM_TREE8: two trees of packed int16 adds over 4 entries each (8 cells,
// exact in int16 for xdrop and scores within the packed gate's limits), each
// folded into c; SUM_TREE16: one tree over all 8 (16 cells, exact in int16
// when xdrop + 1 and every ACGT score are at most G2_NARROW_SCORE + 1 and
// G2_NARROW_SCORE), folded once; SUM_RAW: c is the plain 32-bit sum of every
// entry of the side so far (three v_add3 per word), whose low 16 bits are the
// score sum mod 2^16 -- exact as an int16 while every ACGT score is within
// +-G2_RAW_SCORE, since a side spans at most 128 cells (side_sum reads it).
enum { SUM_TREE8 = 0, SUM_TREE16 = 1, SUM_RAW = 2 };
constexpr int32_t G2_RAW_SCORE = 255;

template <int SUMS = SUM_TREE8>
__device__ __forceinline__ void add_a8(int32_t &c, const uint32_t (&p)[8]) {
    if constexpr (SUMS == SUM_RAW) {
        const uint32_t x = p[0] + p[1] + p[2], y = p[3] + p[4] + p[5], z = p[6] + p[7] + (uint32_t)c;
        c = (int32_t)(x + y + z);
    } else if constexpr (SUMS == SUM_TREE16) {
        c += lo16(pk_add16(pk_add16(pk_add16(p[0], p[1]), pk_add16(p[2], p[3])),
                           pk_add16(pk_add16(p[4], p[5]), pk_add16(p[6], p[7]))));
    } else {
        c += lo16(pk_add16(pk_add16(p[0], p[1]), pk_add16(p[2], p[3])));
        c += lo16(pk_add16(pk_add16(p[4], p[5]), pk_add16(p[6], p[7])));
    }
}

// A side's score sum from its add_a8 accumulator.
template <int SUMS>
__device__ __forceinline__ int32_t side_sum(int32_t c) {
    if constexpr (SUMS == SUM_RAW) return (int32_t)(int16_t)(uint16_t)(uint32_t)c;
    return c;
}

// The packed gate over k-mer-major items: the same X-drop core and upper-bound
// argument as packed_gate_loop; reference words are prepared once per item
// (A0, LA0) and query words once per hit. At 64-cell steps no stream is
// shifted (each step picks its words by step index), so the reference window
// survives for the item's next hit; narrower steps shift working copies that
// every hit restores from A0, LA0. Left-stream words are byte- and
// nibble-swapped at preparation, so both phases form keys the same way.
template <int SW, class Src>
__device__ __forceinline__ void kmer_gate_loop(Src &src, const uint32_t *__restrict__ rplane,
                                               const QRec *__restrict__ recs, const uint8_t *s_map, int32_t xdrop,
                                               int32_t thresh) {
    static_assert(SW == 1 || SW == 2 || SW == 4, "step width must divide the right window");
    constexpr uint32_t SEL[4] = {0x0C0C0400u, 0x0C0C0500u, 0x0C0C0600u, 0x0C0C0700u};
    bool has = false, item = false, rok = false, left = false;
    uint32_t R = 0, Q = 0, tag = 0, s_cur = 0, s_end = 0, grp = 0, limit = 0;
    const uint32_t copy_off = (lane_id() & 31) * 4;
    uint32_t off = copy_off;
    int32_t c = 0, e = xdrop, Mr = 0;
    constexpr bool SHIFT = SW < 4;
    constexpr int WR = SHIFT ? (int)G2_RIGHT_WORDS : 1, WL = SHIFT ? (int)G2_LEFT_WORDS : 1;
    uint32_t A0[G2_RIGHT_WORDS], LA0[G2_LEFT_WORDS], B[G2_RIGHT_WORDS], LB[G2_LEFT_WORDS];
    uint32_t A[WR], LA[WL];  // working copies of the reference streams (narrow steps only)
#pragma unroll
    for (int j = 0; j < (int)G2_RIGHT_WORDS; j++) A0[j] = B[j] = 0;
#pragma unroll
    for (int j = 0; j < (int)G2_LEFT_WORDS; j++) LA0[j] = LB[j] = 0;
#pragma unroll
    for (int j = 0; j < WR; j++) A[j] = 0;
#pragma unroll
    for (int j = 0; j < WL; j++) LA[j] = 0;

    for (;;) {
        bool fresh_r = false, fresh_q = false;
        uint32_t ridx = 0;
        src.refill(has, item, fresh_r, fresh_q, R, s_cur, s_end, rok, ridx);
        if (__ballot(has) == 0) break;

        bool defer = false, done = false, pass = false;
        if (fresh_r && rok) {
            const uint32_t *rp = rplane + (R >> 4);
            uint32_t rend = rp[G2_RIGHT_WORDS], lend = rp[-(int)G2_LEFT_WORDS];
#pragma unroll
            for (int k = 0; k < (int)G2_RIGHT_WORDS; k++) A0[k] = rp[k];
#pragma unroll
            for (int k = 0; k < (int)G2_LEFT_WORDS; k++) LA0[k] = rp[-k];
            uint32_t sr = 2 * (R & 15);
#pragma unroll
            for (int k = 0; k < (int)G2_RIGHT_WORDS; k++)
                A0[k] = __builtin_amdgcn_alignbit(k + 1 < (int)G2_RIGHT_WORDS ? A0[k + 1] : rend, A0[k], sr);
#pragma unroll
            for (int k = 0; k < (int)G2_LEFT_WORDS; k++)
                LA0[k] = nib_swap(__builtin_bswap32(
                    __builtin_amdgcn_alignbit(LA0[k], k + 1 < (int)G2_LEFT_WORDS ? LA0[k + 1] : lend, sr)));
            // the query records are Gray-coded: so are the reference words they meet
#pragma unroll
            for (int k = 0; k < (int)G2_RIGHT_WORDS; k++) A0[k] = gray2(A0[k]);
#pragma unroll
            for (int k = 0; k < (int)G2_LEFT_WORDS; k++) LA0[k] = gray2(LA0[k]);
        }
        if (fresh_q) {
            uint32_t var;
            const uint4 *rec = src.record(recs, s_cur, var);
            s_cur++;
            uint4 v0 = rec[0], v1 = rec[1], v2 = rec[2], v3 = rec[3];
            B[0] = v0.x, B[1] = v0.y, B[2] = v0.z, B[3] = v0.w;
            LB[0] = v1.x, LB[1] = v1.y, LB[2] = v1.z, LB[3] = v1.w;
            LB[4] = v2.x, LB[5] = v2.y, LB[6] = v2.z, LB[7] = v2.w;
            Q = v3.x;
            tag = v3.y + var;
            bool ok = rok && v3.z != 0;
            if (!ok) {
                defer = true;
            } else {
                left = false;
                off = copy_off;
                grp = 0;
                limit = G2_RIGHT_WORDS / SW;
                c = 0;
                e = xdrop;
                if constexpr (SHIFT) {
#pragma unroll
                    for (int k = 0; k < WR; k++) A[k] = A0[k];
#pragma unroll
                    for (int k = 0; k < WL; k++) LA[k] = LA0[k];
                }
            }
        }

        if (has && !defer) {
            int32_t emin = e;
#pragma unroll
            for (int h = 0; h < SW; h++) {
                uint32_t r, q;
                if constexpr (SHIFT) {
                    r = left ? LA[h] : A[h];
                    q = left ? LB[h] : B[h];
                } else {
                    r = left ? pick_word<SW>(LA0, grp, h) : A0[h];
                    q = left ? pick_word<SW>(LB, grp, h) : B[h];
                }
                // key bytes in walk order for both phases (left words are
                // byte- and nibble-swapped): t1 byte k = cell pair 2k, t2 byte k = pair 2k+1
                uint32_t t1 = (r & 0x0F0F0F0Fu) | ((q << 4) & 0xF0F0F0F0u);
                uint32_t t2 = ((r >> 4) & 0x0F0F0F0Fu) | (q & 0xF0F0F0F0u);
                uint32_t p[8];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    p[2 * j] = map_entry(s_map, t1, off, SEL[j]);
                    p[2 * j + 1] = map_entry(s_map, t2, off, SEL[j]);
                }
#pragma unroll
                for (int i = 0; i < 8; i += 2) {
                    int32_t e1 = min(e + lo16(p[i]), hi16(p[i]));
                    e = min(e1 + lo16(p[i + 1]), hi16(p[i + 1]));
                    emin = min(emin, min(e1, e));
                }
                add_a8(c, p);
            }
            if (emin >= 0) {
                grp++;
                if constexpr (SHIFT) {
                    if (!left) {
#pragma unroll
                        for (int j = 0; j + SW < WR; j++) {
                            A[j] = A[j + SW];
                            B[j] = B[j + SW];
                        }
                    } else {
#pragma unroll
                        for (int j = 0; j + SW < WL; j++) {
                            LA[j] = LA[j + SW];
                            LB[j] = LB[j + SW];
                        }
                    }
                }
                if (grp == limit) defer = true;  // outlived the register window
            } else if (!left) {
                Mr = c + (xdrop - e);
                left = true;
                off = copy_off + 128;
                grp = 0;
                limit = G2_LEFT_WORDS / SW;
                c = 0;
                e = xdrop;
            } else {
                int32_t total = Mr + c + (xdrop - e);
                done = true;
                pass = packed_keep(total, thresh);
            }
        }
        src.emit(done, pass, tag, R, Q);
        src.defer(defer, tag, R, Q);
        if (done || defer) has = false;
    }
    src.finish();
}

// min(a, b, c) as one v_min3_i32 (the compiler otherwise builds min trees
// that take an extra instruction per pair).
__device__ __forceinline__ int32_t min3_i32(int32_t a, int32_t b, int32_t c) {
    int32_t d;
    asm volatile("v_min3_i32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

// The 8 cell pairs of one key word pair through their map entries p: steps e
// and folds e after every second pair into emin. Testing every second pair
// (every 4 cells) instead of every pair is a further relaxation of the upper
// bound: a drop whose e recovers within the next pair is seen at a later test
// or not at all (the hit then continues or is deferred), and the maximum
// c + xdrop - e of the stretch does not depend on the tests.
__device__ __forceinline__ void map_pairs8(const uint32_t (&p)[8], int32_t &e, int32_t &emin) {
    int32_t t[4];
#pragma unroll
    for (int i = 0; i < 8; i += 2) {
        int32_t e1 = min(e + lo16(p[i]), hi16(p[i]));
        e = min(e1 + lo16(p[i + 1]), hi16(p[i + 1]));
        t[i / 2] = e;
    }
    emin = min3_i32(emin, t[0], t[1]);
    emin = min3_i32(emin, t[2], t[3]);
}

// One 64-cell stretch (4 word pairs) of a stream through the map table at
// lane offset `off`: updates e, the score sum c, and the minimum of e over
// every second cell pair (map_pairs8).
// With N = 8 words, `hi` picks words 4..7 instead of 0..3 (per lane).
template <int N, int SUMS = SUM_TREE8>
__device__ __forceinline__ void map_stretch(const uint8_t *s_map, const uint32_t (&rw)[N], const uint32_t (&qw)[N],
                                            bool hi, uint32_t off, int32_t &e, int32_t &c, int32_t &emin) {
    constexpr uint32_t SEL[4] = {0x0C0C0400u, 0x0C0C0500u, 0x0C0C0600u, 0x0C0C0700u};
#pragma unroll
    for (int h = 0; h < 4; h++) {
        uint32_t r = rw[h], q = qw[h];
        if constexpr (N == 8) {
            r = hi ? rw[4 + h] : r;
            q = hi ? qw[4 + h] : q;
        }
        uint32_t t1 = (r & 0x0F0F0F0Fu) | ((q << 4) & 0xF0F0F0F0u);
        uint32_t t2 = ((r >> 4) & 0x0F0F0F0Fu) | (q & 0xF0F0F0F0u);
        uint32_t p[8];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            p[2 * j] = map_entry(s_map, t1, off, SEL[j]);
            p[2 * j + 1] = map_entry(s_map, t2, off, SEL[j]);
        }
        map_pairs8(p, e, emin);
        add_a8<SUMS>(c, p);
        // finish this word before the next one's reads: deferring the sums
        // keeps every table entry of the stretch live
        asm volatile("" : "+v"(c), "+v"(e), "+v"(emin));
    }
}

// K words (16 K cells) of a stream of N words through the map table, as
// map_stretch: words [0, K), or [K, 2K) when `hi` (N = 2K).
template <int K, int N, int SUMS = SUM_TREE8>
__device__ __forceinline__ void map_words(const uint8_t *s_map, const uint32_t (&rw)[N], const uint32_t (&qw)[N],
                                          bool hi, uint32_t off, int32_t &e, int32_t &c, int32_t &emin) {
    static_assert(N == K || N == 2 * K, "a stretch is the whole stream or one of its halves");
    constexpr uint32_t SEL[4] = {0x0C0C0400u, 0x0C0C0500u, 0x0C0C0600u, 0x0C0C0700u};
#pragma unroll
    for (int h = 0; h < K; h++) {
        uint32_t r = rw[h], q = qw[h];
        if constexpr (N == 2 * K) {
            r = hi ? rw[K + h] : r;
            q = hi ? qw[K + h] : q;
        }
        uint32_t t1 = (r & 0x0F0F0F0Fu) | ((q << 4) & 0xF0F0F0F0u);
        uint32_t t2 = ((r >> 4) & 0x0F0F0F0Fu) | (q & 0xF0F0F0F0u);
        uint32_t p[8];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            p[2 * j] = map_entry(s_map, t1, off, SEL[j]);
            p[2 * j + 1] = map_entry(s_map, t2, off, SEL[j]);
        }
        map_pairs8(p, e, emin);
        add_a8<SUMS>(c, p);
        asm volatile("" : "+v"(c), "+v"(e), "+v"(emin));
    }
}

// A hit's right stretch (RW words of the 4-word right stream, words [RW, 2 RW)
// when `rhi`) and left stretch (4 words of the 8-word left stream, the upper
// half when `lhi`) through the map table, as two map_words calls, with the
// two sides' words interleaved: each round issues the table reads of one
// right word and one left word before either side's recurrence consumes them,
// so one side's dependent adds and mins run while the other's reads are in
// flight.
template <int RW>
__device__ __forceinline__ void map_words_lr(const uint8_t *s_map, const uint32_t (&ra)[G2_RIGHT_WORDS],
                                             const uint32_t (&qa)[G2_RIGHT_WORDS], bool rhi, uint32_t off_r,
                                             int32_t &er, int32_t &cr, int32_t &eminr,
                                             const uint32_t (&rl)[G2_LEFT_WORDS], const uint32_t (&ql)[G2_LEFT_WORDS],
                                             bool lhi, uint32_t off_l, int32_t &el, int32_t &cl, int32_t &eminl) {
    static_assert(RW == 4 || RW == 2, "a right stretch is the whole right stream or one of its halves");
    constexpr uint32_t SEL[4] = {0x0C0C0400u, 0x0C0C0500u, 0x0C0C0600u, 0x0C0C0700u};
#pragma unroll
    for (int h = 0; h < 4; h++) {
        const bool right = h < RW;
        uint32_t pr[8], pl[8];
        if (right) {
            uint32_t r = ra[h], q = qa[h];
            if constexpr (RW == 2) {
                r = rhi ? ra[2 + h] : r;
                q = rhi ? qa[2 + h] : q;
            }
            uint32_t t1 = (r & 0x0F0F0F0Fu) | ((q << 4) & 0xF0F0F0F0u);
            uint32_t t2 = ((r >> 4) & 0x0F0F0F0Fu) | (q & 0xF0F0F0F0u);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                pr[2 * j] = map_entry(s_map, t1, off_r, SEL[j]);
                pr[2 * j + 1] = map_entry(s_map, t2, off_r, SEL[j]);
            }
        }
        {
            uint32_t r = lhi ? rl[4 + h] : rl[h], q = lhi ? ql[4 + h] : ql[h];
            uint32_t t1 = (r & 0x0F0F0F0Fu) | ((q << 4) & 0xF0F0F0F0u);
            uint32_t t2 = ((r >> 4) & 0x0F0F0F0Fu) | (q & 0xF0F0F0F0u);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                pl[2 * j] = map_entry(s_map, t1, off_l, SEL[j]);
                pl[2 * j + 1] = map_entry(s_map, t2, off_l, SEL[j]);
            }
        }
        if (right) {
            map_pairs8(pr, er, eminr);
            add_a8(cr, pr);
        }
        map_pairs8(pl, el, eminl);
        add_a8(cl, pl);
        // finish this round before the next one's reads, so at most two
        // words' table entries are live
        asm volatile("" : "+v"(cr), "+v"(er), "+v"(eminr), "+v"(cl), "+v"(el), "+v"(eminl));
    }
}

// Per-lane state of a hit's two sides across the merged gates' iterations:
// whether each side has dropped (its maximum recorded) or continues from a
// kept X-drop state and score sum.
struct HitSides {
    bool rdone = false, ldone = false, rcont = false, cont = false;
    int32_t Mr = 0, Ml = 0, er_keep = 0, cr_keep = 0, el_keep = 0, cl_keep = 0;

    __device__ __forceinline__ void start() { rdone = ldone = rcont = cont = false; }
    // The decisions after an iteration's stretches, as selects (nested
    // branches around them cost more issue than the selects): a side that
    // drops records its maximum; one that outlives its last stretch (right:
    // 64 cells, left: 128) defers the hit; one that outlives an earlier
    // stretch keeps its state for the next. `done` once both sides dropped,
    // `pass` if their maxima reach the threshold.
    template <int RW>
    __device__ __forceinline__ void decide(int32_t er, int32_t cr, int32_t eminr, int32_t el, int32_t cl,
                                           int32_t eminl, int32_t xdrop, int32_t thresh, bool &defer, bool &done,
                                           bool &pass) {
        const bool r_drop = !rdone && eminr < 0, l_drop = !ldone && eminl < 0;
        const bool r_open = !rdone && !r_drop, l_open = !ldone && !l_drop;
        const bool r_last = rcont || RW == (int)G2_RIGHT_WORDS;
        Mr = r_drop ? cr + (xdrop - er) : Mr;
        Ml = l_drop ? cl + (xdrop - el) : Ml;
        defer = (r_open && r_last) || (l_open && cont);
        const bool r_next = r_open && !r_last, l_next = l_open && !cont;
        er_keep = r_next ? er : er_keep;
        cr_keep = r_next ? cr : cr_keep;
        el_keep = l_next ? el : el_keep;
        cl_keep = l_next ? cl : cl_keep;
        rdone = rdone || r_drop;
        ldone = ldone || l_drop;
        rcont = rcont || r_next;
        cont = cont || l_next;
        done = !defer && rdone && ldone;
        pass = done && packed_keep(Mr + Ml, thresh);
    }
};

// The k-mer-major gate with merged phases: one iteration evaluates a fresh
// hit's first right stretch and its first 64 left cells; a hit whose left
// extension outlives them takes another iteration for left cells 64..127.
// RW = 4: the right stretch is cells 0..63, so the iteration evaluates
// exactly the cells of kmer_gate_loop<4> and reaches the same decisions,
// with about half the iterations per hit (the refill and bookkeeping of an
// iteration are paid once per hit) and both table offsets fixed per stretch.
// RW = 2: right cells 0..31, and a right side that outlives them continues
// with cells 32..63 in the next iteration (a random right extension mostly
// drops within 32 cells, so most hits skip 32 cells of work), with the same
// upper-bound decisions.
template <int RW, int SUMS, class Src>
__device__ __forceinline__ void kmer_gate_merged(Src &src, const uint32_t *__restrict__ rplane,
                                                 const uint4 *__restrict__ rrec, const QRec *__restrict__ recs,
                                                 const uint8_t *s_map, int32_t xdrop, int32_t thresh) {
    static_assert(RW == 4 || RW == 2, "a right stretch is the whole right stream or one of its halves");
    bool has = false, item = false, rok = false;
    HitSides st;
    uint32_t R = 0, Q = 0, tag = 0, s_cur = 0, s_end = 0;
    const uint32_t off_fwd = (lane_id() & 31) * 4, off_rev = off_fwd + 128;
    uint32_t A0[G2_RIGHT_WORDS], LA0[G2_LEFT_WORDS], B[G2_RIGHT_WORDS], LB[G2_LEFT_WORDS];
#pragma unroll
    for (int j = 0; j < (int)G2_RIGHT_WORDS; j++) A0[j] = B[j] = 0;
#pragma unroll
    for (int j = 0; j < (int)G2_LEFT_WORDS; j++) LA0[j] = LB[j] = 0;

    for (;;) {
        bool fresh_r = false, fresh_q = false;
        uint32_t ridx = 0;
        src.refill(has, item, fresh_r, fresh_q, R, s_cur, s_end, rok, ridx);
        if (__ballot(has) == 0) break;

        bool defer = false, done = false, pass = false;
        if (rrec && fresh_r && rok) {  // the item's precomputed reference record
            const uint4 *rr = rrec + 3 * (size_t)ridx;
            uint4 w0 = rr[0], w1 = rr[1], w2 = rr[2];
            A0[0] = w0.x, A0[1] = w0.y, A0[2] = w0.z, A0[3] = w0.w;
            LA0[0] = w1.x, LA0[1] = w1.y, LA0[2] = w1.z, LA0[3] = w1.w;
            LA0[4] = w2.x, LA0[5] = w2.y, LA0[6] = w2.z, LA0[7] = w2.w;
        } else if (fresh_r && rok) {
            const uint32_t *rp = rplane + (R >> 4);
            uint32_t rend = rp[G2_RIGHT_WORDS], lend = rp[-(int)G2_LEFT_WORDS];
#pragma unroll
            for (int k = 0; k < (int)G2_RIGHT_WORDS; k++) A0[k] = rp[k];
#pragma unroll
            for (int k = 0; k < (int)G2_LEFT_WORDS; k++) LA0[k] = rp[-k];
            uint32_t sr = 2 * (R & 15);
#pragma unroll
            for (int k = 0; k < (int)G2_RIGHT_WORDS; k++)
                A0[k] = __builtin_amdgcn_alignbit(k + 1 < (int)G2_RIGHT_WORDS ? A0[k + 1] : rend, A0[k], sr);
#pragma unroll
            for (int k = 0; k < (int)G2_LEFT_WORDS; k++)
                LA0[k] = nib_swap(__builtin_bswap32(
                    __builtin_amdgcn_alignbit(LA0[k], k + 1 < (int)G2_LEFT_WORDS ? LA0[k + 1] : lend, sr)));
            // the query records are Gray-coded: so are the reference words they meet
#pragma unroll
            for (int k = 0; k < (int)G2_RIGHT_WORDS; k++) A0[k] = gray2(A0[k]);
#pragma unroll
            for (int k = 0; k < (int)G2_LEFT_WORDS; k++) LA0[k] = gray2(LA0[k]);
        }
        if (fresh_q) {
            uint32_t var;
            const uint4 *rec = src.record(recs, s_cur, var);
            s_cur++;
            uint4 v0 = rec[0], v1 = rec[1], v2 = rec[2], v3 = rec[3];
            B[0] = v0.x, B[1] = v0.y, B[2] = v0.z, B[3] = v0.w;
            LB[0] = v1.x, LB[1] = v1.y, LB[2] = v1.z, LB[3] = v1.w;
            LB[4] = v2.x, LB[5] = v2.y, LB[6] = v2.z, LB[7] = v2.w;
            Q = v3.x;
            tag = v3.y + var;
            st.start();
            if (!(rok && v3.z != 0)) defer = true;
        }

        if (has && !defer) {
            // the right stretch: cells 0 .. 16 RW - 1, or the next ones for a continuing right side
            int32_t er = st.rcont ? st.er_keep : xdrop, cr = st.rcont ? st.cr_keep : 0, eminr = er;
            map_words<RW, G2_RIGHT_WORDS, SUMS>(s_map, A0, B, st.rcont, off_fwd, er, cr, eminr);
            // left cells 0..63, or 64..127 for a continuing left side
            int32_t el = st.cont ? st.el_keep : xdrop, cl = st.cont ? st.cl_keep : 0, eminl = el;
            map_stretch<8, SUMS>(s_map, LA0, LB, st.cont, off_rev, el, cl, eminl);
            st.decide<RW>(er, side_sum<SUMS>(cr), eminr, el, side_sum<SUMS>(cl), eminl, xdrop, thresh, defer, done,
                          pass);
        }
        src.emit(done, pass, tag, R, Q);
        src.defer(defer, tag, R, Q);
        if (done || defer) {
            has = false;
            st.cont = st.rcont = false;
        }
    }
    src.finish();
}

// A 16-byte load marked non-temporal (data read once, streamed past the caches).
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 nt_load4(const uint4 *p) {
    u32x4_t v = __builtin_nontemporal_load((const u32x4_t *)p);
    return make_uint4(v.x, v.y, v.z, v.w);
}

// kmer_gate_merged with every global load issued one iteration before its
// data is used, so a wave computes one hit while the loads for its next hit
// and next item are in flight. Each lane holds, besides its current item and
// hit: the next item (its reference record, R, clean bit and query-record
// range), the item claimed in the last iteration (loads in flight), and the
// query record of its next fresh hit (prefetched). The load instructions of
// an iteration are issued unconditionally (idle lanes read entry 0), so the
// counter wait at the top of the next iteration covers exactly them. Results
// of a finished hit are appended at the top of the next iteration. Hits,
// their cells and decisions are those of kmer_gate_merged; only the order in
// which lanes take items differs. Reads the reference records (k_ref_records).
// RW = 4: right cells 0..63 in a hit's first iteration (the cells of
// kmer_gate_merged). RW = 2: right cells 0..31, and a right side that outlives
// them continues with cells 32..63 in the next iteration, while the left side
// keeps its 64-cell stretches; fewer cells per hit (a right side mostly drops
// within 32), the same upper-bound decisions.
template <int RW>
__device__ __forceinline__ void kmer_gate_ahead(KmerSource &src, const uint4 *__restrict__ rrec,
                                                const QRec *__restrict__ recs, const uint8_t *s_map, int32_t xdrop,
                                                int32_t thresh) {
    const uint32_t off_fwd = (lane_id() & 31) * 4, off_rev = off_fwd + 128;
    // current item and hit (per side: continuing, done, kept state, maximum)
    bool cur = false, rok = false, has = false, qok = false;
    HitSides st;
    uint32_t R = 0, s_cur = 0, s_end = 0, Q = 0, tag = 0;
    uint32_t A0[G2_RIGHT_WORDS], LA0[G2_LEFT_WORDS], B[G2_RIGHT_WORDS], LB[G2_LEFT_WORDS];
    // next item, and the item whose loads are in flight
    bool nxt = false, nrok = false, fly = false;
    uint32_t nR = 0, ns0 = 0, ns1 = 0, fsh = 0, fs0 = 0, fs1 = 0;
    uint32_t NA[G2_RIGHT_WORDS + G2_LEFT_WORDS];
    uint4 T0 = make_uint4(0, 0, 0, 0), T1 = T0, T2 = T0;
    uint32_t tpos = 0, tclean = 0;
    // prefetched query record of the next fresh hit
    bool pv = false;
    uint32_t pidx = 0, pvar = 0;
    uint4 P0 = T0, P1 = T0, P2 = T0, P3 = T0;
    // the hit finished in the previous iteration, appended at the top of this one
    bool e_pass = false, e_defer = false;
    uint32_t eR = 0, eQ = 0, etag = 0;
#pragma unroll
    for (int j = 0; j < (int)G2_RIGHT_WORDS; j++) A0[j] = B[j] = 0;
#pragma unroll
    for (int j = 0; j < (int)G2_LEFT_WORDS; j++) LA0[j] = LB[j] = 0;
#pragma unroll
    for (int j = 0; j < (int)(G2_RIGHT_WORDS + G2_LEFT_WORDS); j++) NA[j] = 0;

    for (;;) {
        // The record's pad word is read nowhere; keeping it live to here stops
        // the compiler from reusing its register while the load is in flight
        // (which would stall the hit's cells on that load).
        asm volatile("" ::"v"(P3.w));
        // 1. results of the previous iteration's loads and hit
        if (fly) {
            NA[0] = T0.x, NA[1] = T0.y, NA[2] = T0.z, NA[3] = T0.w;
            NA[4] = T1.x, NA[5] = T1.y, NA[6] = T1.z, NA[7] = T1.w;
            NA[8] = T2.x, NA[9] = T2.y, NA[10] = T2.z, NA[11] = T2.w;
            nR = tpos + src.seed_size;
            nrok = ((tclean >> fsh) & 1u) != 0;
            ns0 = fs0;
            ns1 = fs1;
            nxt = true;
            fly = false;
        }
        src.emit(true, e_pass, etag, eR, eQ);
        src.defer(e_defer, etag, eR, eQ);
        e_pass = e_defer = false;

        // 2. a lane without a hit starts its item's next one, promoting the
        //    next item when the current one is used up
        bool defer = false;
        if (!has) {
            if (cur && s_cur >= s_end) cur = false;
            if (!cur && nxt) {
#pragma unroll
                for (int k = 0; k < (int)G2_RIGHT_WORDS; k++) A0[k] = NA[k];
#pragma unroll
                for (int k = 0; k < (int)G2_LEFT_WORDS; k++) LA0[k] = NA[G2_RIGHT_WORDS + k];
                R = nR;
                rok = nrok;
                s_cur = ns0;
                s_end = ns1;
                cur = true;
                nxt = false;
            }
            if (cur) {
                uint4 v0 = P0, v1 = P1, v2 = P2, v3 = P3;
                uint32_t var = pvar;
                if (!(pv && pidx == s_cur)) {  // not prefetched (kept for safety; the schedule prefetches it)
                    const uint4 *rec = src.record(recs, s_cur, var);
                    v0 = rec[0], v1 = rec[1], v2 = rec[2], v3 = rec[3];
                }
                B[0] = v0.x, B[1] = v0.y, B[2] = v0.z, B[3] = v0.w;
                LB[0] = v1.x, LB[1] = v1.y, LB[2] = v1.z, LB[3] = v1.w;
                LB[4] = v2.x, LB[5] = v2.y, LB[6] = v2.z, LB[7] = v2.w;
                Q = v3.x;
                tag = v3.y + var;
                qok = v3.z != 0;
                s_cur++;
                has = true;
                st.start();
                if (!(rok && qok)) defer = true;
            }
        }

        // 3. lanes without a next item claim one
        uint32_t ce = 0, c0 = 0, c1 = 0;
        bool got = src.claim_item(!nxt && !fly, ce, c0, c1);
        bool pending = cur && s_cur < s_end;
        if (__ballot(has || pending || nxt || got || e_pass || e_defer) == 0) break;

        // 4. this iteration's loads, for every lane: the claimed item's record,
        //    pos entry and clean word; the query record of the next fresh hit
        {
            uint32_t e = got ? ce : 0u;
            const uint4 *rr = rrec + 3 * (size_t)e;
            // reference records and pos entries are read once per item: streamed past the caches
            T0 = nt_load4(rr + 0);
            T1 = nt_load4(rr + 1);
            T2 = nt_load4(rr + 2);
            tpos = __builtin_nontemporal_load(src.pos + e);
            tclean = src.clean[e >> 5];
            fsh = e & 31;
            fs0 = c0;
            fs1 = c1;
            fly = got;
            uint32_t idx = pending ? s_cur : nxt ? ns0 : c0;
            pv = pending || nxt || got;
            pidx = idx;
            const uint4 *rec = src.record(recs, pv ? idx : 0u, pvar);
            P0 = rec[0], P1 = rec[1], P2 = rec[2], P3 = rec[3];
        }

        // 5. the current hit: its next right stretch (RW words) and its next
        //    left stretch (cells 0..63, or 64..127 for a continuing left side)
        bool done = false, pass = false;
        if (has && !defer) {
            int32_t er = st.rcont ? st.er_keep : xdrop, cr = st.rcont ? st.cr_keep : 0, eminr = er;
            int32_t el = st.cont ? st.el_keep : xdrop, cl = st.cont ? st.cl_keep : 0, eminl = el;
            map_words_lr<RW>(s_map, A0, B, st.rcont, off_fwd, er, cr, eminr, LA0, LB, st.cont, off_rev, el, cl,
                             eminl);
            st.decide<RW>(er, cr, eminr, el, cl, eminl, xdrop, thresh, defer, done, pass);
        }
        if (has && (done || defer)) {
            e_pass = done && pass;
            e_defer = defer;
            eR = R;
            eQ = Q;
            etag = tag;
            has = false;
        }
    }
    src.finish();
}

// The lookahead k-mer-major gate (kmer_gate_ahead<RW>); same arguments as
// k_bgate_kmer, rrec required. Four waves per SIMD (512-thread blocks, two per
// CU) leave it the 128 VGPRs its in-flight state needs.
template <int RW>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_bgate_ahead(
    const uint32_t *__restrict__ rplane, const uint32_t *__restrict__ pairtab, int32_t xdrop,
    int32_t thresh, const QRec *__restrict__ recs, const uint32_t *__restrict__ skm,
    const uint32_t *__restrict__ sval, const uint32_t *__restrict__ goff, uint32_t nk,
    const uint32_t *__restrict__ index, const uint32_t *__restrict__ pos, const uint32_t *__restrict__ clean,
    const uint4 *__restrict__ rrec, uint32_t seed_size, uint32_t claim, unsigned long long *__restrict__ work,
    SurvRec *__restrict__ surv,
    uint32_t *__restrict__ surv_n, uint32_t cap_surv, SurvRec *__restrict__ slow, uint32_t *__restrict__ slow_n,
    uint32_t cap_slow, uint32_t *__restrict__ err) {
    __shared__ uint32_t s_map[G2_MAP_WORDS];
    __shared__ uint4 s_q[2][512 / WAVE][WQ_CAP];
    load_map_table(s_map, pairtab, 512);
    KmerSource src;
    src.skm = skm;
    src.sval = sval;
    src.goff = goff;
    src.index = index;
    src.pos = pos;
    src.clean = clean;
    src.work = work;
    src.total = __builtin_amdgcn_readfirstlane(goff[nk]);
    src.claim = claim;
    src.seed_size = seed_size;
    src.surv = surv;
    src.surv_n = surv_n;
    src.cap_surv = cap_surv;
    src.slow = slow;
    src.slow_n = slow_n;
    src.cap_slow = cap_slow;
    src.err = err;
    src.set_lists(s_q[0][threadIdx.x / WAVE], s_q[1][threadIdx.x / WAVE]);
    src.init();
    (void)rplane;
    kmer_gate_ahead<RW>(src, rrec, recs, (const uint8_t *)s_map, xdrop, thresh);
}

// The batched gate in k-mer-major order (same launch shape as k_bgate_packed):
// the merged gate for SW = 8 (128-cell steps) and SW = 6 (96-cell steps, 32
// right cells at a time), kmer_gate_loop<SW> for the others. SUMS: the
// merged gate's score-sum form (add_a8), the cheapest the engine's matrix allows.
template <int SW, int BT, int SUMS = SUM_TREE8>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(2 * BT / 256, 8))) void k_bgate_kmer(
    const uint32_t *__restrict__ rplane, const uint32_t *__restrict__ pairtab, int32_t xdrop,
    int32_t thresh, const QRec *__restrict__ recs, const uint32_t *__restrict__ skm,
    const uint32_t *__restrict__ sval, const uint32_t *__restrict__ goff, uint32_t nk,
    const uint32_t *__restrict__ index, const uint32_t *__restrict__ pos, const uint32_t *__restrict__ clean,
    const uint4 *__restrict__ rrec, uint32_t seed_size, uint32_t claim, unsigned long long *__restrict__ work,
    SurvRec *__restrict__ surv,
    uint32_t *__restrict__ surv_n, uint32_t cap_surv, SurvRec *__restrict__ slow, uint32_t *__restrict__ slow_n,
    uint32_t cap_slow, uint32_t *__restrict__ err) {
    __shared__ uint32_t s_map[G2_MAP_WORDS];
    __shared__ uint4 s_dq[BT / WAVE][WQ_CAP];
    load_map_table(s_map, pairtab, BT);
    KmerSource src;
    src.skm = skm;
    src.sval = sval;
    src.goff = goff;
    src.index = index;
    src.pos = pos;
    src.clean = clean;
    src.work = work;
    src.total = __builtin_amdgcn_readfirstlane(goff[nk]);
    src.claim = claim;
    src.seed_size = seed_size;
    src.surv = surv;
    src.surv_n = surv_n;
    src.cap_surv = cap_surv;
    src.slow = slow;
    src.slow_n = slow_n;
    src.cap_slow = cap_slow;
    src.err = err;
    src.set_lists(nullptr, s_dq[threadIdx.x / WAVE]);
    src.init();
    if constexpr (SW == 8)
        kmer_gate_merged<4, SUMS>(src, rplane, rrec, recs, (const uint8_t *)s_map, xdrop, thresh);
    else if constexpr (SW == 6)
        kmer_gate_merged<2, SUMS>(src, rplane, rrec, recs, (const uint8_t *)s_map, xdrop, thresh);
    else kmer_gate_loop<SW>(src, rplane, recs, (const uint8_t *)s_map, xdrop, thresh);
}

// ---------------------------------------------------------------- coarse bound
//
// A cheaper upper bound of a hit's X-drop total that rejects almost every
// random hit before the two-cell maps run. Each cell is scored by the best
// score of its class under the engine's matrix: match, transition (codes
// xor 2) or transversion (xor odd), so the class score u >= the cell's score.
// Over a side, U_k = sum of u (prefix) >= P_k (the true prefix), and for j <=
// k, U_j - U_k <= P_j - P_k; so wherever the u-process has fallen more than X
// below its running maximum, the true process already has, and the true
// extension ends no later. Checking the fall only at 16-cell word boundaries
// (whose U are exact) declares a drop only where the u-process has dropped.
// The true maximum is then at most the largest U_start(word) + (positive
// class scores of the word) over the words up to that boundary: B. A hit
// whose sides both drop within the windows and whose B_right + B_left fails
// the gate's keep test (monotone in the total) fails the exact test too. Per
// word: the classes come from the xor of the two Gray-coded words (gray2) as
// two popcounts.

// A 16-cell word whose Gray-coded xor x has nd = popc((x | x >> 1) & odd-cell
// mask) differing cells and np = popc(x) set bits (a transversion sets one, a
// transition two): nt = np - nd transitions, nv = 2 nd - np transversions,
// 16 - nd matches. Its class-score sum is s0 + sd * nd + sp * np (s0 = 16 um,
// sd = 2 uv - ut - um, sp = ut - uv), the sum of its positive class scores
// p0 + pd * nd + pp * np (the same with pm, pt, pv = max(u, 0)).
// The paired form takes the same counts with biases: where small integers
// bd, bp with sd bd + sp bp = s0 exist, the counts are taken as nd' = nd + bd
// and np' = np + bp (free: the popcount instruction adds an accumulator), so a
// word's sum is sd nd' + sp np' with no constant term (c0 = 0; else c0 = s0
// and no biases). The tracked bound U + pd nd' + pp np' is the true one less
// the constant pc = p0 - pd bd - pp bp, added back at the decision.
struct CoarseCoef {
    int32_t s0, sd, sp, p0, pd, pp;
    int32_t c0, bd, bp, pc;  // paired form; c0 = s0, or 0 with biases
    int32_t lm;              // paired form: class-score sum that sends an undropped candidate to the long gate
};

constexpr int CW_RIGHT = 4, CW_LEFT = 7;  // coarse windows: 64 right cells, 112 left cells
static_assert(CW_RIGHT <= (int)G2_RIGHT_WORDS && CW_LEFT <= (int)G2_LEFT_WORDS, "windows within the records");

// Differing cells of a Gray-coded xor: one bit per cell at its even position.
__device__ __forceinline__ uint32_t differ_bits(uint32_t x) { return (x | (x >> 1)) & 0x55555555u; }

// The counts are at most 32 and the class scores small, so the products are
// 24-bit multiply-adds (full rate; a 32-bit multiply is quarter rate).
__device__ __forceinline__ void coarse_word(uint32_t r, uint32_t q, const CoarseCoef &cc, int32_t xdrop,
                                            int32_t &U, int32_t &mx, int32_t &B, bool &dropped) {
    const uint32_t x = r ^ q;
    const int32_t nd = __popc(differ_bits(x)), np = __popc(x);
    const int32_t bnd = U + cc.p0 + __mul24(cc.pd, nd) + __mul24(cc.pp, np);
    B = max(B, dropped ? (int32_t)0x80000000 : bnd);
    U += cc.s0 + __mul24(cc.sd, nd) + __mul24(cc.sp, np);
    mx = max(mx, U);
    dropped = dropped || mx - U > xdrop;
}

// Candidate record: (R, reference-record index, query-record index |
// variant << CAND_VAR_SHIFT, Q): what k_bgate_exact loads directly, or, for
// the plane-reading candidates' gate, where its windows lie.
constexpr uint32_t CAND_VAR_SHIFT = 32 - GV_SHIFT;

// The coarse gate over the k-mer-major walk (KmerSource, reference records
// required): every hit with clean windows is bounded over CW_RIGHT right and
// CW_LEFT left words; a hit not rejected by the bound (or whose u-process
// outlives a window) becomes a candidate for k_bgate_exact; a hit whose
// windows meet an exception is deferred as in k_bgate_kmer. One iteration per
// hit.
template <int BT>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(2 * BT / 256, 8))) void k_bgate_coarse(
    const QRec *__restrict__ recs, const uint32_t *__restrict__ skm, const uint32_t *__restrict__ sval,
    const uint32_t *__restrict__ goff, uint32_t nk, const uint32_t *__restrict__ index,
    const uint32_t *__restrict__ pos, const uint32_t *__restrict__ clean, const uint4 *__restrict__ rrec,
    uint32_t seed_size, uint32_t claim, unsigned long long *__restrict__ work, CoarseCoef cc, int32_t xdrop,
    int32_t thresh, uint4 *__restrict__ cand, uint32_t *__restrict__ cand_n, uint32_t cap_cand,
    SurvRec *__restrict__ slow, uint32_t *__restrict__ slow_n, uint32_t cap_slow, uint32_t *__restrict__ err) {
    __shared__ uint4 s_q[2][BT / WAVE][WQ_CAP];
    KmerSource src;
    src.skm = skm;
    src.sval = sval;
    src.goff = goff;
    src.index = index;
    src.pos = pos;
    src.clean = clean;
    src.work = work;
    src.total = __builtin_amdgcn_readfirstlane(goff[nk]);
    src.claim = claim;
    src.seed_size = seed_size;
    src.surv = (SurvRec *)cand;
    src.surv_n = cand_n;
    src.cap_surv = cap_cand;
    src.slow = slow;
    src.slow_n = slow_n;
    src.cap_slow = cap_slow;
    src.err = err;
    src.set_lists(s_q[0][threadIdx.x / WAVE], s_q[1][threadIdx.x / WAVE]);
    src.init();
    bool has = false, item = false, rok = false;
    uint32_t R = 0, s_cur = 0, s_end = 0, iidx = 0;
    uint32_t A0[CW_RIGHT], LA0[G2_LEFT_WORDS];
#pragma unroll
    for (int k = 0; k < CW_RIGHT; k++) A0[k] = 0;
#pragma unroll
    for (int k = 0; k < (int)G2_LEFT_WORDS; k++) LA0[k] = 0;
    for (;;) {
        bool fresh_r = false, fresh_q = false;
        uint32_t ridx = 0;
        src.refill(has, item, fresh_r, fresh_q, R, s_cur, s_end, rok, ridx);
        if (__ballot(has) == 0) break;
        if (fresh_r && rok) {
            const uint4 *rr = rrec + 3 * (size_t)ridx;
            const uint4 w0 = rr[0], w1 = rr[1], w2 = rr[2];
            A0[0] = w0.x, A0[1] = w0.y, A0[2] = w0.z, A0[3] = w0.w;
            LA0[0] = w1.x, LA0[1] = w1.y, LA0[2] = w1.z, LA0[3] = w1.w;
            LA0[4] = w2.x, LA0[5] = w2.y, LA0[6] = w2.z, LA0[7] = w2.w;
        }
        iidx = fresh_r ? ridx : iidx;
        // every lane with a hit has a fresh one: a hit takes one iteration
        const uint32_t gv = src.sval[has ? s_cur : 0u];
        const uint32_t var = gv & ((1u << GV_SHIFT) - 1), qi = gv >> GV_SHIFT;
        const uint4 *rec = (const uint4 *)(recs + qi);
        const uint4 v0 = rec[0], v1 = rec[1], v2 = rec[2], v3 = rec[3];
        s_cur += has ? 1u : 0u;
        const uint32_t B[CW_RIGHT] = {v0.x, v0.y, v0.z, v0.w};
        const uint32_t LB[G2_LEFT_WORDS] = {v1.x, v1.y, v1.z, v1.w, v2.x, v2.y, v2.z, v2.w};
        int32_t Ur = 0, mxr = 0, Br = 0, Ul = 0, mxl = 0, Bl = 0;
        bool dr = false, dl = false;
#pragma unroll
        for (int k = 0; k < CW_LEFT; k++) {
            if (k < CW_RIGHT) coarse_word(A0[k], B[k], cc, xdrop, Ur, mxr, Br, dr);
            coarse_word(LA0[k], LB[k], cc, xdrop, Ul, mxl, Bl, dl);
        }
        const bool ok = rok && v3.z != 0;
        const bool defer = has && !ok;
        const bool pass = has && ok && (!(dr && dl) || packed_keep(max(Br, 0) + max(Bl, 0), thresh));
        src.sq.push(pass, make_uint4(R, iidx, qi | var << CAND_VAR_SHIFT, v3.x), err);
        src.defer(defer, v3.y + var, R, v3.x);
        has = false;
    }
    src.finish();
}

// ------------------------------------------------- coarse bound, paired hits
//
// The same bound for two hits of one item at once, in packed 16-bit
// arithmetic: a lane takes its item's next two query records (hit a in the
// low halves, hit b in the high halves; the reference words are the item's,
// shared), so every word's sums, maximum and drop test serve both hits in
// one instruction and the per-iteration bookkeeping is paid once per two
// hits. The bound B is taken over every window word, drop or not (a larger
// value than the maximum up to the drop, so still an upper bound of the true
// maximum). Values in range: with every class score within +-a, a side's
// prefix sums over 8 words lie within +-128a, U - max within [-256a, 16a] and
// the tracked bound within +-(128a + |pc|), so a <= 127 and 128a + |pc| <=
// 32767 keep every compared value within 16 bits (the host checks both);
// sums and products wrap like the 16-bit ring they are, exact wherever the
// result is in range.
typedef short s16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ s16x2 as_s16x2(uint32_t v) { return __builtin_bit_cast(s16x2, v); }
__device__ __forceinline__ s16x2 splat16(int32_t v) { return as_s16x2(((uint32_t)v & 0xffffu) * 0x10001u); }

constexpr int32_t COARSE_PAIR_MAX_CLASS = 127;  // largest |class score| the packed form holds

struct CoarsePair {
    s16x2 sd, sp, c0, pd, pp;
    int32_t bd, bp;
};

// Lanes of `lo` in the low half and of `hi` in the high half (a byte select).
__device__ __forceinline__ uint32_t pack16(uint32_t lo, uint32_t hi) {
    return __builtin_amdgcn_perm(hi, lo, 0x05040100u);
}

// One 16-cell word of two chains (Gray-coded query words qa and qb against
// reference word r); U, mx, mg (minimum of U - running maximum) and B (the
// bound less pc) per half. PP: the set-bit count adds to the positive part.
// C0: the sums carry a constant (no bias pair exists).
template <bool PP, bool C0>
__device__ __forceinline__ void coarse_pair_word(uint32_t r, uint32_t qa, uint32_t qb, const CoarsePair &cp,
                                                 s16x2 &U, s16x2 &mx, s16x2 &mg, s16x2 &B) {
    const uint32_t xa = r ^ qa, xb = r ^ qb;
    const s16x2 np = as_s16x2(pack16(__popc(xa) + cp.bp, __popc(xb) + cp.bp));
    const s16x2 nd = as_s16x2(pack16(__popc(differ_bits(xa)) + cp.bd, __popc(differ_bits(xb)) + cp.bd));
    s16x2 bnd = U + cp.pd * nd;
    if constexpr (PP) bnd = bnd + cp.pp * np;
    B = __builtin_elementwise_max(B, bnd);
    s16x2 Un = U + cp.sd * nd + cp.sp * np;
    if constexpr (C0) Un = Un + cp.c0;
    mg = __builtin_elementwise_min(mg, Un - mx);
    mx = __builtin_elementwise_max(mx, Un);
    U = Un;
}

// k_bgate_coarse with two hits per lane and iteration (the second half idle
// when the item has one record left); same arguments, lists and decisions.
template <int BT, bool PP, bool C0>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(2 * BT / 256, 8))) void k_bgate_coarse2(
    const QRec *__restrict__ recs, const uint32_t *__restrict__ skm, const uint32_t *__restrict__ sval,
    const uint32_t *__restrict__ goff, uint32_t nk, const uint32_t *__restrict__ index,
    const uint32_t *__restrict__ pos, const uint32_t *__restrict__ clean, const uint4 *__restrict__ rrec,
    uint32_t seed_size, uint32_t claim, unsigned long long *__restrict__ work, CoarseCoef cc, int32_t xdrop,
    int32_t thresh, uint4 *__restrict__ cand, uint32_t *__restrict__ cand_n, uint32_t cap_cand,
    SurvRec *__restrict__ slow, uint32_t *__restrict__ slow_n, uint32_t cap_slow, uint32_t *__restrict__ err) {
    __shared__ uint4 s_q[2][BT / WAVE][WQ_CAP];
    KmerSource src;
    src.skm = skm;
    src.sval = sval;
    src.goff = goff;
    src.index = index;
    src.pos = pos;
    src.clean = clean;
    src.work = work;
    src.total = __builtin_amdgcn_readfirstlane(goff[nk]);
    src.claim = claim;
    src.seed_size = seed_size;
    src.surv = (SurvRec *)cand;
    src.surv_n = cand_n;
    src.cap_surv = cap_cand;
    src.slow = slow;
    src.slow_n = slow_n;
    src.cap_slow = cap_slow;
    src.err = err;
    src.set_lists(s_q[0][threadIdx.x / WAVE], s_q[1][threadIdx.x / WAVE]);
    src.init();
    const CoarsePair cp = {splat16(cc.sd), splat16(cc.sp), splat16(cc.c0), splat16(cc.pd), splat16(cc.pp), cc.bd,
                           cc.bp};
    const s16x2 pc = splat16(cc.pc);
    const s16x2 zero = splat16(0), lowest = splat16(-32768);
    bool has = false, item = false, rok = false;
    uint32_t R = 0, s_cur = 0, s_end = 0, iidx = 0;
    uint32_t A0[CW_RIGHT], LA0[G2_LEFT_WORDS];
#pragma unroll
    for (int k = 0; k < CW_RIGHT; k++) A0[k] = 0;
#pragma unroll
    for (int k = 0; k < (int)G2_LEFT_WORDS; k++) LA0[k] = 0;
    for (;;) {
        bool fresh_r = false, fresh_q = false;
        uint32_t ridx = 0;
        src.refill(has, item, fresh_r, fresh_q, R, s_cur, s_end, rok, ridx);
        if (__ballot(has) == 0) break;
        if (fresh_r && rok) {
            const uint4 *rr = rrec + 3 * (size_t)ridx;
            const uint4 w0 = rr[0], w1 = rr[1], w2 = rr[2];
            A0[0] = w0.x, A0[1] = w0.y, A0[2] = w0.z, A0[3] = w0.w;
            LA0[0] = w1.x, LA0[1] = w1.y, LA0[2] = w1.z, LA0[3] = w1.w;
            LA0[4] = w2.x, LA0[5] = w2.y, LA0[6] = w2.z, LA0[7] = w2.w;
        }
        iidx = fresh_r ? ridx : iidx;
        // a lane with a hit has a fresh one; hit b is its item's next record, if any
        const bool two = has && s_cur + 1 < s_end;
        const uint32_t ga = src.sval[has ? s_cur : 0u], gb = src.sval[two ? s_cur + 1 : 0u];
        s_cur += two ? 2u : has ? 1u : 0u;
        const uint32_t va = ga & ((1u << GV_SHIFT) - 1), qa = ga >> GV_SHIFT;
        const uint32_t vb = gb & ((1u << GV_SHIFT) - 1), qb = gb >> GV_SHIFT;
        const uint4 *ra = (const uint4 *)(recs + qa), *rb = (const uint4 *)(recs + qb);
        const uint4 a0 = ra[0], a1 = ra[1], a2 = ra[2], a3 = ra[3];
        const uint4 b0 = rb[0], b1 = rb[1], b2 = rb[2], b3 = rb[3];
        const uint32_t QA[CW_RIGHT] = {a0.x, a0.y, a0.z, a0.w}, QB[CW_RIGHT] = {b0.x, b0.y, b0.z, b0.w};
        const uint32_t LQA[G2_LEFT_WORDS] = {a1.x, a1.y, a1.z, a1.w, a2.x, a2.y, a2.z, a2.w};
        const uint32_t LQB[G2_LEFT_WORDS] = {b1.x, b1.y, b1.z, b1.w, b2.x, b2.y, b2.z, b2.w};
        s16x2 Ur = zero, mxr = zero, mgr = zero, Br = lowest, Ul = zero, mxl = zero, mgl = zero, Bl = lowest;
#pragma unroll
        for (int k = 0; k < CW_LEFT; k++) {
            if (k < CW_RIGHT) coarse_pair_word<PP, C0>(A0[k], QA[k], QB[k], cp, Ur, mxr, mgr, Br);
            coarse_pair_word<PP, C0>(LA0[k], LQA[k], LQB[k], cp, Ul, mxl, mgl, Bl);
        }
        // both sides dropped: the larger of the two minima below -X; the bound's
        // total (each side's bound at least 0, the sum within 0..65535) in u16
        const s16x2 dd = __builtin_elementwise_max(mgr, mgl);
        const s16x2 tt = __builtin_elementwise_max(Br + pc, zero) + __builtin_elementwise_max(Bl + pc, zero);
        const int32_t X = xdrop;
        const bool da = dd.x < -X, db = dd.y < -X;
        const int32_t ta = (uint16_t)tt.x, tb = (uint16_t)tt.y;
        const bool oka = rok && a3.z != 0, okb = rok && b3.z != 0;
        const bool pa = has && oka && (!da || packed_keep(ta, thresh));
        const bool pb = two && okb && (!db || packed_keep(tb, thresh));
        src.sq.push(pa, make_uint4(R, iidx, qa | va << CAND_VAR_SHIFT, a3.x), err);
        src.sq.push(pb, make_uint4(R, iidx, qb | vb << CAND_VAR_SHIFT, b3.x), err);
        src.defer(has && !oka, a3.y + va, R, a3.x);
        src.defer(two && !okb, b3.y + vb, R, b3.x);
        has = false;
    }
    src.finish();
}

// One 16-cell word of two chains that share the query word q and differ in
// the reference word (ra for the low halves, rb for the high halves);
// otherwise coarse_pair_word.
template <bool PP, bool C0>
__device__ __forceinline__ void coarse_pair_word_q(uint32_t ra, uint32_t rb, uint32_t q, const CoarsePair &cp,
                                                   s16x2 &U, s16x2 &mx, s16x2 &mg, s16x2 &B) {
    coarse_pair_word<PP, C0>(q, ra, rb, cp, U, mx, mg, B);
}

// The coarse bound with two reference positions per lane: a lane takes two
// consecutive items of a k-mer group (the second absent at the group's end)
// and walks the group's query records once for both, one record per
// iteration, so the hits of an iteration share their query words (hit a: the
// first item in the low halves, hit b: the second in the high halves). Both
// items of a lane meet the same records, so they finish together and a lane
// is refilled once per group walk; only a group with an odd number of
// positions leaves a half idle. Same arguments, lists and decisions as
// k_bgate_coarse2.
template <int BT, bool PP, bool C0>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(2 * BT / 256, 8))) void k_bgate_coarse2p(
    const QRec *__restrict__ recs, const uint32_t *__restrict__ skm, const uint32_t *__restrict__ sval,
    const uint32_t *__restrict__ goff, uint32_t nk, const uint32_t *__restrict__ index,
    const uint32_t *__restrict__ pos, const uint32_t *__restrict__ clean, const uint4 *__restrict__ rrec,
    uint32_t seed_size, uint32_t claim, unsigned long long *__restrict__ work, CoarseCoef cc, int32_t xdrop,
    int32_t thresh, uint4 *__restrict__ cand, uint32_t *__restrict__ cand_n, uint32_t cap_cand,
    SurvRec *__restrict__ slow, uint32_t *__restrict__ slow_n, uint32_t cap_slow, uint32_t *__restrict__ err) {
    __shared__ uint4 s_q[2][BT / WAVE][WQ_CAP];
    KmerSource src;
    src.skm = skm;
    src.sval = sval;
    src.goff = goff;
    src.index = index;
    src.pos = pos;
    src.clean = clean;
    src.work = work;
    src.total = __builtin_amdgcn_readfirstlane(goff[nk]);
    src.claim = claim;
    src.seed_size = seed_size;
    src.surv = (SurvRec *)cand;
    src.surv_n = cand_n;
    src.cap_surv = cap_cand;
    src.slow = slow;
    src.slow_n = slow_n;
    src.cap_slow = cap_slow;
    src.err = err;
    src.set_lists(s_q[0][threadIdx.x / WAVE], s_q[1][threadIdx.x / WAVE]);
    src.init();
    const CoarsePair cp = {splat16(cc.sd), splat16(cc.sp), splat16(cc.c0), splat16(cc.pd), splat16(cc.pp), cc.bd,
                           cc.bp};
    const s16x2 pc = splat16(cc.pc);
    const s16x2 zero = splat16(0), lowest = splat16(-32768);
    // the lane's items: reference positions e (hit a) and e + 1 (hit b, if `two`)
    bool has = false, item = false, two = false, roka = false, rokb = false;
    uint32_t Ra = 0, Rb = 0, s_cur = 0, s_end = 0, e = 0;
    uint32_t A[CW_RIGHT], LA[CW_LEFT], B[CW_RIGHT], LB[CW_LEFT];
#pragma unroll
    for (int k = 0; k < CW_RIGHT; k++) A[k] = B[k] = 0;
#pragma unroll
    for (int k = 0; k < CW_LEFT; k++) LA[k] = LB[k] = 0;
    for (;;) {
        // a lane continues with its items' next query record, else takes a new pair of items
        if (!has && item) {
            if (s_cur < s_end) has = true;
            else item = false;
        }
        bool fresh = false;
        uint64_t need = __ballot(!has);
        while (need) {
            if (src.j >= src.n) {
                if (!src.next_group()) break;
                continue;
            }
            const uint32_t avail = src.n - src.j, rank = mbcnt(need);
            if (!has && 2 * rank < avail) {
                e = src.b0 + src.j + 2 * rank;
                two = 2 * rank + 1 < avail;
                s_cur = src.g0;
                s_end = src.g1;
                has = item = fresh = true;
            }
            src.j += min(2 * (uint32_t)__popcll(need), avail);
            need = __ballot(!has);
        }
        if (__ballot(has) == 0) break;
        if (fresh) {
            const uint32_t pa = pos[e], pb = two ? pos[e + 1] : 0u;
            Ra = pa + seed_size;
            Rb = pb + seed_size;
            roka = ((clean[e >> 5] >> (e & 31)) & 1u) != 0;
            rokb = two && ((clean[(e + 1) >> 5] >> ((e + 1) & 31)) & 1u) != 0;
            const uint4 *ra = rrec + 3 * (size_t)e;
            const uint4 a0 = ra[0], a1 = ra[1], a2 = ra[2];
            const uint4 *rb = rrec + 3 * (size_t)(two ? e + 1 : e);
            const uint4 b0 = rb[0], b1 = rb[1], b2 = rb[2];
            A[0] = a0.x, A[1] = a0.y, A[2] = a0.z, A[3] = a0.w;
            LA[0] = a1.x, LA[1] = a1.y, LA[2] = a1.z, LA[3] = a1.w, LA[4] = a2.x, LA[5] = a2.y, LA[6] = a2.z;
            B[0] = b0.x, B[1] = b0.y, B[2] = b0.z, B[3] = b0.w;
            LB[0] = b1.x, LB[1] = b1.y, LB[2] = b1.z, LB[3] = b1.w, LB[4] = b2.x, LB[5] = b2.y, LB[6] = b2.z;
        }
        const uint32_t g = src.sval[has ? s_cur : 0u];
        s_cur += has ? 1u : 0u;
        const uint32_t var = g & ((1u << GV_SHIFT) - 1), qi = g >> GV_SHIFT;
        const uint4 *rq = (const uint4 *)(recs + qi);
        const uint4 q0 = rq[0], q1 = rq[1], q2 = rq[2], q3 = rq[3];
        const uint32_t QW[CW_RIGHT] = {q0.x, q0.y, q0.z, q0.w};
        const uint32_t LQ[CW_LEFT] = {q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z};
        s16x2 Ur = zero, mxr = zero, mgr = zero, Br = lowest, Ul = zero, mxl = zero, mgl = zero, Bl = lowest;
#pragma unroll
        for (int k = 0; k < CW_LEFT; k++) {
            if (k < CW_RIGHT) coarse_pair_word_q<PP, C0>(A[k], B[k], QW[k], cp, Ur, mxr, mgr, Br);
            coarse_pair_word_q<PP, C0>(LA[k], LB[k], LQ[k], cp, Ul, mxl, mgl, Bl);
        }
        const s16x2 dd = __builtin_elementwise_max(mgr, mgl);
        const s16x2 tt = __builtin_elementwise_max(Br + pc, zero) + __builtin_elementwise_max(Bl + pc, zero);
        const int32_t X = xdrop;
        const bool da = dd.x < -X, db = dd.y < -X;
        const int32_t ta = (uint16_t)tt.x, tb = (uint16_t)tt.y;
        const bool qok = q3.z != 0;
        const bool hb = has && two;
        const bool oka = roka && qok, okb = rokb && qok;
        const bool pa = has && oka && (!da || packed_keep(ta, thresh));
        const bool pb = hb && okb && (!db || packed_keep(tb, thresh));
        const uint32_t qv = qi | var << CAND_VAR_SHIFT;
        // A candidate with a side whose u-process never dropped and a side
        // whose class-score sum reached cc.lm is most likely a homologous
        // extension that outlives the two-cell maps' windows (which would
        // defer it after loading them): it goes straight to the deferred-hit
        // gate, which decides every hit from its seed.
        const s16x2 uu = __builtin_elementwise_max(Ur, Ul);
        const bool la = pa && !da && uu.x >= cc.lm, lb = pb && !db && uu.y >= cc.lm;
        src.sq.push(pa && !la, make_uint4(Ra, e, qv, q3.x), err);
        src.sq.push(pb && !lb, make_uint4(Rb, e + 1, qv, q3.x), err);
        src.defer((has && !oka) || la, q3.y + var, Ra, q3.x);
        src.defer((hb && !okb) || lb, q3.y + var, Rb, q3.x);
        has = false;
    }
    src.finish();
}

// Candidates binned by reference position before the candidates' gate reads
// their windows from the planes: bin = R >> shift, at most CAND_BINS bins (a
// shift that makes each bin's reference-plane segment a few hundred KiB, which
// the L2 of an XCD holds while its waves work through the bin).
constexpr uint32_t CAND_BINS = 1024;
// Query bases of a pass up to which binning pays (a 3 MiB query plane): the
// binned order splits the pairs of hits that share a query window, so the
// pass's query plane has to stay in an XCD's L2 beside the bins' reference
// lines (4 MiB).
constexpr uint32_t CAND_BIN_MAX_QUERY = 12u << 20;
// XCDs of the device: consecutive workgroups go to consecutive XCDs.
constexpr uint32_t XCDS = 8;
// Entries of a wave's chunk in interleaved shares of a binned list.
constexpr uint32_t CAND_CHUNK = 64;

__device__ __forceinline__ uint32_t cand_bin(uint32_t R, uint32_t shift) { return min(R >> shift, CAND_BINS - 1); }

// Block b's share [c0, c1) of a list of n entries; the histogram and the
// scatter split the list alike.
__device__ __forceinline__ void cand_share(uint32_t n, uint32_t b, uint32_t nb, uint32_t &c0, uint32_t &c1) {
    c0 = (uint32_t)((uint64_t)n * b / nb);
    c1 = (uint32_t)((uint64_t)n * (b + 1) / nb);
}

// Per-block bin sizes: block b counts the bins of its share of the candidate
// list into table[bin * gridDim.x + b]; CAND_BINS threads.
__global__ __launch_bounds__(CAND_BINS) void k_cand_hist(const uint4 *__restrict__ cand,
                                                         const uint32_t *__restrict__ cand_n, uint32_t cap,
                                                         uint32_t shift, uint32_t *__restrict__ table) {
    __shared__ uint32_t h[CAND_BINS];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t c0, c1;
    cand_share(min(*cand_n, cap), blockIdx.x, gridDim.x, c0, c1);
    const uint32_t *r = (const uint32_t *)cand;  // word 0 of each record is R
    for (uint32_t i = c0 + threadIdx.x; i < c1; i += CAND_BINS) atomicAdd(&h[cand_bin(r[4 * (size_t)i], shift)], 1u);
    __syncthreads();
    table[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// Each bin's row of the table (one entry per scatter block) to its exclusive
// prefix sums, and the bin's size to sizes[bin]; one 256-thread block per bin.
__global__ __launch_bounds__(256) void k_cand_rows(uint32_t *__restrict__ table, uint32_t nblk,
                                                   uint32_t *__restrict__ sizes) {
    uint32_t *row = table + (size_t)blockIdx.x * nblk;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nblk; b0 += 256) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < nblk ? row[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan<256>(v, &tot);
        if (i < nblk) row[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) sizes[blockIdx.x] = carry;
}

// The bins' first entries: bases = exclusive scan of sizes; one block of
// CAND_BINS threads.
__global__ __launch_bounds__(CAND_BINS) void k_cand_scan(const uint32_t *__restrict__ sizes,
                                                         uint32_t *__restrict__ bases) {
    uint32_t tot;
    bases[threadIdx.x] = block_excl_scan<CAND_BINS>(sizes[threadIdx.x], &tot);
}

// The candidates in bin order: block b writes the records of its share that
// fall in a bin from bases[bin] + table[bin * gridDim.x + b] on. Each tile of
// CPT records per thread is counting-sorted by bin in LDS first, so that a
// bin's records of the tile go out as one run of consecutive stores (order
// within a bin is free: the gate's decisions do not depend on it); the next
// tile's records are loaded while the current one is sorted.
template <int CPT>
__global__ __launch_bounds__(CAND_BINS) void k_cand_scatter(const uint4 *__restrict__ cand,
                                                            const uint32_t *__restrict__ cand_n, uint32_t cap,
                                                            uint32_t shift, const uint32_t *__restrict__ table,
                                                            const uint32_t *__restrict__ bases,
                                                            uint4 *__restrict__ out) {
    constexpr uint32_t TILE = CPT * CAND_BINS;
    __shared__ uint4 s_rec[TILE];
    __shared__ uint32_t s_cnt[CAND_BINS], s_start[CAND_BINS], s_cur[CAND_BINS];
    const uint32_t t = threadIdx.x;
    s_cur[t] = bases[t] + table[(size_t)t * gridDim.x + blockIdx.x];
    uint32_t c0, c1;
    cand_share(min(*cand_n, cap), blockIdx.x, gridDim.x, c0, c1);
    uint4 nxt[CPT];
#pragma unroll
    for (int k = 0; k < CPT; k++) {
        const uint32_t i = c0 + t + k * CAND_BINS;
        nxt[k] = i < c1 ? cand[i] : make_uint4(0, 0, 0, 0);
    }
    for (uint32_t t0 = c0; t0 < c1; t0 += TILE) {
        const uint32_t m = min(TILE, c1 - t0);
        uint4 rec[CPT];
        uint32_t bin[CPT], rank[CPT];
#pragma unroll
        for (int k = 0; k < CPT; k++) {
            rec[k] = nxt[k];
            const uint32_t i = t0 + TILE + t + k * CAND_BINS;
            nxt[k] = i < c1 ? cand[i] : make_uint4(0, 0, 0, 0);
        }
        s_cnt[t] = 0;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < CPT; k++) {
            bin[k] = cand_bin(rec[k].x, shift);
            rank[k] = t + k * CAND_BINS < m ? atomicAdd(&s_cnt[bin[k]], 1u) : 0u;
        }
        __syncthreads();
        const uint32_t cnt = s_cnt[t];
        uint32_t tot;
        s_start[t] = block_excl_scan<CAND_BINS>(cnt, &tot);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < CPT; k++)
            if (t + k * CAND_BINS < m) s_rec[s_start[bin[k]] + rank[k]] = rec[k];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < CPT; k++) {
            const uint32_t j = t + k * CAND_BINS;
            if (j < m) {
                const uint4 r = s_rec[j];
                const uint32_t b = cand_bin(r.x, shift);
                out[s_cur[b] + j - s_start[b]] = r;
            }
        }
        __syncthreads();
        s_cur[t] += cnt;
    }
}

// Hit source over the coarse gate's candidates: waves claim `claim`
// candidates at a time from a grid-wide counter and hand them to idle lanes,
// each one hit on its own item. A lane's s_cur is the candidate's query-record
// index with the variant in its top bits.
struct CandSource : GateOut {
    const uint4 *list;
    unsigned long long *work;
    uint32_t n, claim, c_cur, c_end;
    uint64_t s_next;                   // interleaved shares: the wave's next chunk,
    uint32_t s_step, s_chunk, s_hi;    // the chunks' stride and size, the share's end (s_step 0: claims)
    bool dry;

    __device__ __forceinline__ bool take() {
        if (dry) return false;
        if (s_step) {
            if (s_next >= s_hi) {
                dry = true;
                return false;
            }
            c_cur = (uint32_t)s_next;
            c_end = s_next + s_chunk < (uint64_t)s_hi ? (uint32_t)(s_next + s_chunk) : s_hi;
            s_next += s_step;
            return true;
        }
        uint32_t leader = (uint32_t)(__ffsll((unsigned long long)__ballot(true)) - 1);
        unsigned long long v = 0;
        if (lane_id() == leader) v = atomicAdd(work, (unsigned long long)claim);
        const uint64_t start = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) |
                               (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
        if (start >= (uint64_t)n) {
            dry = true;
            return false;
        }
        c_cur = (uint32_t)start;
        c_end = start + claim < (uint64_t)n ? (uint32_t)(start + claim) : n;
        return true;
    }
    // A fixed share of the list instead of claims: wave w of `waves` takes
    // entries [n w / waves, n (w + 1) / waves) and no atomic (a claim's
    // returned value would make the wave wait on every load in flight).
    __device__ __forceinline__ void share(uint32_t w, uint32_t waves) {
        c_cur = (uint32_t)((uint64_t)n * w / waves);
        c_end = (uint32_t)((uint64_t)n * (w + 1) / waves);
        dry = true;
    }
    // Interleaved shares of [lo, hi): wave w of `waves` takes the chunks of
    // `chunk` entries w, w + waves, w + 2 waves, ..., so that the waves work on
    // neighbouring entries at any time (no atomic either).
    __device__ __forceinline__ void interleave(uint32_t lo, uint32_t hi, uint32_t w, uint32_t waves,
                                               uint32_t chunk) {
        c_cur = c_end = 0;
        s_next = (uint64_t)lo + (uint64_t)w * chunk;
        s_step = waves * chunk;
        s_chunk = chunk;
        s_hi = hi;
        dry = false;
    }
    __device__ __forceinline__ void init() {
        c_cur = c_end = 0;
        s_step = 0;
        dry = false;
    }
    // Hands each lane with `want` its next candidate entry; whether it got one.
    // (The load is issued after the loop: inside it, the compiler would wait
    // for every load in flight before each pass.)
    __device__ __forceinline__ bool claim_entry(bool want, uint4 &e) {
        bool got = false;
        uint32_t at = 0;
        uint64_t need = __ballot(want);
        while (need) {
            if (c_cur >= c_end) {
                if (!take()) break;
                continue;
            }
            const uint32_t avail = c_end - c_cur, rank = mbcnt(need);
            if (want && !got && rank < avail) {
                at = c_cur + rank;
                got = true;
            }
            c_cur += min((uint32_t)__popcll(need), avail);
            need = __ballot(want && !got);
        }
        if (got) e = list[at];
        return got;
    }
    __device__ __forceinline__ const uint4 *record(const QRec *recs, uint32_t p, uint32_t &var) const {
        var = p >> CAND_VAR_SHIFT;
        return (const uint4 *)(recs + (p & ((1u << CAND_VAR_SHIFT) - 1)));
    }
    __device__ __forceinline__ void refill(bool &has, bool &item, bool &fresh_r, bool &fresh_q, uint32_t &R,
                                           uint32_t &s_cur, uint32_t &s_end, bool &rok, uint32_t &ridx) {
        uint64_t need = __ballot(!has);
        while (need) {
            if (c_cur >= c_end) {
                if (!take()) break;
                continue;
            }
            const uint32_t avail = c_end - c_cur, rank = mbcnt(need);
            if (!has && rank < avail) {
                const uint4 e = list[c_cur + rank];
                R = e.x;
                ridx = e.y;
                s_cur = e.z;
                s_end = e.z + 1;
                rok = true;
                has = item = fresh_r = fresh_q = true;
            }
            c_cur += min((uint32_t)__popcll(need), avail);
            need = __ballot(!has);
        }
    }
};

// The merged two-cell-map gate (kmer_gate_merged, 96-cell steps) over the
// coarse gate's candidates: survivors and deferred hits as k_bgate_kmer's.
template <int BT, int SUMS>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(2 * BT / 256, 8))) void k_bgate_exact(
    const uint32_t *__restrict__ rplane, const uint32_t *__restrict__ pairtab, int32_t xdrop, int32_t thresh,
    const QRec *__restrict__ recs, const uint32_t *__restrict__ sval, const uint4 *__restrict__ rrec,
    const uint4 *__restrict__ cand, const uint32_t *__restrict__ cand_n, uint32_t cap_cand, uint32_t claim,
    unsigned long long *__restrict__ work, SurvRec *__restrict__ surv, uint32_t *__restrict__ surv_n,
    uint32_t cap_surv, SurvRec *__restrict__ slow, uint32_t *__restrict__ slow_n, uint32_t cap_slow,
    uint32_t *__restrict__ err, const uint32_t *__restrict__ qplane, uint32_t per, uint32_t interleave) {
    (void)qplane;
    (void)per;
    (void)interleave;
    __shared__ uint32_t s_map[G2_MAP_WORDS];
    __shared__ uint4 s_dq[BT / WAVE][WQ_CAP];
    load_map_table(s_map, pairtab, BT);
    CandSource src;
    src.list = cand;
    src.work = work;
    (void)sval;
    src.n = __builtin_amdgcn_readfirstlane(min(*cand_n, cap_cand));
    src.claim = claim;
    src.surv = surv;
    src.surv_n = surv_n;
    src.cap_surv = cap_surv;
    src.slow = slow;
    src.slow_n = slow_n;
    src.cap_slow = cap_slow;
    src.err = err;
    src.set_lists(nullptr, s_dq[threadIdx.x / WAVE]);
    src.init();
    kmer_gate_merged<2, SUMS>(src, rplane, rrec, recs, (const uint8_t *)s_map, xdrop, thresh);
}

// The 13 plane words around anchor P (word 8 holds base P) of a window whose
// margins were checked when the hit became a candidate.
__device__ __forceinline__ void load_window(const uint32_t *__restrict__ plane, uint32_t P, uint32_t (&W)[13]) {
    const uint32_t *wp = plane + (P >> 4) - G2_LEFT_WORDS;
#pragma unroll
    for (int k = 0; k < 13; k++) W[k] = wp[k];
}

// A window's right and left streams as the merged gate consumes them (the
// layout of the records, in plane code).
__device__ __forceinline__ void window_streams(const uint32_t (&W)[13], uint32_t P, uint32_t (&A)[G2_RIGHT_WORDS],
                                               uint32_t (&LA)[G2_LEFT_WORDS]) {
    const uint32_t sh = 2 * (P & 15);
#pragma unroll
    for (int k = 0; k < (int)G2_RIGHT_WORDS; k++) A[k] = __builtin_amdgcn_alignbit(W[9 + k], W[8 + k], sh);
#pragma unroll
    for (int k = 0; k < (int)G2_LEFT_WORDS; k++)
        LA[k] = nib_swap(__builtin_bswap32(__builtin_amdgcn_alignbit(W[8 - k], W[7 - k], sh)));
}

// The merged gate (RW = 2) over the coarse gate's candidates with its global
// loads issued ahead, in three stages per lane: a claimed candidate entry
// (loaded in one iteration), its next hit's windows (loads issued in the
// next, from the landed entry), and the current hit (promoted from those
// windows once the lane is free). Every load is issued one iteration before
// its data is used, so the counter wait at the top of an iteration finds them
// landed while the wave computed. Cells, decisions and lists are those of
// kmer_gate_merged<2, SUMS> over CandSource; a finished hit's results are
// appended at the top of the next iteration. PLANES = false: a candidate's windows are its
// reference and query records (Gray-coded; the Gray-keyed maps). PLANES =
// true: they are read from the 2-bit planes at R and Q (the reference plane
// is 250 MB at 1 Gbp, the pass's query plane a few MB, against 48 B per
// indexed position of reference records), in plane code (the plane-keyed
// maps), and the hit's seed slot is rank * per + variant.
template <int SUMS, bool PLANES = false>
__device__ __forceinline__ void cand_gate_ahead(CandSource &src, const uint4 *__restrict__ rrec,
                                                const QRec *__restrict__ recs, const uint32_t *__restrict__ rplane,
                                                const uint32_t *__restrict__ qplane, uint32_t per,
                                                const uint8_t *s_map, int32_t xdrop, int32_t thresh) {
    constexpr int RW = 2;
    const uint32_t off_fwd = (lane_id() & 31) * 4, off_rev = off_fwd + 128;
    // current hit
    bool has = false;
    HitSides st;
    uint32_t R = 0, Q = 0, tag = 0;
    uint32_t A0[G2_RIGHT_WORDS], LA0[G2_LEFT_WORDS], B[G2_RIGHT_WORDS], LB[G2_LEFT_WORDS];
    // next hit: its records (loads issued last iteration)
    bool nxt = false;
    uint32_t nR = 0, nvar = 0;
    const uint4 zero4 = make_uint4(0, 0, 0, 0);
    uint4 T0 = zero4, T1 = zero4, T2 = zero4, P0 = zero4, P1 = zero4, P2 = zero4, P3 = zero4;
    uint32_t WR[13], WQ[13];  // PLANES: the next hit's window words
    uint32_t nQ = 0;
#pragma unroll
    for (int k = 0; k < 13; k++) WR[k] = WQ[k] = 0;
    // claimed entry (load issued last iteration)
    bool ent = false;
    uint4 E = zero4;
    // the hit finished in the previous iteration
    bool e_pass = false, e_defer = false;
    uint32_t eR = 0, eQ = 0, etag = 0;
#pragma unroll
    for (int j = 0; j < (int)G2_RIGHT_WORDS; j++) A0[j] = B[j] = 0;
#pragma unroll
    for (int j = 0; j < (int)G2_LEFT_WORDS; j++) LA0[j] = LB[j] = 0;

    for (;;) {
        // The records' pad word and the entry's last word are read nowhere;
        // keeping them live to here stops the compiler from reusing their
        // registers while the loads are in flight (which would wait on them).
        if constexpr (PLANES) asm volatile("" ::"v"(E.y));
        else asm volatile("" ::"v"(P3.w), "v"(E.w));
        src.emit(true, e_pass, etag, eR, eQ);
        src.defer(e_defer, etag, eR, eQ);
        e_pass = e_defer = false;

        // 1. a free lane takes its next hit
        bool defer = false;
        if (!has && nxt) {
            if constexpr (PLANES) {
                window_streams(WR, nR, A0, LA0);
                window_streams(WQ, nQ, B, LB);
                Q = nQ;
                tag = nvar;
            } else {
                A0[0] = T0.x, A0[1] = T0.y, A0[2] = T0.z, A0[3] = T0.w;
                LA0[0] = T1.x, LA0[1] = T1.y, LA0[2] = T1.z, LA0[3] = T1.w;
                LA0[4] = T2.x, LA0[5] = T2.y, LA0[6] = T2.z, LA0[7] = T2.w;
                B[0] = P0.x, B[1] = P0.y, B[2] = P0.z, B[3] = P0.w;
                LB[0] = P1.x, LB[1] = P1.y, LB[2] = P1.z, LB[3] = P1.w;
                LB[4] = P2.x, LB[5] = P2.y, LB[6] = P2.z, LB[7] = P2.w;
                Q = P3.x;
                tag = P3.y + nvar;
                if (P3.z == 0) defer = true;
            }
            R = nR;
            has = true;
            nxt = false;
            st.start();
        }
        // 2. the landed entry's windows, for the lane's next hit
        if (!nxt && ent) {
            if constexpr (PLANES) {
                // every candidate's windows are clean (the coarse gate deferred the others)
                load_window(rplane, E.x, WR);
                load_window(qplane, E.w, WQ);
                const uint32_t rank = E.z & ((1u << CAND_VAR_SHIFT) - 1);
                nvar = rank * per + (E.z >> CAND_VAR_SHIFT);
                nQ = E.w;
            } else {
                const uint4 *rr = rrec + 3 * (size_t)E.y;
                T0 = rr[0], T1 = rr[1], T2 = rr[2];
                uint32_t var;
                const uint4 *rec = src.record(recs, E.z, var);
                P0 = rec[0], P1 = rec[1], P2 = rec[2], P3 = rec[3];
                nvar = var;
            }
            nR = E.x;
            nxt = true;
            ent = false;
        }
        // 3. a lane without an entry claims one
        if (src.claim_entry(!ent, E)) ent = true;
        if (__ballot(has || nxt || ent) == 0) break;

        // 4. the current hit's next stretches
        bool done = false, pass = false;
        if (has && !defer) {
            int32_t er = st.rcont ? st.er_keep : xdrop, cr = st.rcont ? st.cr_keep : 0, eminr = er;
            map_words<RW, G2_RIGHT_WORDS, SUMS>(s_map, A0, B, st.rcont, off_fwd, er, cr, eminr);
            int32_t el = st.cont ? st.el_keep : xdrop, cl = st.cont ? st.cl_keep : 0, eminl = el;
            map_stretch<8, SUMS>(s_map, LA0, LB, st.cont, off_rev, el, cl, eminl);
            st.decide<RW>(er, side_sum<SUMS>(cr), eminr, el, side_sum<SUMS>(cl), eminl, xdrop, thresh, defer, done,
                          pass);
        }
        if (has && (done || defer)) {
            e_pass = done && pass;
            e_defer = defer;
            eR = R;
            eQ = Q;
            etag = tag;
            has = false;
            st.cont = st.rcont = false;
        }
    }
    src.finish();
}

// cand_gate_ahead over the coarse gate's candidates: 512-thread blocks, four
// waves per SIMD (the 128 VGPRs its in-flight records need); same arguments
// and lists as k_bgate_exact.
template <int SUMS, bool PLANES = false>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_bgate_exact_ahead(
    const uint32_t *__restrict__ rplane, const uint32_t *__restrict__ pairtab, int32_t xdrop, int32_t thresh,
    const QRec *__restrict__ recs, const uint32_t *__restrict__ sval, const uint4 *__restrict__ rrec,
    const uint4 *__restrict__ cand, const uint32_t *__restrict__ cand_n, uint32_t cap_cand, uint32_t claim,
    unsigned long long *__restrict__ work, SurvRec *__restrict__ surv, uint32_t *__restrict__ surv_n,
    uint32_t cap_surv, SurvRec *__restrict__ slow, uint32_t *__restrict__ slow_n, uint32_t cap_slow,
    uint32_t *__restrict__ err, const uint32_t *__restrict__ qplane, uint32_t per, uint32_t interleave) {
    __shared__ uint32_t s_map[G2_MAP_WORDS];
    __shared__ uint4 s_dq[512 / WAVE][WQ_CAP];
    load_map_table(s_map, pairtab, 512);
    CandSource src;
    src.list = cand;
    src.work = work;
    (void)sval;
    src.n = __builtin_amdgcn_readfirstlane(min(*cand_n, cap_cand));
    src.claim = claim;
    src.surv = surv;
    src.surv_n = surv_n;
    src.cap_surv = cap_surv;
    src.slow = slow;
    src.slow_n = slow_n;
    src.cap_slow = cap_slow;
    src.err = err;
    src.set_lists(nullptr, s_dq[threadIdx.x / WAVE]);
    src.init();
    (void)work;
    // Blocks are dealt to the eight XCDs in turn. With binned candidates, the
    // waves of one XCD interleave small chunks of one eighth of the list, so
    // that they move through its bins together and the XCD's L2 holds the
    // reference-plane lines of the bin they are in.
    if (interleave && gridDim.x % XCDS == 0) {
        const uint32_t x = blockIdx.x % XCDS, per_x = gridDim.x / XCDS * (512 / WAVE);
        const uint32_t w = blockIdx.x / XCDS * (512 / WAVE) + threadIdx.x / WAVE;
        src.interleave((uint32_t)((uint64_t)src.n * x / XCDS), (uint32_t)((uint64_t)src.n * (x + 1) / XCDS), w,
                       per_x, CAND_CHUNK);
    } else {
        src.share(blockIdx.x * (512 / WAVE) + threadIdx.x / WAVE, gridDim.x * (512 / WAVE));
    }
    cand_gate_ahead<SUMS, PLANES>(src, rrec, recs, rplane, qplane, per, (const uint8_t *)s_map, xdrop, thresh);
}

// First index of each batch's records in a scope-sorted list (scope = batch
// * cpb + max_hits chunk): starts[b] = lower bound of b * cpb, for b in
// [0, nb] (starts[nb] = n).
__global__ void k_scope_starts(const uint32_t *__restrict__ scopes, uint32_t n, uint32_t nb, uint32_t cpb,
                               uint32_t *__restrict__ starts) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nb) return;
    const uint64_t key = (uint64_t)b * cpb;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (scopes[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    starts[b] = lo;
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

// Deferred hits extended one lane per hit past the main gate's fixed windows:
// each side in 64-cell stretches read from the 2-bit planes through the same
// two-cell maps (so the same upper bound of the exact score: every cell up to
// the drop, plus the rest of its stretch), until it drops. A hit one of whose
// stretches would touch an exception bit or come within a word of a sequence
// end goes to the exact wave gate instead (list `exact`). Both sides restart
// at the anchor.
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_bgate_long(
    const uint32_t *__restrict__ rplane, const uint32_t *__restrict__ rcoarse, uint32_t ref_len,
    const uint32_t *__restrict__ qplane, const uint32_t *__restrict__ qcoarse, uint32_t qlen,
    const uint32_t *__restrict__ pairtab, int32_t xdrop, int32_t thresh, const SurvRec *__restrict__ slow,
    const uint32_t *__restrict__ slow_n, uint32_t cap_slow, SurvRec *__restrict__ surv,
    uint32_t *__restrict__ surv_n, uint32_t cap_surv, SurvRec *__restrict__ exact,
    uint32_t *__restrict__ exact_n, uint32_t *__restrict__ err) {
    __shared__ uint32_t s_map[G2_MAP_WORDS];
    load_map_table(s_map, pairtab, 1024);
    const uint32_t n = min(*slow_n, cap_slow);
    const uint32_t off_fwd = (lane_id() & 31) * 4, off_rev = off_fwd + 128;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < n; base += stride) {
        const uint32_t i = base + lane_id();
        const bool have = i < n;
        SurvRec rec = {0, 0, 0, 0};
        if (have) rec = slow[i];
        bool ex = false;
        int32_t M[2] = {0, 0};
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const bool left = side == 1;
            int32_t e = xdrop, c = 0;
            bool live = have && !ex;
            for (uint32_t k0 = 0; __ballot(live) != 0; k0 += 64) {
                if (!live) continue;
                // a stretch plus one word of alignment slack on either side stays inside both planes
                bool inb = left ? (rec.R >= k0 + 80 && rec.Q >= k0 + 80)
                                : ((uint64_t)rec.R + k0 + 80 <= ref_len && (uint64_t)rec.Q + k0 + 80 <= qlen);
                if (!inb) {
                    ex = true;
                    live = false;
                    continue;
                }
                const uint32_t rb = left ? rec.R - k0 : rec.R + k0, qb = left ? rec.Q - k0 : rec.Q + k0;
                const uint32_t rlo = left ? rb - 64 : rb, qlo = left ? qb - 64 : qb;
                const uint32_t *rc = rcoarse + (rlo >> 11), *qc = qcoarse + (qlo >> 11);
                if (coarse_bits(rc[0], rc[1], rlo, rlo + 63) || coarse_bits(qc[0], qc[1], qlo, qlo + 63)) {
                    ex = true;
                    live = false;
                    continue;
                }
                // five plane words around the stretch, aligned to its first cell
                const uint32_t *rp = rplane + (rb >> 4) - (left ? 4 : 0), *qp = qplane + (qb >> 4) - (left ? 4 : 0);
                uint32_t RW[5], QW[5];
#pragma unroll
                for (int k = 0; k < 5; k++) RW[k] = rp[k], QW[k] = qp[k];
                const uint32_t sr = 2 * (rb & 15), sq = 2 * (qb & 15);
                uint32_t A[4], B[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    if (left) {
                        A[k] = nib_swap(__builtin_bswap32(__builtin_amdgcn_alignbit(RW[4 - k], RW[3 - k], sr)));
                        B[k] = nib_swap(__builtin_bswap32(__builtin_amdgcn_alignbit(QW[4 - k], QW[3 - k], sq)));
                    } else {
                        A[k] = __builtin_amdgcn_alignbit(RW[k + 1], RW[k], sr);
                        B[k] = __builtin_amdgcn_alignbit(QW[k + 1], QW[k], sq);
                    }
                }
                int32_t emin = e;
                map_stretch<4>((const uint8_t *)s_map, A, B, false, left ? off_rev : off_fwd, e, c, emin);
                if (emin < 0) {
                    M[side] = c + (xdrop - e);
                    live = false;
                }
            }
        }
        const bool pass = have && !ex && gate_keep((int32_t)((uint32_t)M[0] + (uint32_t)M[1]), thresh);
        SeedSource::append(pass, surv, surv_n, cap_surv, err, rec.R, rec.Q, rec.seed);
        SeedSource::append(have && ex, exact, exact_n, cap_slow, err, rec.R, rec.Q, rec.seed);
    }
}

// One 64-cell tile of wave_extend's recurrence (without positions): folds the
// tile's cell scores into the running score and maximum; true when the
// extension ends in this tile (a drop, or its last cell at a sequence edge),
// with the final maximum in prev_max.
__device__ __forceinline__ bool wave_tile(int32_t sc, bool edge, int32_t xdrop, int32_t &prev_score,
                                          int32_t &prev_max) {
    int32_t ps = (int32_t)((uint32_t)wave_incl_scan_add_i32(sc) + (uint32_t)prev_score);
    int32_t cand = ps > prev_max ? ps : prev_max;
    int32_t scanned = wave_incl_scan_max_i32(cand);
    uint64_t dropped = __ballot((int32_t)((uint32_t)scanned - (uint32_t)ps) > xdrop);
    uint32_t first = dropped ? (uint32_t)(__ffsll((unsigned long long)dropped) - 1) : WAVE;
    prev_max = first == 0 ? prev_max : __builtin_amdgcn_readlane(scanned, first - 1);
    if (dropped || edge) return true;
    prev_score = __builtin_amdgcn_readlane(ps, WAVE - 1);
    return false;
}

// The maximum of one X-drop extension (= wave_extend<false>'s score) with the
// bytes of its first P tiles loaded at once, so that an extension within
// 64 * P cells costs one memory round trip; later tiles load one at a time.
template <int P>
__device__ __forceinline__ int32_t wave_extend_pre(const uint8_t *__restrict__ ref, uint32_t rlen,
                                                   const uint8_t *__restrict__ qry, uint32_t qlen,
                                                   const int32_t *s_sub, int32_t xdrop, uint32_t R, uint32_t Q,
                                                   bool left) {
    const uint32_t lane = lane_id();
    auto cell = [&](int64_t tile, uint8_t &rb, uint8_t &qb) -> bool {
        int64_t off = tile + lane + (left ? 1 : 0);
        bool v = left ? ((int64_t)R >= off && (int64_t)Q >= off)
                      : ((int64_t)R + off < (int64_t)rlen && (int64_t)Q + off < (int64_t)qlen);
        rb = v ? ref[left ? R - off : R + off] : 0;
        qb = v ? qry[left ? Q - off : Q + off] : 0;
        return v;
    };
    auto at_edge = [&](int64_t tile) -> bool {
        int64_t last = left ? tile + WAVE : tile + WAVE - 1;
        return left ? ((int64_t)R < last || (int64_t)Q < last)
                    : ((int64_t)R + last >= (int64_t)rlen || (int64_t)Q + last >= (int64_t)qlen);
    };
    uint8_t rb[P], qb[P];
    bool v[P];
#pragma unroll
    for (int t = 0; t < P; t++) v[t] = cell((int64_t)t * WAVE, rb[t], qb[t]);
    int32_t ps = 0, pm = 0;
#pragma unroll
    for (int t = 0; t < P; t++) {
        int32_t sc = v[t] ? s_sub[rb[t] * 8 + qb[t]] : 0;
        if (wave_tile(sc, at_edge((int64_t)t * WAVE), xdrop, ps, pm)) return pm;
    }
    for (int64_t tile = (int64_t)P * WAVE;; tile += WAVE) {
        uint8_t r1, q1;
        bool v1 = cell(tile, r1, q1);
        int32_t sc = v1 ? s_sub[r1 * 8 + q1] : 0;
        if (wave_tile(sc, at_edge(tile), xdrop, ps, pm)) return pm;
    }
}

// Exact gate for deferred hits (one wave per record). With arena_coords the
// extensions run in arena coordinates: every block is fenced by separator
// codes that force a drop under this matrix, so the maxima equal the
// block-local ones and no block lookup is needed; otherwise block-local bounds
// come from the hit's batch. The right extension preloads 2 tiles, the left 3
// (a hit is deferred when its right side outlives 64 cells or its left 128).
__global__ __launch_bounds__(256) void k_bgate_wave(const uint8_t *__restrict__ ref, uint32_t ref_len,
                                                    const uint8_t *__restrict__ arena, uint32_t arena_len,
                                                    int arena_coords, BlockOf bo,
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
        uint32_t q0 = 0, qlen = arena_len;
        if (!arena_coords) {
            uint32_t b = bo.batch_of(rec.seed);
            uint32_t blk = bo.batches[b].block;
            q0 = bo.blk_off[blk];
            qlen = bo.blk_len[blk];
        }
        int32_t mr = wave_extend_pre<2>(ref, ref_len, arena + q0, qlen, s_sub, xdrop, rec.R, rec.Q - q0, false);
        int32_t ml = wave_extend_pre<3>(ref, ref_len, arena + q0, qlen, s_sub, xdrop, rec.R, rec.Q - q0, true);
        bool pass = gate_keep((int32_t)((uint32_t)mr + (uint32_t)ml), thresh);
        if (pass && lane_id() == 0) {
            uint32_t at = atomicAdd(surv_n, 1u);
            if (at < cap_surv) surv[at] = rec;
            else atomicOr(err, 8u);
        }
    }
}

// find_hsps for the batched survivors (one wave per record, or per listed
// record: the ones k_bmaterialize_lane handed on): exact extension
// with positions in the chunk-local block, entropy, keep test. Output record
// i: HSP in block coordinates, its scope key (batch * MAX_CHUNKS + chunk) and
// the keep flag.
__global__ __launch_bounds__(256) void k_bmaterialize(
    const uint8_t *__restrict__ ref, uint32_t ref_len, const uint8_t *__restrict__ arena, BlockOf bo,
    const uint32_t *__restrict__ starts, const uint32_t *__restrict__ nchunks,
    const int32_t *__restrict__ sub, int32_t xdrop, int32_t thresh, int32_t noentropy, double log4,
    const SurvRec *__restrict__ surv, const uint32_t *__restrict__ surv_n, uint32_t cap_surv,
    SegmentPair *__restrict__ hsp, uint32_t *__restrict__ scope, uint8_t *__restrict__ keep,
    const uint32_t *__restrict__ ids, const uint32_t *__restrict__ ids_n, uint32_t cap_ids) {
    __shared__ int32_t s_sub[64];
    if (threadIdx.x < 64) s_sub[threadIdx.x] = sub[threadIdx.x];
    __syncthreads();
    const uint32_t lane = lane_id();
    // every survivor, or (ids) the survivors listed there
    const uint32_t n = ids ? min(*ids_n, cap_ids) : min(*surv_n, cap_surv);
    const uint32_t nwaves = gridDim.x * (blockDim.x / WAVE);
    for (uint32_t j = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; j < n; j += nwaves) {
        const uint32_t i = ids ? ids[j] : j;
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

// Sixteen byte codes from p on (p may lie a few bytes outside the sequence:
// the reference and the arena are padded with code 7 on both sides) as four
// words, word g byte j = the code at p + 4 g + j, from five aligned loads.
__device__ __forceinline__ void bytes16(const uint8_t *__restrict__ base, int64_t p, uint32_t (&v)[4]) {
    const uintptr_t u = (uintptr_t)(base + p);
    const uint32_t *w = (const uint32_t *)(u & ~(uintptr_t)3);
    const uint32_t sh = (uint32_t)(u & 3);
    uint32_t x[5];
#pragma unroll
    for (int g = 0; g < 5; g++) x[g] = w[g];
#pragma unroll
    for (int g = 0; g < 4; g++) v[g] = __builtin_amdgcn_alignbyte(x[g + 1], x[g], sh);
}

// Cells one lane extends per side before it hands the survivor to the wave
// materializer instead.
constexpr uint32_t LANE_EXTEND_CAP = 1024;
// Survivors a wave of the lane materializer claims at a time.
constexpr uint32_t LANE_CLAIM = 64;

// find_hsps for the batched survivors, one lane per survivor: the exact
// X-drop extension of wave_extend<true> written as the scalar loop it scans
// (kernels.rs:1279-1543: running sum, maximum with its earliest position, stop
// at the first cell more than xdrop below the maximum or at a sequence end),
// sixteen cells per iteration, then the entropy counts over the extent and the
// same f64 keep test and record as k_bmaterialize. Waves claim LANE_CLAIM
// survivors at a time and hand the next one to a lane as soon as the lane
// finishes (right side, left side, entropy counts), so lanes with short
// extensions do not wait for long ones. A side longer than LANE_EXTEND_CAP
// cells sends the survivor to `ids` for the wave-per-record kernel (which
// writes its keep flag; the caller zero-filled keep).
__global__ __launch_bounds__(256) void k_bmaterialize_lane(
    const uint8_t *__restrict__ ref, uint32_t ref_len, const uint8_t *__restrict__ arena, BlockOf bo,
    const uint32_t *__restrict__ starts, const uint32_t *__restrict__ nchunks,
    const int32_t *__restrict__ sub, int32_t xdrop, int32_t thresh, int32_t noentropy, double log4,
    const SurvRec *__restrict__ surv, const uint32_t *__restrict__ surv_n, uint32_t cap_surv,
    SegmentPair *__restrict__ hsp, uint32_t *__restrict__ scope, uint8_t *__restrict__ keep,
    uint32_t *__restrict__ ids, uint32_t *__restrict__ ids_n, uint32_t cap_ids, uint32_t *__restrict__ work,
    uint32_t *__restrict__ err) {
    __shared__ int32_t s_sub[64];
    if (threadIdx.x < 64) s_sub[threadIdx.x] = sub[threadIdx.x];
    __syncthreads();
    const uint32_t n = min(*surv_n, cap_surv);
    uint32_t c_cur = 0, c_end = 0;
    bool dry = false;
    bool act = false;
    uint32_t i = 0, b = 0, R = 0, Q = 0, k = 0, lim = 0, phase = 0;
    const uint8_t *qry = arena;
    int32_t score = 0, best = 0, bpos = 0, Mr = 0, rpos = 0, total = 0, extent = 0;
    uint32_t cnt0 = 0, cnt1 = 0, cnt2 = 0, cnt3 = 0;
    for (;;) {
        // lanes without a survivor take the next ones of the wave's claims
        uint64_t need = __ballot(!act);
        while (need) {
            if (c_cur >= c_end) {
                if (dry) break;
                uint32_t v = 0;
                if (lane_id() == (uint32_t)(__ffsll((unsigned long long)__ballot(true)) - 1))
                    v = atomicAdd(work, LANE_CLAIM);
                v = __builtin_amdgcn_readfirstlane(__shfl(v, (int)(__ffsll((unsigned long long)__ballot(true)) - 1), WAVE));
                if (v >= n) {
                    dry = true;
                    break;
                }
                c_cur = v;
                c_end = min(v + LANE_CLAIM, n);
            }
            const uint32_t rank = mbcnt(need);
            if (!act && rank < c_end - c_cur) {
                i = c_cur + rank;
                const SurvRec rec = surv[i];
                b = bo.batch_of(rec.seed);
                const uint32_t blk = bo.batches[b].block, q0 = bo.blk_off[blk], qlen = bo.blk_len[blk];
                qry = arena + q0;
                R = rec.R;
                Q = rec.Q - q0;
                phase = 0;
                k = 0;
                score = best = 0;
                bpos = -1;
                lim = min(ref_len - R, qlen - Q);
                act = true;
            }
            c_cur += min((uint32_t)__popcll(need), c_end - c_cur);
            need = __ballot(!act);
        }
        if (__ballot(act) == 0) break;
        if (!act) continue;
        bool fin = false;
        if (phase < 2) {
            // right (phase 0): cells R + k, Q + k; left (phase 1): R - 1 - k, Q - 1 - k
            const bool lft = phase == 1;
            uint32_t rw[4], qw[4];
            bytes16(ref, lft ? (int64_t)R - 16 - k : (int64_t)R + k, rw);
            bytes16(qry, lft ? (int64_t)Q - 16 - k : (int64_t)Q + k, qw);
            bool stop = false;
#pragma unroll
            for (int g = 0; g < 4; g++) {
                // left: word 3 - g byte-swapped holds cells 4 g .. 4 g + 3 in walk order
                const uint32_t r4 = lft ? __builtin_bswap32(rw[3 - g]) : rw[g];
                const uint32_t q4 = lft ? __builtin_bswap32(qw[3 - g]) : qw[g];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const uint32_t cell = k + 4 * g + j;
                    if (!stop && cell < lim) {
                        score += s_sub[((r4 >> (8 * j)) & 0xFFu) * 8 + ((q4 >> (8 * j)) & 0xFFu)];
                        const int32_t pos = (int32_t)cell + (lft ? 1 : 0);
                        bpos = score > best ? pos : bpos;
                        best = max(best, score);
                        stop = (int32_t)((uint32_t)best - (uint32_t)score) > xdrop;
                    } else {
                        stop = true;  // a sequence end: the cells past it score 0 and change nothing
                    }
                }
            }
            k += 16;
            if (!stop && k >= LANE_EXTEND_CAP) {  // too long for a lane: the wave materializer takes it
                const uint32_t at = atomicAdd(ids_n, 1u);
                if (at < cap_ids) ids[at] = i;
                else atomicOr(err, 8u);
                act = false;
            } else if (stop && !lft) {
                Mr = best;
                rpos = bpos;
                phase = 1;
                k = 0;
                score = best = bpos = 0;
                lim = min(R, Q);
            } else if (stop) {
                total = (int32_t)((uint32_t)Mr + (uint32_t)best);
                extent = (int32_t)((uint32_t)rpos + (uint32_t)bpos);
                rpos = bpos;  // from here: the left extent
                if (total >= thresh && (int64_t)total <= 3 * (int64_t)thresh && noentropy == 0) {
                    phase = 2;
                    k = 0;
                    lim = (uint32_t)(extent + 1);
                    cnt0 = cnt1 = cnt2 = cnt3 = 0;
                } else {
                    fin = true;
                }
            }
        } else {
            // entropy counts over positions 0..extent from R - left extent
            const uint32_t rs = R - (uint32_t)rpos, qs = Q - (uint32_t)rpos;
            uint32_t rw[4], qw[4];
            bytes16(ref, (int64_t)rs + k, rw);
            bytes16(qry, (int64_t)qs + k, qw);
#pragma unroll
            for (int g = 0; g < 4; g++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const uint32_t r = (rw[g] >> (8 * j)) & 0xFFu, q = (qw[g] >> (8 * j)) & 0xFFu;
                    const bool m = k + 4 * g + j < lim && r == q;
                    cnt0 += m && r == 0 ? 1u : 0u;
                    cnt1 += m && r == 1 ? 1u : 0u;
                    cnt2 += m && r == 2 ? 1u : 0u;
                    cnt3 += m && r == 3 ? 1u : 0u;
                }
            k += 16;
            fin = k >= lim;
        }
        if (fin) {
            double entropy = 1.0;
            if (phase == 2 && cnt0 + cnt1 + cnt2 + cnt3 >= 20) {
                const uint32_t cnt[4] = {cnt0, cnt1, cnt2, cnt3};
                const double denom = (double)(extent + 1);
                double acc = 0.0;
                for (int m = 0; m < 4; m++) {
                    double pb = (double)cnt[m] / denom;
                    double l = cnt[m] != 0 ? ref_ln(pb) : 0.0;
                    double t = pb * l;
                    acc = acc + t;
                }
                entropy = -acc / log4;
            }
            const uint32_t leftx = (uint32_t)rpos;
            const double gated = (double)(float)total * entropy;
            const bool kept = trunc_sat(gated) >= thresh;
            if (kept) {
                SegmentPair o;
                o.ref_start = R - leftx;
                o.query_start = Q - leftx;
                o.len = (uint32_t)extent;
                o.score = entropy > 0.0 ? trunc_sat((double)total * entropy) : 0;
                hsp[i] = o;
                const uint32_t nc = nchunks[b];
                const uint32_t ch = nc > 1 ? find_le(starts + (uint64_t)b * MAX_CHUNKS, nc, surv[i].seed) : 0;
                scope[i] = b * MAX_CHUNKS + ch;
            }
            keep[i] = kept ? 1 : 0;
            act = false;
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

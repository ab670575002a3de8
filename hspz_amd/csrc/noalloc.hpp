// noalloc.hpp — hspz_seed_and_filter_noalloc: the plain per-batch pipeline of
// hspz_seed_and_filter over a file-scope device scratch array instead of the
// engine's grow-only buffers, so that a call allocates nothing (no hipMalloc,
// no hipFree, no allocating library routine). Included by engine.hip, which
// owns the engine struct; the array and the lookup of its address live here,
// in the same source file.
//
// Scratch: SCRATCH_BYTES per device, held from module load until the process
// ends. Every call on a device shares that device's array; calls that overlap
// on one device race. The array keeps the last call's contents, so the
// counters a call needs at zero are cleared by a fill kernel (never by
// hipMemsetAsync) before use.
//
// Layout of one call (n = seeds in the slot, T = hits per tile, U = survivors
// of a max_hits chunk): counters and chunk limits | hit counts u32[n] | scan
// tiles | tile: anchors u64[T], flags u8[T], deferred ids u32[T], survivor ids
// u32[T] | survivors: anchors u64[U], ids u32[U], HSPs 16 B[U], keep u8[U],
// kept ids u32[U], kept HSPs 16 B[U]. The hits of a max_hits chunk are
// expanded and gated T at a time; each tile's survivors are appended (their
// anchors copied) to the chunk's survivor list, which the materializer then
// extends exactly. Survivors keep hit order, and dedup_and_order gives the
// same lists for any order, so the output equals hspz_seed_and_filter's.
#pragma once

namespace {

constexpr size_t SCRATCH_BYTES = (size_t)256 << 20;
constexpr uint32_t NA_MAX_CHUNKS = 256;  // max_hits chunks of one call

}  // namespace

// The scratch array. Host code never uses this name as an address: it takes
// the per-device address from hipGetSymbolAddress (scratch_ptr).
static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

namespace {

// Per-device address of `scratch`, looked up once per device with that device
// current (each device holds its own copy at its own address).
std::mutex &scratch_mutex() {
    static std::mutex mu;
    return mu;
}

int scratch_ptr(int device, uint8_t **out) {
    static uint8_t *ptrs[256] = {nullptr};
    if (device < 0 || device >= 256) return set_error(HSPZ_E_ARG, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(scratch_mutex());
    if (!ptrs[device]) {
        int cur = 0;
        HIP_TRY(hipGetDevice(&cur));
        HIP_TRY(hipSetDevice(device));
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return set_error(-(int)e, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        ptrs[device] = (uint8_t *)p;
    }
    *out = ptrs[device];
    return 0;
}

// gpu/mod.rs:2397-2420 (the KegAlign lower_bound walk) on the device, over the
// inclusive hit-count prefix of n seeds: out[0] = number of chunks, or
// 0xFFFFFFFF on the reference's underflow, 0xFFFFFFFE beyond `cap` chunks;
// out[1 + i] = last seed of chunk i. One thread.
__global__ void k_chunk_walk(const uint32_t *__restrict__ prefix, uint32_t n, uint32_t max_hits, uint32_t cap,
                             uint32_t *__restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const uint32_t num_iter = prefix[n - 1] / max_hits + 1;
    if (num_iter > cap) {
        out[0] = 0xFFFFFFFEu;
        return;
    }
    uint32_t lim = max_hits;
    for (uint32_t it = 0; it + 1 < num_iter; it++) {
        uint32_t lo = 0, hi = n;  // first p with prefix[p] >= lim
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (prefix[mid] < lim) lo = mid + 1;
            else hi = mid;
        }
        if (lo == 0) {
            out[0] = 0xFFFFFFFFu;
            return;
        }
        out[1 + it] = lo - 1;
        lim = prefix[lo - 1] + max_hits;
    }
    out[num_iter] = n - 1;
    out[0] = num_iter;
}

// Hits [t0, t1) of the chunk whose seeds are [s0, s0 + nseeds) and whose first
// hit is start_hit, in the plain path's anchor layout (k_find_hits_wave:
// prefix - 1 - start_hit - j for hit j of a seed), written to tile[a - t0].
// One wave per seed; seeds whose hits lie outside the tile are skipped.
__global__ __launch_bounds__(256) void k_find_hits_tile(const uint32_t *__restrict__ index,
                                                        const uint32_t *__restrict__ pos,
                                                        const uint64_t *__restrict__ seeds, uint32_t seed_size,
                                                        const uint32_t *__restrict__ prefix,
                                                        uint32_t start_seed, uint32_t start_hit, uint32_t nseeds,
                                                        uint32_t t0, uint32_t t1, uint64_t *__restrict__ tile) {
    const uint32_t lane = lane_id();
    const uint32_t nwaves = gridDim.x * (blockDim.x / WAVE);
    for (uint32_t s = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; s < nseeds; s += nwaves) {
        const uint32_t top = prefix[start_seed + s] - start_hit;  // one past this seed's highest anchor index
        const uint64_t sd = seeds[start_seed + s];
        const uint32_t k = (uint32_t)(sd >> 32);
        const uint32_t end = index[k], start = k ? index[k - 1] : 0;
        const uint32_t cnt = end - start;
        // anchor indices of this seed: [top - cnt, top)
        if (cnt == 0 || top <= t0 || top - cnt >= t1) continue;
        const uint64_t q = (uint64_t)((uint32_t)sd + seed_size) << 32;
        // j such that a = top - 1 - j lies in [t0, t1)
        const uint32_t j_lo = top > t1 ? top - t1 : 0;
        const uint32_t j_hi = min(cnt, top - t0);
        for (uint32_t j = j_lo + lane; j < j_hi; j += WAVE)
            tile[top - 1 - j - t0] = (uint64_t)(pos[start + j] + seed_size) | q;
    }
}

// out[i] = in[ids[i]] for i < *n.
__global__ void k_gather_u64(const uint64_t *__restrict__ in, const uint32_t *__restrict__ ids,
                             const uint32_t *__restrict__ n_ptr, uint64_t *__restrict__ out) {
    const uint32_t n = *n_ptr;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = in[ids[i]];
}

__global__ void k_iota_u32(uint32_t *__restrict__ p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = i;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Counter slots at the start of the scratch array.
enum { NA_TOTAL = 0, NA_SLOW = 1, NA_COUNT = 2, NA_N = 64 };

struct NoallocLayout {
    size_t hit_num, tiles, anchors, flags, slow, ids, sanchors, sids, hsp, done, kids, reduced, walk, end;
    uint32_t tile_hits, surv_cap;
};

// Places the regions of one call with n seeds: after the counters, the chunk
// limits and the hit counts, 1 MiB holds the scan tiles of every scan of the
// call; of the rest, the tile gets three quarters and the survivor lists a
// quarter (tile_cap, HSPZ_FLAG_NOALLOC_TILE_HITS, narrows the tile for tests).
// False when n seeds leave too little.
bool noalloc_layout(uint64_t n, uint32_t tile_cap, NoallocLayout *L) {
    size_t at = align256(NA_N * 4);
    L->walk = at;
    at = align256(at + (NA_MAX_CHUNKS + 1) * 4);
    L->hit_num = at;
    at = align256(at + n * 4);
    L->tiles = at;
    at += 1u << 20;
    if (at + (64u << 20) > SCRATCH_BYTES) return false;
    const size_t rest = SCRATCH_BYTES - at;
    // per tile hit: anchor 8, flag 1, deferred id 4, survivor id 4 (+ slack)
    uint64_t t = (rest / 4 * 3) / 18;
    // per survivor: anchor 8, id 4, HSP 16, keep 1, kept id 4, kept HSP 16 (+ slack)
    uint64_t u = (rest / 4) / 50;
    if (tile_cap) t = std::min<uint64_t>(t, tile_cap);
    t &= ~(uint64_t)63;
    u &= ~(uint64_t)63;
    // the scans of the call (seeds, tile flags, survivor flags) fit the 1 MiB of tiles
    const uint64_t scan_n = std::max<uint64_t>(std::max<uint64_t>(n, t), u);
    if (((scan_n + TILE - 1) / TILE + 2) * 4 > (1u << 20)) return false;
    L->tile_hits = (uint32_t)t;
    L->surv_cap = (uint32_t)u;
    L->anchors = at;
    at = align256(at + t * 8);
    L->flags = at;
    at = align256(at + t + 16);
    L->slow = at;
    at = align256(at + t * 4);
    L->ids = at;
    at = align256(at + t * 4);
    L->sanchors = at;
    at = align256(at + u * 8);
    L->sids = at;
    at = align256(at + u * 4);
    L->hsp = at;
    at = align256(at + u * 16);
    L->done = at;
    at = align256(at + u + 16);
    L->kids = at;
    at = align256(at + u * 4);
    L->reduced = at;
    at = align256(at + u * 16);
    L->end = at;
    return L->end <= SCRATCH_BYTES && t >= 64 && u >= 64;
}

// Stable compaction of nonzero flags[0..n) into ids over scratch tiles; the
// count lands in *count_dev and is read back (one blocking round trip).
int noalloc_compact(hspz_engine *h, const uint8_t *flags, uint32_t n, uint32_t *tiles, uint32_t *ids,
                    uint32_t *count_dev, uint32_t *count) {
    uint32_t nt = std::max<uint32_t>(ceil_div(n, TILE), 1);
    hipLaunchKernelGGL(k_tile_count_u8, dim3(nt), dim3(TILE_THREADS), 0, h->stream, flags, n, tiles);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_THREADS), 0, h->stream, tiles, nt, count_dev);
    hipLaunchKernelGGL(k_tile_emit_u8, dim3(nt), dim3(TILE_THREADS), 0, h->stream, flags, n, tiles, ids);
    h->launches += 3;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(h->h_small, count_dev, 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->pipeline_syncs++;
    *count = h->h_small[0];
    return 0;
}

// hspz_seed_and_filter over the scratch array; kept HSPs of every max_hits
// chunk land in out[0..cap) (each chunk's pre-dedup list must fit there, after
// the lists kept before it).
int seed_and_filter_noalloc(hspz_engine *h, int slot, int rev, hspz_segment_pair *out, size_t cap,
                            size_t *n_out, uint32_t *num_hits_out, uint32_t *raw_out) {
    HIP_TRY(hipSetDevice(h->device));
    if (int rc = join_seed_copy(h, slot)) return rc;
    uint8_t *base = h->scratch;
    if (!base) return set_error(HSPZ_E_ARG, "engine has no scratch address");
    *n_out = 0;
    *num_hits_out = 0;
    *raw_out = 0;
    const uint32_t n = h->seed_len[slot];
    if (n == 0) return 0;
    NoallocLayout L;
    if (!noalloc_layout(n, h->noalloc_tile_hits, &L))
        return set_error(HSPZ_E_OOM, "%u seeds need more than the %zu-byte scratch array holds", n, SCRATCH_BYTES);
    uint32_t *ctr = (uint32_t *)base;
    uint32_t *walk = (uint32_t *)(base + L.walk);
    uint32_t *hit_num = (uint32_t *)(base + L.hit_num), *tiles = (uint32_t *)(base + L.tiles);
    uint64_t *anchors = (uint64_t *)(base + L.anchors), *sanchors = (uint64_t *)(base + L.sanchors);
    uint8_t *flags = base + L.flags, *done = base + L.done;
    uint32_t *slow = (uint32_t *)(base + L.slow), *ids = (uint32_t *)(base + L.ids);
    uint32_t *sids = (uint32_t *)(base + L.sids), *kids = (uint32_t *)(base + L.kids);
    SegmentPair *hsp = (SegmentPair *)(base + L.hsp), *reduced = (SegmentPair *)(base + L.reduced);

    const uint8_t *q = (rev ? h->qr : h->qf).as<uint8_t>() + PAD;
    const uint8_t *r = h->ref.as<uint8_t>() + PAD;
    const uint64_t *seeds = h->seeds[slot].as<uint64_t>();
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, h->stream, ctr, (uint32_t)NA_N, 0u);
    hipLaunchKernelGGL(k_find_num_hits, dim3(std::min<uint32_t>(ceil_div(n, 256), 4096)), dim3(256), 0, h->stream,
                       seeds, n, h->index.as<uint32_t>(), hit_num);
    const uint32_t nt = std::max<uint32_t>(ceil_div(n, TILE), 1);
    hipLaunchKernelGGL(k_tile_sum_u32, dim3(nt), dim3(TILE_THREADS), 0, h->stream, hit_num, n, tiles);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_THREADS), 0, h->stream, tiles, nt, ctr + NA_TOTAL);
    hipLaunchKernelGGL(k_tile_apply_u32, dim3(nt), dim3(TILE_THREADS), 0, h->stream, hit_num, n, tiles);
    h->launches += 5;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(h->h_small, ctr + NA_TOTAL, 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->pipeline_syncs++;
    const uint32_t num_hits = h->h_small[0];
    *num_hits_out = num_hits;
    if (num_hits == 0) return 0;

    // max_hits chunks: (first seed, last seed) pairs, on the device when split
    uint32_t nchunks = 1;
    uint32_t last[NA_MAX_CHUNKS];
    last[0] = n - 1;
    if (num_hits > h->max_hits) {
        hipLaunchKernelGGL(k_chunk_walk, dim3(1), dim3(64), 0, h->stream, hit_num, n, h->max_hits, NA_MAX_CHUNKS,
                           walk);
        h->launches++;
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(h->h_small, walk, (NA_MAX_CHUNKS + 1) * 4, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        h->pipeline_syncs++;
        const uint32_t w0 = h->h_small[0];
        if (w0 == 0xFFFFFFFFu) return set_error(HSPZ_E_CHUNK, "a single seed exceeds max_hits %u", h->max_hits);
        if (w0 == 0xFFFFFFFEu || w0 > NA_MAX_CHUNKS)
            return set_error(HSPZ_E_OOM, "more than %u max_hits chunks in one call", NA_MAX_CHUNKS);
        nchunks = w0;
        for (uint32_t i = 0; i < nchunks; i++) last[i] = h->h_small[1 + i];
    }
    // the chunk's first hit = the prefix at the previous chunk's last seed
    uint32_t prev_hits[NA_MAX_CHUNKS];
    if (nchunks > 1) {
        for (uint32_t i = 0; i + 1 < nchunks; i++)
            HIP_TRY(hipMemcpyAsync(h->h_small + i, hit_num + last[i], 4, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        h->pipeline_syncs++;
        for (uint32_t i = 0; i + 1 < nchunks; i++) prev_hits[i] = h->h_small[i];
    }
    size_t at = 0;
    uint32_t s0 = 0, h0 = 0;
    for (uint32_t c = 0; c < nchunks; c++) {
        const uint32_t s1 = last[c], h1 = c + 1 < nchunks ? prev_hits[c] : num_hits;
        const uint32_t nseeds = s1 + 1 - s0, nh = h1 - h0;
        uint32_t nsurv = 0;
        for (uint32_t t0 = 0; t0 < nh; t0 += L.tile_hits) {
            const uint32_t t1 = std::min<uint64_t>((uint64_t)t0 + L.tile_hits, nh), tn = t1 - t0;
            hipLaunchKernelGGL(k_find_hits_tile, dim3(std::min<uint32_t>(ceil_div(nseeds, 4), 16384)), dim3(256), 0,
                               h->stream, h->index.as<uint32_t>(), h->pos.as<uint32_t>(), seeds, h->seed_size,
                               hit_num, s0, h0, nseeds, t0, t1, anchors);
            h->launches++;
            if (h->fast_gate) {
                hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, h->stream, ctr + NA_SLOW, 1u, 0u);
                const uint32_t *q2 = (rev ? h->qr2 : h->qf2).as<uint32_t>();
                const uint32_t *qx = (rev ? h->qrx : h->qfx).as<uint32_t>();
                uint32_t gblocks = std::max<uint32_t>(1, std::min<uint32_t>(ceil_div(tn, GATE_CHUNK * 8), h->gate_blocks));
                hipLaunchKernelGGL(k_gate_packed, dim3(gblocks), dim3(512), 0, h->stream, h->ref2.as<uint32_t>(),
                                   h->refx.as<uint32_t>(), h->ref_len, q2, qx, h->query_len,
                                   h->pairtab.as<uint32_t>(), h->xdrop, h->hspthresh, anchors, tn, flags, slow,
                                   ctr + NA_SLOW);
                uint32_t sblocks = std::max<uint32_t>(1, std::min<uint32_t>(ceil_div(tn, 64), 2048));
                hipLaunchKernelGGL(k_gate_wave, dim3(sblocks), dim3(256), 0, h->stream, r, h->ref_len, q,
                                   h->query_len, h->sub.as<int32_t>(), h->xdrop, h->hspthresh, anchors, flags, slow,
                                   ctr + NA_SLOW);
                h->launches += 3;
            } else {
                uint32_t gblocks = std::max<uint32_t>(1, std::min<uint32_t>(ceil_div(tn, GATE_CHUNK * 4), h->gate_blocks));
                hipLaunchKernelGGL(k_gate, dim3(gblocks), dim3(256), 0, h->stream, r, h->ref_len, q, h->query_len,
                                   h->sub.as<int32_t>(), h->xdrop, h->hspthresh, anchors, tn, flags,
                                   (const uint32_t *)nullptr, (const uint32_t *)nullptr);
                h->launches++;
            }
            HIP_TRY(hipGetLastError());
            uint32_t ns = 0;
            if (int rc = noalloc_compact(h, flags, tn, tiles, ids, ctr + NA_COUNT, &ns)) return rc;
            h->gate_hits += tn;
            if (ns == 0) continue;
            if ((uint64_t)nsurv + ns > L.surv_cap)
                return set_error(HSPZ_E_OOM, "a max_hits chunk has more than %u gate survivors (scratch limit)",
                                 L.surv_cap);
            hipLaunchKernelGGL(k_gather_u64, dim3(std::min<uint32_t>(ceil_div(ns, 256), 4096)), dim3(256), 0,
                               h->stream, anchors, ids, ctr + NA_COUNT, sanchors + nsurv);
            h->launches++;
            nsurv += ns;
        }
        h->gate_launches++;
        if (nsurv) {
            hipLaunchKernelGGL(k_iota_u32, dim3(std::min<uint32_t>(ceil_div(nsurv, 256), 4096)), dim3(256), 0,
                               h->stream, sids, nsurv);
            hipLaunchKernelGGL(k_materialize, dim3(std::min<uint32_t>(ceil_div(nsurv, 4), 16384)), dim3(256), 0,
                               h->stream, r, h->ref_len, q, h->query_len, h->sub.as<int32_t>(), h->xdrop,
                               h->hspthresh, h->noentropy, 1.3862943649291992, sanchors, sids, nsurv, hsp, done);
            h->launches += 2;
            HIP_TRY(hipGetLastError());
            uint32_t nraw = 0;
            if (int rc = noalloc_compact(h, done, nsurv, tiles, kids, ctr + NA_COUNT, &nraw)) return rc;
            *raw_out += nraw;
            if (nraw) {
                if (at + nraw > cap)
                    return set_error(HSPZ_E_OOM, "output holds %zu records; %zu are needed before dedup", cap,
                                     at + nraw);
                hipLaunchKernelGGL(k_gather_hsp, dim3(ceil_div(nraw, 256)), dim3(256), 0, h->stream, hsp, kids,
                                   nraw, reduced);
                h->launches++;
                HIP_TRY(hipGetLastError());
                HIP_TRY(hipMemcpyAsync(out + at, reduced, (size_t)nraw * 16, hipMemcpyDeviceToHost, h->stream));
                HIP_TRY(hipStreamSynchronize(h->stream));
                h->pipeline_syncs++;
                at += dedup_and_order(out + at, nraw);
            }
        }
        s0 = s1 + 1;
        h0 = h1;
    }
    *n_out = at;
    return 0;
}

}  // namespace

extern "C" int hspz_seed_and_filter_noalloc(hspz_engine *h, int slot, int rev, hspz_segment_pair *out, size_t cap,
                                            size_t *n_out, uint32_t *num_hits, uint32_t *raw_hsps) {
    clear_error();
    if (!h || !n_out || slot < 0 || slot > 1 || (cap && !out)) return set_error(HSPZ_E_ARG, "bad argument");
    *n_out = 0;
    if (num_hits) *num_hits = 0;
    if (raw_hsps) *raw_hsps = 0;
    if (h->query_len == 0)
        return set_error(HSPZ_E_NO_QUERY, "Engine has no query: swap_query() must be called before seed_and_filter()");
    uint32_t nh = 0, raw = 0;
    int rc = seed_and_filter_noalloc(h, slot, rev, out, cap, n_out, &nh, &raw);
    if (rc == 0) rc = mark_slot_computed(h, slot);
    if (num_hits) *num_hits = nh;
    if (raw_hsps) *raw_hsps = raw;
    return rc;
}

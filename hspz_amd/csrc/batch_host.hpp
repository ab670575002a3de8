// batch_host.hpp — host side of hspz_filter_batch (included by engine.hip,
// which owns the engine struct and the buffer helpers).
//
// One call = one device pass over every batch of every chunk:
//   arena upload + 2-bit planes -> batch seeding (validity, stable compaction,
//   scatter) -> bucket sizes -> u64 inclusive scan -> per-batch max_hits
//   chunks -> fused hit expansion + packed gate -> exact gate for deferred
//   hits -> materializer -> compaction of kept HSPs with scope keys
// then one read-back, per-scope dedup_and_order on host threads, and the
// per-chunk, per-interval, fw-then-rc assembly of run.rs:278-299 / 443-451.
#pragma once

#include <chrono>
#include <thread>

#include "dedup_dev.hpp"

namespace {

// One seed-and-filter batch of a call (run.rs:278-299): query positions
// [lo, hi) of chunk `chunk`'s forward (rev = 0) or reverse-complement block,
// in interval `interval` of that chunk.
struct HostBatch {
    uint32_t chunk, interval;
    int rev;
    uint32_t lo, hi;
};

// Device buffers of the batched path, grow-only, owned by the engine.
struct BatchBufs {
    DevBuf arena, plane, coarse, desc, poff, blk_off, blk_len, valid, ids, voff, seeds, counts, prefix,
        tiles, nchunks, starts, bhits, surv, slow, mhsp, mscope, keep, kids, ohsp, oscope, ctr, gcnt, sq, skm, qrec,
        kvalo, ktmp, exact, kkey, kgaps, cand, cand2, cbins;
};

// Largest survivor or deferred-hit list of one pass, in records (16 B each).
constexpr uint32_t LIST_CAP_MAX = 1u << 30;

// Capacity after an overflow. The pass's list counter keeps counting past the
// capacity, so `count` is what the list needed, and a rerun of the same pass
// needs exactly that (the gate's decisions do not depend on the capacity):
// the larger of count + 1/4 and 1.5 x the old capacity, computed in 64 bits
// and clamped to LIST_CAP_MAX. False when the count itself is beyond
// LIST_CAP_MAX (growing could never hold it).
bool grow_list_cap(uint32_t cap, uint32_t count, uint32_t *out) {
    if (count > LIST_CAP_MAX) return false;
    uint64_t want = std::max<uint64_t>((uint64_t)count + count / 4, (uint64_t)cap + cap / 2);
    *out = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(want, 64), LIST_CAP_MAX);
    return *out >= count;
}

}  // namespace

// Test hook (not part of the ABI in include/hspz.h): the staged read's piece
// size, so that a GPU test covers many pieces with a small output.
extern "C" int hspz_test_set_read_piece(hspz_engine *h, size_t bytes) {
    if (!h || bytes < 4096 || bytes % 16) return 1;
    h->read_piece = bytes;
    return 0;
}

// Test hook (not part of the ABI in include/hspz.h): grow_list_cap for the
// CPU unit test of the growth arithmetic; 1 when the grown list holds count.
extern "C" int hspz_test_grow_list_cap(uint32_t cap, uint32_t count, uint32_t *out) {
    return grow_list_cap(cap, count, out) ? 1 : 0;
}

namespace {

// Device counters of one pass; CTR_WORK is the 64-bit seed-claim counter of
// the batched gate (two slots, 8-byte aligned).
// CTR_CAND counts the coarse gate's candidates and CTR_WORK2 (64-bit, slots
// 10-11) is the candidate-claim counter of k_bgate_exact. CTR_LONG counts the
// survivors the lane materializer hands to the wave materializer, CTR_MWORK
// its survivor-claim counter.
enum { CTR_NVALID = 0, CTR_SURV = 1, CTR_SLOW = 2, CTR_ERR = 3, CTR_KEPT = 4, CTR_EXACT = 5, CTR_WORK = 6,
       CTR_KGAPS = 8, CTR_CAND = 9, CTR_WORK2 = 10, CTR_LONG = 12, CTR_MWORK = 13, CTR_N = 14 };

typedef void (*BGateFn)(const uint32_t *, const uint32_t *, uint32_t, const uint32_t *, const uint32_t *, uint32_t,
                        const uint32_t *, int32_t, int32_t, const uint64_t *, const uint32_t *, uint32_t,
                        const uint32_t *, const uint32_t *, uint32_t, uint32_t, unsigned long long *, SurvRec *,
                        uint32_t *, uint32_t, SurvRec *, uint32_t *, uint32_t, uint32_t *);

typedef void (*KGateFn)(const uint32_t *, const uint32_t *, int32_t, int32_t, const QRec *, const uint32_t *,
                        const uint32_t *, const uint32_t *, uint32_t, const uint32_t *, const uint32_t *, const uint32_t *,
                        const uint4 *, uint32_t, uint32_t, unsigned long long *, SurvRec *, uint32_t *, uint32_t,
                        SurvRec *, uint32_t *, uint32_t, uint32_t *);

// The k-mer-major gate instantiation for a step width and block size; the
// merged steps take the cheapest score-sum form the engine's matrix allows.
template <int BT>
KGateFn kgate_kernel_bt(int step_cells, int sums) {
    switch (step_cells) {
    case 16: return k_bgate_kmer<1, BT>;
    case 64: return k_bgate_kmer<4, BT>;
    case 96:
        return sums == SUM_RAW ? k_bgate_kmer<6, BT, SUM_RAW>
               : sums == SUM_TREE16 ? k_bgate_kmer<6, BT, SUM_TREE16> : k_bgate_kmer<6, BT, SUM_TREE8>;
    case 128:
        return sums == SUM_RAW ? k_bgate_kmer<8, BT, SUM_RAW>
               : sums == SUM_TREE16 ? k_bgate_kmer<8, BT, SUM_TREE16> : k_bgate_kmer<8, BT, SUM_TREE8>;
    default: return k_bgate_kmer<2, BT>;
    }
}

KGateFn kgate_kernel(int step_cells, int threads, int sums) {
    return threads == 1024 ? kgate_kernel_bt<1024>(step_cells, sums) : kgate_kernel_bt<512>(step_cells, sums);
}

// The lookahead gate (k_bgate_ahead) applies to merged steps (96 or 128
// cells) in 512-thread blocks and needs the reference records; the other
// gates run a 96-cell step as 128.
bool lookahead_wanted(const hspz_engine *h) {
    return h->lookahead && h->bgate_step >= 96 && h->bgate_threads == 512;
}

// The batched gate instantiation for a step width and block size.
BGateFn bgate_kernel(int step_cells, int threads) {
    if (step_cells >= 96) step_cells = 64;  // the merged-phase variants exist in k-mer-major order only
    if (threads == 1024)
        return step_cells == 16 ? k_bgate_packed<1, 1024> : step_cells == 64 ? k_bgate_packed<4, 1024> : k_bgate_packed<2, 1024>;
    return step_cells == 16 ? k_bgate_packed<1, 512> : step_cells == 64 ? k_bgate_packed<4, 512> : k_bgate_packed<2, 512>;
}

}  // namespace

// Everything a captured device pass depends on: if a call's key equals the
// cached one, the instantiated graph is replayed instead of relaunching.
struct GraphKey {
    const void *ptrs[48];
    uint64_t vals[24];
    uint32_t shape[33];
};

struct hspz_batch_state {
    BatchBufs b;
    uint32_t cap_surv = 1u << 22, cap_slow = 1u << 22, cap_cand = 1u << 22;
    bool have_graph = false;
    GraphKey key;
    // before the grouping sort, after it, the gate, after the gate
    hipGraph_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
    hipGraphExec_t exec[4] = {nullptr, nullptr, nullptr, nullptr};
    uint64_t graph_captures = 0, graph_replays = 0;
    // device dedup_and_order scratch
    DevBuf dd_key_a, dd_key_b, dd_id_a, dd_id_b, dd_keep, dd_kids, dd_tmp;
};

namespace {

// Expected hits per query position and strand against the engine's index:
// per seeds, each meeting the mean bucket (indexed positions / 4^k; repeats
// only add to it).
double hits_per_position(const hspz_engine *h, uint32_t per) {
    return (double)per * (double)h->pos_len / (double)std::max<size_t>(h->index_len, 1);
}

// Builds the engine's reference records (k_ref_records) on first use, when a
// gate wants them (HSPZ_FLAG_REF_RECORDS, or a merged-step gate) and device
// memory leaves room for them and a margin; otherwise the k-mer-major gate
// reads its windows from the plane.
int ensure_ref_records(hspz_engine *h, bool want) {
    // (an engine whose target is shared reads the holder's records or none)
    if (!want || h->rrec.p || h->rrec_declined || h->pos_len == 0 || h->resident) return 0;
    size_t need = h->pos_len * 48;
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t));
    if (f < need + (8ull << 30)) {
        h->rrec_declined = true;
        return 0;
    }
    if (int rc = reserve(h, h->rrec, need, true)) return rc;
    hipLaunchKernelGGL(k_ref_records, dim3(std::min<uint32_t>(ceil_div(h->pos_len, 256), 65536)), dim3(256), 0,
                       h->stream, h->pos.as<uint32_t>(), (uint32_t)h->pos_len, h->clean.as<uint32_t>(),
                       h->ref2.as<uint32_t>(), h->seed_size, h->rrec.as<uint4>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->launches++;
    return 0;
}

// Stable radix sort of (key, u32 value) pairs on the engine's stream, over
// the key's low `bits` bits (the rest are zero).
template <class K>
int dd_sort(hspz_engine *h, DevBuf &tmp, const K *ki, K *ko, const uint32_t *vi, uint32_t *vo, uint32_t n,
            uint32_t bits = 8 * sizeof(K)) {
    size_t bytes = 0;
    HIP_TRY(rocprim::radix_sort_pairs(nullptr, bytes, ki, ko, vi, vo, n, 0, bits, h->stream));
    if (int rc = reserve(h, tmp, bytes + 256)) return rc;
    HIP_TRY(rocprim::radix_sort_pairs(tmp.p, bytes, ki, ko, vi, vo, n, 0, bits, h->stream));
    return 0;
}

// dedup_and_order of every scope of the pass's kept list (ohsp/oscope, n
// records) on the device; the result, grouped by ascending scope and in the
// reference's output order within each scope, lands in mhsp/mscope and its
// length in *kept. Scopes are below nscopes.
int device_dedup(hspz_engine *h, BatchBufs &B, uint32_t n, uint64_t nscopes, uint32_t *kept) {
    hspz_batch_state *bs = h->batch;
    int rc;
    if ((rc = reserve(h, bs->dd_key_a, (size_t)n * 8 + 8)) || (rc = reserve(h, bs->dd_key_b, (size_t)n * 8 + 8)) ||
        (rc = reserve(h, bs->dd_id_a, (size_t)n * 4 + 4)) || (rc = reserve(h, bs->dd_id_b, (size_t)n * 4 + 4)) ||
        (rc = reserve(h, bs->dd_keep, (size_t)n + 16)) || (rc = reserve(h, bs->dd_kids, (size_t)n * 4 + 4)))
        return rc;
    uint32_t nt = std::max<uint32_t>(ceil_div(n, TILE), 1);
    if ((rc = reserve(h, h->tiles, (size_t)nt * 4))) return rc;
    const SegmentPair *hsp = B.ohsp.as<SegmentPair>();
    const uint32_t *scope = B.oscope.as<uint32_t>();
    uint64_t *ka = bs->dd_key_a.as<uint64_t>(), *kb = bs->dd_key_b.as<uint64_t>();
    uint32_t *ia = bs->dd_id_a.as<uint32_t>(), *ib = bs->dd_id_b.as<uint32_t>();
    uint32_t *k32a = (uint32_t *)ka, *k32b = (uint32_t *)kb;
    auto grid = [](uint32_t m) { return dim3(std::max<uint32_t>(1, std::min<uint32_t>(ceil_div(m, 256), 4096))); };
    // (scope << 32 | low word) keys: only the scope's bits above the low word
    uint32_t sbits = 32;
    while (sbits < 64 && (nscopes - 1) >> (sbits - 32)) sbits++;
    // order 1 over all records: (scope, diagonal, ref_start, len, -score)
    hipLaunchKernelGGL(k_dd_key_len_score, grid(n), dim3(256), 0, h->stream, hsp, nullptr, n, ka, ia);
    if ((rc = dd_sort<uint64_t>(h, bs->dd_tmp, ka, kb, ia, ib, n))) return rc;
    hipLaunchKernelGGL(k_dd_key_ref, grid(n), dim3(256), 0, h->stream, hsp, ib, n, k32a);
    if ((rc = dd_sort<uint32_t>(h, bs->dd_tmp, k32a, k32b, ib, ia, n))) return rc;
    hipLaunchKernelGGL(k_dd_key_scope, grid(n), dim3(256), 0, h->stream, hsp, scope, ia, n, 1, ka);
    if ((rc = dd_sort<uint64_t>(h, bs->dd_tmp, ka, kb, ia, ib, n, sbits))) return rc;
    // containment against the previous record, compaction
    hipLaunchKernelGGL(k_dd_mark, grid(n), dim3(256), 0, h->stream, hsp, scope, ib, n, bs->dd_keep.as<uint8_t>());
    hipLaunchKernelGGL(k_tile_count_u8, dim3(nt), dim3(TILE_THREADS), 0, h->stream, bs->dd_keep.as<uint8_t>(), n,
                       h->tiles.as<uint32_t>());
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_THREADS), 0, h->stream, h->tiles.as<uint32_t>(), nt,
                       h->total.as<uint32_t>());
    hipLaunchKernelGGL(k_tile_emit_u8, dim3(nt), dim3(TILE_THREADS), 0, h->stream, bs->dd_keep.as<uint8_t>(), n,
                       h->tiles.as<uint32_t>(), bs->dd_kids.as<uint32_t>());
    HIP_TRY(hipGetLastError());
    uint32_t k = 0;
    if ((rc = read_total(h, &k))) return rc;
    if (k > n) return set_error(HSPZ_E_ARG, "device dedup kept %u of %u records", k, n);
    hipLaunchKernelGGL(k_dd_compose, grid(k), dim3(256), 0, h->stream, ib, bs->dd_kids.as<uint32_t>(), k, ia);
    // order 2 over the kept records: (scope, query_start, ref_start, len, -score)
    hipLaunchKernelGGL(k_dd_key_len_score, grid(k), dim3(256), 0, h->stream, hsp, ia, k, ka, ib);
    if ((rc = dd_sort<uint64_t>(h, bs->dd_tmp, ka, kb, ib, ia, k))) return rc;
    hipLaunchKernelGGL(k_dd_key_ref, grid(k), dim3(256), 0, h->stream, hsp, ia, k, k32a);
    if ((rc = dd_sort<uint32_t>(h, bs->dd_tmp, k32a, k32b, ia, ib, k))) return rc;
    hipLaunchKernelGGL(k_dd_key_scope, grid(k), dim3(256), 0, h->stream, hsp, scope, ib, k, 0, ka);
    if ((rc = dd_sort<uint64_t>(h, bs->dd_tmp, ka, kb, ib, ia, k, sbits))) return rc;
    hipLaunchKernelGGL(k_dd_gather, grid(k), dim3(256), 0, h->stream, hsp, scope, ia, k, B.mhsp.as<SegmentPair>(),
                       B.mscope.as<uint32_t>());
    HIP_TRY(hipGetLastError());
    h->launches += 17;
    *kept = k;
    return 0;
}

// One device pass over the batches hb (chunk indices into chunks[0..n_chunks),
// in output order): kept HSPs with their scope keys (local batch index *
// MAX_CHUNKS + max_hits chunk), grouped by ascending scope and deduplicated
// when *deduped comes back true.
// Copies n bytes of device memory to pageable host memory through two pinned
// staging halves: the DMA of one piece overlaps the host threads that copy
// the previous piece out (and take the destination's first-touch page faults
// in parallel). Small reads go directly.
int read_to_host(hspz_engine *h, void *dst, const void *src, size_t n) {
    const size_t READ_STAGE_BYTES = h->read_piece;  // bytes per staging half (64 MiB)
    if (n < READ_STAGE_BYTES / 8) {
        HIP_TRY(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        return 0;
    }
    if (h->read_stage && h->read_stage_piece != READ_STAGE_BYTES) {
        (void)hipHostFree(h->read_stage);
        h->read_stage = nullptr;
    }
    if (!h->read_stage) {
        if (hipHostMalloc((void **)&h->read_stage, 2 * READ_STAGE_BYTES, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            h->read_stage = nullptr;
            HIP_TRY(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, h->stream));
            HIP_TRY(hipStreamSynchronize(h->stream));
            return 0;
        }
        h->read_stage_piece = READ_STAGE_BYTES;
    }
    const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto copy_out = [&](size_t off, size_t len, const uint8_t *from) {
        std::vector<std::thread> th;
        const size_t per = (len + nth - 1) / nth;
        for (unsigned t = 0; t < nth && (size_t)t * per < len; t++) {
            const size_t a = (size_t)t * per, b = std::min(len, a + per);
            th.emplace_back([=] { memcpy((uint8_t *)dst + off + a, from + a, b - a); });
        }
        for (auto &x : th) x.join();
    };
    const size_t pieces = (n + READ_STAGE_BYTES - 1) / READ_STAGE_BYTES;
    for (size_t k = 0; k <= pieces; k++) {
        if (k < pieces) {  // this piece's DMA into its half (that half's last piece was copied out already)
            const size_t off = k * READ_STAGE_BYTES, len = std::min(READ_STAGE_BYTES, n - off);
            HIP_TRY(hipMemcpyAsync(h->read_stage + (k & 1) * READ_STAGE_BYTES, (const uint8_t *)src + off, len,
                                   hipMemcpyDeviceToHost, h->stream));
            HIP_TRY(hipEventRecord(h->ev_stage[k & 1], h->stream));
        }
        if (k > 0) {  // the previous piece has landed: out to the destination while this one copies
            const size_t j = k - 1, off = j * READ_STAGE_BYTES, len = std::min(READ_STAGE_BYTES, n - off);
            HIP_TRY(hipEventSynchronize(h->ev_stage[j & 1]));
            copy_out(off, len, h->read_stage + (j & 1) * READ_STAGE_BYTES);
        }
    }
    return 0;
}

// Where a device-deduplicated pass puts its lists: appended to a growing
// host array (already in the call's output order: scopes ascend in batch
// order, and a list's batches are consecutive), with each batch's record
// count in batch_counts. Passes deduplicated on the host fill hsps/scopes.
struct DirectOut {
    hspz_segment_pair *buf = nullptr;
    size_t len = 0, cap = 0;
    std::vector<uint32_t> batch_counts;
};

// Blocks of the candidate binning (histogram and scatter alike): one per CU,
// whose scatter tile (8 records per thread, 128 KiB) fills most of its LDS.
uint32_t cand_scatter_blocks(const hspz_engine *h) { return (uint32_t)std::max(h->num_cus, 1); }

int batch_pass(hspz_engine *h, const Shape &s, const hspz_pass_params *p,
               const hspz_query_chunk *chunks, size_t n_chunks, const std::vector<HostBatch> &hb,
               std::vector<hspz_segment_pair> &hsps, std::vector<uint32_t> &scopes,
               hspz_pass_stats *st, bool *deduped, bool use_graph, DirectOut *dout) {
    *deduped = false;
    BatchBufs &B = h->batch->b;
    auto t_begin = std::chrono::steady_clock::now();
    // ---- layout: blocks (chunk c fwd = 2c, rc = 2c + 1), each after QPAD code-7 bytes
    std::vector<uint32_t> blk_off(2 * n_chunks), blk_len(2 * n_chunks);
    uint64_t off = QPAD;
    for (size_t c = 0; c < n_chunks; c++)
        for (int r = 0; r < 2; r++) {
            blk_off[2 * c + r] = (uint32_t)off;
            blk_len[2 * c + r] = (uint32_t)chunks[c].len;
            off += chunks[c].len + QPAD;
        }
    if (off >= 0xFFFFFFFFull - 2 * PAD) return set_error(HSPZ_E_ARG, "batch arena beyond 4 GiB");
    uint32_t arena_len = (uint32_t)off;
    // the coarse bound's candidates binned by reference position: every pass, or
    // (the default) passes whose query plane stays in an XCD's L2
    const bool bin_want = h->cand_bins == 2 || (h->cand_bins == 1 && arena_len <= CAND_BIN_MAX_QUERY);
    // encoded chunks are copied into a host image of the arena; raw chunks
    // (rc == NULL) are concatenated and encoded into it on the device
    bool any_encoded = false;
    std::vector<uint8_t> arena, rawcat;
    std::vector<EncSeg> segs;
    for (size_t c = 0; c < n_chunks; c++) {
        if (!chunks[c].len) continue;
        if (chunks[c].rc) {
            if (!any_encoded) arena.assign(arena_len, 7);
            any_encoded = true;
            memcpy(arena.data() + blk_off[2 * c], chunks[c].fwd, chunks[c].len);
            memcpy(arena.data() + blk_off[2 * c + 1], chunks[c].rc, chunks[c].len);
        } else {
            encode_segments(segs, (uint32_t)rawcat.size(), (uint32_t)chunks[c].len, blk_off[2 * c],
                            blk_off[2 * c + 1]);
            rawcat.insert(rawcat.end(), chunks[c].fwd, chunks[c].fwd + chunks[c].len);
        }
    }
    // ---- the pass's batches, in the order given
    std::vector<BatchDesc> desc;
    std::vector<uint32_t> poff;
    desc.reserve(hb.size());
    poff.reserve(hb.size() + 1);
    uint64_t npos = 0;
    for (const HostBatch &b : hb) {
        desc.push_back({2 * b.chunk + (uint32_t)b.rev, b.lo, b.hi});
        poff.push_back((uint32_t)std::min<uint64_t>(npos, 0xFFFFFFFFull));
        npos += b.hi - b.lo;
    }
    // scope keys are b * MAX_CHUNKS + chunk in 32 bits
    if (desc.size() >= 0xFFFFFFFFull / MAX_CHUNKS)
        return set_error(HSPZ_E_ARG, "batch too large: %zu seed-and-filter batches", desc.size());
    uint32_t nb = (uint32_t)desc.size();
    poff.push_back((uint32_t)npos);
    if (npos >= GV_MAX_POSITIONS) return set_error(HSPZ_E_ARG, "batch pass too large: %llu positions", (unsigned long long)npos);
    if (nb == 0) return 0;
    const uint32_t per = p->transitions ? 1 + s.kmer_size : 1;
    const uint64_t cap_seeds = npos * per;
    if (cap_seeds >= 0xFFFFFFFFull) return set_error(HSPZ_E_ARG, "batch too large: %llu seeds", (unsigned long long)cap_seeds);

    // ---- device buffers
    int rc;
    if (any_encoded) {
        if ((rc = upload_seq(h, B.arena, arena.data(), arena_len))) return rc;
    } else {
        if ((rc = reserve(h, B.arena, (size_t)arena_len + 2 * PAD))) return rc;
        HIP_TRY(hipMemsetAsync(B.arena.p, 7, (size_t)arena_len + 2 * PAD, h->stream));
    }
    if ((rc = encode_on_device(h, rawcat.data(), rawcat.size(), segs, B.arena.as<uint8_t>() + PAD,
                               B.arena.as<uint8_t>() + PAD)))
        return rc;
    if ((rc = build_planes(h, B.arena, arena_len, B.plane, B.coarse))) return rc;
    if ((rc = reserve(h, B.desc, (size_t)nb * sizeof(BatchDesc)))) return rc;
    if ((rc = reserve(h, B.poff, (size_t)(nb + 1) * 4))) return rc;
    if ((rc = reserve(h, B.blk_off, blk_off.size() * 4))) return rc;
    if ((rc = reserve(h, B.blk_len, blk_len.size() * 4))) return rc;
    if ((rc = reserve(h, B.valid, npos + 16))) return rc;
    if ((rc = reserve(h, B.ids, npos * 4 + 4))) return rc;
    if ((rc = reserve(h, B.voff, (size_t)(nb + 1) * 4))) return rc;
    if ((rc = reserve(h, B.seeds, cap_seeds * 8 + 8))) return rc;
    if ((rc = reserve(h, B.counts, cap_seeds * 4 + 4))) return rc;
    if ((rc = reserve(h, B.prefix, cap_seeds * 8 + 8))) return rc;
    uint32_t cap_tiles = ceil_div(std::max<uint64_t>(cap_seeds, 1), TILE);
    if ((rc = reserve(h, B.tiles, (size_t)cap_tiles * 8 + 8))) return rc;
    if ((rc = reserve(h, B.nchunks, (size_t)nb * 4))) return rc;
    if ((rc = reserve(h, B.starts, (size_t)nb * MAX_CHUNKS * 4))) return rc;
    if ((rc = reserve(h, B.bhits, (size_t)nb * 8))) return rc;
    if ((rc = reserve(h, B.ctr, CTR_N * 4))) return rc;
    HIP_TRY(hipMemcpyAsync(B.desc.p, desc.data(), (size_t)nb * sizeof(BatchDesc), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(B.poff.p, poff.data(), (size_t)(nb + 1) * 4, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(B.blk_off.p, blk_off.data(), blk_off.size() * 4, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(B.blk_len.p, blk_len.data(), blk_len.size() * 4, hipMemcpyHostToDevice, h->stream));

    const uint8_t *ar = B.arena.as<uint8_t>() + PAD;
    uint32_t *ctr = B.ctr.as<uint32_t>();
    ShapeArg sa;
    memset(&sa, 0, sizeof sa);
    memcpy(sa.pos, s.pos, s.kmer_size * 4);
    sa.kmer_size = s.kmer_size;

    {
        // start the survivor and deferred-hit lists at a size the pass is
        // likely to need (a sixteenth and a quarter of its seed slots, or
        // 1/4096 and 1/256 of its expected hits if more), so that large passes
        // rarely pay an overflow rerun
        hspz_batch_state *bs = h->batch;
        if (h->list_cap) {
            bs->cap_surv = bs->cap_slow = bs->cap_cand = h->list_cap;
        } else {
            // the coarse gate's candidates: ~1.5% of random hits and every homologous one
            const double hits = (double)npos * hits_per_position(h, per);
            const uint64_t surv = std::max<uint64_t>(cap_seeds / 16, (uint64_t)(hits / 4096));
            const uint64_t slow = std::max<uint64_t>(cap_seeds / 4, (uint64_t)(hits / 256));
            const uint64_t cand = std::max<uint64_t>(cap_seeds / 4, (uint64_t)(hits / 32));
            bs->cap_surv = (uint32_t)std::max<uint64_t>(bs->cap_surv, std::min<uint64_t>(surv, 1u << 28));
            bs->cap_slow = (uint32_t)std::max<uint64_t>(bs->cap_slow, std::min<uint64_t>(slow, 1u << 28));
            bs->cap_cand = (uint32_t)std::max<uint64_t>(bs->cap_cand, std::min<uint64_t>(cand, 1u << 29));
        }
    }
    for (;;) {  // one attempt per survivor/slow capacity
        hspz_batch_state *bs = h->batch;
        if ((rc = reserve(h, B.surv, (size_t)bs->cap_surv * sizeof(SurvRec)))) return rc;
        if ((rc = reserve(h, B.slow, (size_t)bs->cap_slow * sizeof(SurvRec)))) return rc;
        if ((rc = reserve(h, B.exact, (size_t)bs->cap_slow * sizeof(SurvRec)))) return rc;
        if ((rc = reserve(h, B.mhsp, (size_t)bs->cap_surv * 16))) return rc;
        if ((rc = reserve(h, B.mscope, (size_t)bs->cap_surv * 4))) return rc;
        if ((rc = reserve(h, B.keep, (size_t)bs->cap_surv + 16))) return rc;
        if ((rc = reserve(h, B.kids, (size_t)bs->cap_surv * 4))) return rc;
        if ((rc = reserve(h, B.ohsp, (size_t)bs->cap_surv * 16))) return rc;
        if ((rc = reserve(h, B.oscope, (size_t)bs->cap_surv * 4))) return rc;

        uint32_t nt = std::max<uint32_t>(ceil_div(npos, TILE), 1);
        uint32_t kt = ceil_div(bs->cap_surv, TILE);
        // hit order: k-mer-major once the batch holds about a seed per two k-mers
        const uint32_t nkm = (uint32_t)h->index_len;
        const bool kmer_order = h->clean.p && (h->batch_order == 1 || (h->batch_order == 2 && cap_seeds * 2 >= nkm));
        uint32_t ntk = kmer_order ? ceil_div((uint64_t)nkm + 1, TILE) : 0;
        if ((rc = reserve(h, h->tiles, (size_t)std::max(std::max(kt, nt), ntk) * 4))) return rc;
        if (kmer_order) {
            if ((rc = reserve(h, B.gcnt, ((size_t)nkm + 1) * 4))) return rc;
            if (h->coarse_gate && (rc = reserve(h, B.cand, (size_t)bs->cap_cand * 16))) return rc;
            // the binned copy of the candidates, when the device has room for it
            if (h->coarse_gate && h->cand_planes && h->lookahead && bin_want &&
                B.cand2.cap < (size_t)bs->cap_cand * 16) {
                size_t f = 0, t = 0;
                HIP_TRY(hipMemGetInfo(&f, &t));
                if ((uint64_t)f >= (uint64_t)bs->cap_cand * 18 + (8ull << 30)) {
                    if ((rc = reserve(h, B.cand2, (size_t)bs->cap_cand * 16))) return rc;
                    // bin sizes, bin bases, then one row of per-block counts per bin
                    if ((rc = reserve(h, B.cbins, (size_t)(2 + cand_scatter_blocks(h)) * CAND_BINS * 4))) return rc;
                }
            }
            if ((rc = reserve(h, B.kgaps, (size_t)kgap_capacity(nkm) * sizeof(uint4)))) return rc;
            if ((rc = reserve(h, B.kkey, cap_seeds * 4 + 4))) return rc;   // sort keys (u32)
            if ((rc = reserve(h, B.sq, cap_seeds * 4 + 4))) return rc;     // sort values (u32), in
            if ((rc = reserve(h, B.kvalo, cap_seeds * 4 + 4))) return rc;  // and out
            if ((rc = reserve(h, B.skm, cap_seeds * 4 + 4))) return rc;
            if ((rc = reserve(h, B.qrec, npos * sizeof(QRec) + sizeof(QRec)))) return rc;  // one per position
            if ((rc = ensure_ref_records(h, h->rrec_mode > 0 ||
                                                (h->rrec_mode == 0 && (h->bgate_step >= 96 || h->coarse_gate)))))
                return rc;
        }
        // the coarse gate runs on the reference records, in k-mer-major order
        const bool coarse = kmer_order && h->coarse_gate && h->rrec.p;
        uint32_t *kkey = kmer_order ? B.kkey.as<uint32_t>() : nullptr;
        uint32_t *kval = kmer_order ? B.sq.as<uint32_t>() : nullptr;
        const int key_bits = 32 - __builtin_clz(nkm - 1);  // keys are in [0, nk - 1]
        if (kmer_order) {  // sort temp storage before any capture
            size_t bytes = 0;
            HIP_TRY(rocprim::radix_sort_pairs(nullptr, bytes, kkey, B.skm.as<uint32_t>(), kval, B.kvalo.as<uint32_t>(),
                                              (uint32_t)cap_seeds, 0, key_bits, h->stream));
            if ((rc = reserve(h, B.ktmp, bytes + 256))) return rc;
        }
        BlockOf bo;
        bo.seeds = B.seeds.as<uint64_t>();
        bo.voff = B.voff.as<uint32_t>();
        bo.blk_off = B.blk_off.as<uint32_t>();
        bo.blk_len = B.blk_len.as<uint32_t>();
        bo.batches = B.desc.as<BatchDesc>();
        bo.nb = nb;
        bo.per = per;
        const uint8_t *r = h->ref.as<uint8_t>() + PAD;
        auto enqueue_pre = [&]() {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, h->stream, ctr, (uint32_t)CTR_N, 0u);
        // seeding
        hipLaunchKernelGGL(k_bseed_valid, dim3(std::min<uint32_t>(ceil_div(npos, 256), 16384)), dim3(256), 0, h->stream,
                           ar, B.blk_off.as<uint32_t>(), B.desc.as<BatchDesc>(), B.poff.as<uint32_t>(), nb,
                           (uint32_t)npos, s.size, B.valid.as<uint8_t>());
        hipLaunchKernelGGL(k_tile_count_u8, dim3(nt), dim3(TILE_THREADS), 0, h->stream, B.valid.as<uint8_t>(),
                           (uint32_t)npos, h->tiles.as<uint32_t>());
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_THREADS), 0, h->stream, h->tiles.as<uint32_t>(), nt,
                           ctr + CTR_NVALID);
        hipLaunchKernelGGL(k_tile_emit_u8, dim3(nt), dim3(TILE_THREADS), 0, h->stream, B.valid.as<uint8_t>(),
                           (uint32_t)npos, h->tiles.as<uint32_t>(), B.ids.as<uint32_t>());
        hipLaunchKernelGGL(k_bvoff, dim3(ceil_div(nb + 1, 256)), dim3(256), 0, h->stream, B.ids.as<uint32_t>(),
                           ctr + CTR_NVALID, B.poff.as<uint32_t>(), nb, B.voff.as<uint32_t>());
        hipLaunchKernelGGL(k_bseed_scatter, dim3(std::min<uint32_t>(ceil_div(npos, 256), 16384)), dim3(256), 0,
                           h->stream, ar, B.blk_off.as<uint32_t>(), B.desc.as<BatchDesc>(), B.poff.as<uint32_t>(),
                           nb, B.ids.as<uint32_t>(), ctr + CTR_NVALID, sa, p->transitions, B.seeds.as<uint64_t>());
        // counts, u64 prefix, per-batch chunks
        hipLaunchKernelGGL(k_bnum_tiles, dim3(cap_tiles), dim3(TILE_THREADS), 0, h->stream, B.seeds.as<uint64_t>(),
                           ctr + CTR_NVALID, per, h->index.as<uint32_t>(), B.counts.as<uint32_t>(),
                           B.tiles.as<uint64_t>());
        hipLaunchKernelGGL(k_scan_tiles_u64, dim3(1), dim3(SCAN_THREADS), 0, h->stream, B.tiles.as<uint64_t>(),
                           ctr + CTR_NVALID, per);
        hipLaunchKernelGGL(k_tile_apply_u64, dim3(cap_tiles), dim3(TILE_THREADS), 0, h->stream,
                           B.counts.as<uint32_t>(), ctr + CTR_NVALID, per, B.tiles.as<uint64_t>(),
                           B.prefix.as<uint64_t>());
        hipLaunchKernelGGL(k_bchunks, dim3(ceil_div(nb, 64)), dim3(64), 0, h->stream, B.prefix.as<uint64_t>(),
                           B.voff.as<uint32_t>(), nb, per, h->max_hits, B.nchunks.as<uint32_t>(),
                           B.starts.as<uint32_t>(), B.bhits.as<uint64_t>(), ctr + CTR_ERR);
        if (kmer_order)  // keys of the grouping sort: k-mer for the valid seeds, nk - 1 for unused slots
            hipLaunchKernelGGL(k_kkeys, dim3(std::min<uint32_t>(ceil_div(cap_seeds, 256), 16384)), dim3(256), 0,
                               h->stream, B.seeds.as<uint64_t>(), ctr + CTR_NVALID, per, (uint32_t)cap_seeds, nkm,
                               kkey, kval);
        };
        // k-mer grouping: a stable radix sort of all seed slots by key (run
        // eagerly between the graphs: rocPRIM issues its own fills), then the
        // group starts and one query record per grouped seed
        auto enqueue_sort = [&]() -> int {
            if (!kmer_order) return 0;
            size_t bytes = 0;
            HIP_TRY(rocprim::radix_sort_pairs(nullptr, bytes, kkey, B.skm.as<uint32_t>(), kval, B.kvalo.as<uint32_t>(),
                                              (uint32_t)cap_seeds, 0, key_bits, h->stream));
            if (int e = reserve(h, B.ktmp, bytes + 256)) return e;
            HIP_TRY(rocprim::radix_sort_pairs(B.ktmp.p, bytes, kkey, B.skm.as<uint32_t>(), kval,
                                              B.kvalo.as<uint32_t>(), (uint32_t)cap_seeds, 0, key_bits, h->stream));
            return 0;
        };
        auto enqueue_group = [&]() {
            if (!kmer_order) return;
            hipLaunchKernelGGL(k_kbounds, dim3(std::min<uint32_t>(ceil_div(cap_seeds, 256), 16384)), dim3(256), 0,
                               h->stream, B.skm.as<uint32_t>(), ctr + CTR_NVALID, per, (uint32_t)cap_seeds, nkm,
                               B.gcnt.as<uint32_t>(), B.kgaps.as<uint4>(), ctr + CTR_KGAPS, kgap_capacity(nkm));
            hipLaunchKernelGGL(k_kgaps, dim3(1024), dim3(256), 0, h->stream, B.kgaps.as<uint4>(), ctr + CTR_KGAPS,
                               kgap_capacity(nkm), B.gcnt.as<uint32_t>());
            hipLaunchKernelGGL(k_qrecords, dim3(std::min<uint32_t>(ceil_div(npos, 256), 16384)), dim3(256), 0,
                               h->stream, B.seeds.as<uint64_t>(), ctr + CTR_NVALID, per,
                               B.plane.as<uint32_t>(), B.coarse.as<uint32_t>(), arena_len, h->seed_size,
                               B.qrec.as<QRec>());
        };
        // fused expansion + packed gate, then the exact gate for deferred hits
        BGateFn gate_fn = bgate_kernel(h->bgate_step, h->bgate_threads);
        KGateFn kgate_fn = kgate_kernel(h->bgate_step, h->bgate_threads, h->sum_mode);
        if (kmer_order && lookahead_wanted(h) && h->rrec.p)
            kgate_fn = h->bgate_step == 96 ? k_bgate_ahead<2> : k_bgate_ahead<4>;
        typedef void (*CGateFn)(const QRec *, const uint32_t *, const uint32_t *, const uint32_t *, uint32_t,
                                const uint32_t *, const uint32_t *, const uint32_t *, const uint4 *, uint32_t,
                                uint32_t, unsigned long long *, CoarseCoef, int32_t, int32_t, uint4 *, uint32_t *,
                                uint32_t, SurvRec *, uint32_t *, uint32_t, uint32_t *);
        typedef void (*XGateFn)(const uint32_t *, const uint32_t *, int32_t, int32_t, const QRec *,
                                const uint32_t *, const uint4 *, const uint4 *, const uint32_t *, uint32_t, uint32_t,
                                unsigned long long *, SurvRec *, uint32_t *, uint32_t, SurvRec *, uint32_t *,
                                uint32_t, uint32_t *, const uint32_t *, uint32_t, uint32_t);
        // the paired form (two hits per lane, packed 16-bit sums) where the class scores fit
        // 1: two reference positions per lane sharing each query record; 3: two query
        // records per lane sharing one reference position; 2: one hit per lane
        const bool pair = (h->coarse_gate == 1 || h->coarse_gate == 3) && h->coarse_fits16;
        const bool items2 = pair && h->coarse_gate == 1;
        const bool t1k = h->bgate_threads == 1024, pv = h->coarse_pv, c0 = h->coarse_cc.c0 != 0;
        const CGateFn cgate_fn =
            !pair ? (t1k ? k_bgate_coarse<1024> : k_bgate_coarse<512>)
            : items2
                ? (t1k ? (pv ? (c0 ? k_bgate_coarse2p<1024, true, true> : k_bgate_coarse2p<1024, true, false>)
                             : (c0 ? k_bgate_coarse2p<1024, false, true> : k_bgate_coarse2p<1024, false, false>))
                       : (pv ? (c0 ? k_bgate_coarse2p<512, true, true> : k_bgate_coarse2p<512, true, false>)
                             : (c0 ? k_bgate_coarse2p<512, false, true> : k_bgate_coarse2p<512, false, false>)))
            : t1k ? (pv ? (c0 ? k_bgate_coarse2<1024, true, true> : k_bgate_coarse2<1024, true, false>)
                        : (c0 ? k_bgate_coarse2<1024, false, true> : k_bgate_coarse2<1024, false, false>))
                  : (pv ? (c0 ? k_bgate_coarse2<512, true, true> : k_bgate_coarse2<512, true, false>)
                        : (c0 ? k_bgate_coarse2<512, false, true> : k_bgate_coarse2<512, false, false>));
        XGateFn xgate_fn;
        // the candidates' gate with its loads issued ahead (512-thread blocks) unless the lookahead is off
        const bool xahead = h->lookahead;
        const int xthreads = xahead ? 512 : h->bgate_threads;
        // with the lookahead, the candidates' windows come from the planes unless asked otherwise
        const bool xplanes = xahead && h->cand_planes;
        // and binned by reference position, when the binned copy fits
        const uint32_t cblk = cand_scatter_blocks(h);
        const bool binned = coarse && xplanes && bin_want && B.cand2.p && B.cand2.cap >= (size_t)bs->cap_cand * 16 &&
                            B.cbins.p && B.cbins.cap >= (size_t)(2 + cblk) * CAND_BINS * 4;
        uint32_t cand_shift = 0;  // bins of 2^shift reference bases, at most CAND_BINS of them
        while ((h->ref_len >> cand_shift) >= CAND_BINS) cand_shift++;
        if (xplanes)
            xgate_fn = h->sum_mode == SUM_RAW      ? k_bgate_exact_ahead<SUM_RAW, true>
                       : h->sum_mode == SUM_TREE16 ? k_bgate_exact_ahead<SUM_TREE16, true>
                                                   : k_bgate_exact_ahead<SUM_TREE8, true>;
        else if (xahead)
            xgate_fn = h->sum_mode == SUM_RAW      ? k_bgate_exact_ahead<SUM_RAW>
                       : h->sum_mode == SUM_TREE16 ? k_bgate_exact_ahead<SUM_TREE16>
                                                   : k_bgate_exact_ahead<SUM_TREE8>;
        else if (h->bgate_threads == 1024)
            xgate_fn = h->sum_mode == SUM_RAW      ? k_bgate_exact<1024, SUM_RAW>
                       : h->sum_mode == SUM_TREE16 ? k_bgate_exact<1024, SUM_TREE16>
                                                   : k_bgate_exact<1024, SUM_TREE8>;
        else
            xgate_fn = h->sum_mode == SUM_RAW      ? k_bgate_exact<512, SUM_RAW>
                       : h->sum_mode == SUM_TREE16 ? k_bgate_exact<512, SUM_TREE16>
                                                   : k_bgate_exact<512, SUM_TREE8>;
        int grid = h->bgate_blocks, xgrid = h->bgate_blocks;
        if (coarse && xahead) {  // the candidates' gate keeps its resident grid
            int xper_cu = 0;
            HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&xper_cu, xgate_fn, xthreads, 0));
            xgrid = std::max(xper_cu, 1) * h->num_cus;
        }
        if (grid == 0) {  // resident grid: occupancy of this instantiation x CUs
            int per_cu = 0, xper_cu = 0;
            if (coarse) {
                HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, cgate_fn, h->bgate_threads, 0));
                HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&xper_cu, xgate_fn, xthreads, 0));
            } else if (kmer_order) {
                HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kgate_fn, h->bgate_threads, 0));
            } else {
                HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, gate_fn, h->bgate_threads, 0));
            }
            grid = std::max(per_cu, 1) * h->num_cus;
            if (!(coarse && xahead)) xgrid = std::max(xper_cu, 1) * h->num_cus;
        }
        h->bgate_grid_used = grid;
        h->batch_order_used = kmer_order ? 1 : 0;
        h->coarse_used = !coarse ? 0 : !pair ? 1 : items2 ? 3 : 2;
        h->cand_binned = binned ? 1 : 0;
        // the coarse coefficients, with the long-gate routing off when the flag says so
        CoarseCoef ccl = h->coarse_cc;
        if (!h->coarse_long) ccl.lm = 32768;  // above every 16-bit sum
        auto enqueue_gate = [&]() {
        if (coarse) {
            // coarse bound over every hit, then the two-cell maps over its candidates
            hipLaunchKernelGGL(cgate_fn, dim3(grid), dim3(h->bgate_threads), 0, h->stream, B.qrec.as<QRec>(),
                               B.skm.as<uint32_t>(), B.kvalo.as<uint32_t>(), B.gcnt.as<uint32_t>(), nkm,
                               h->index.as<uint32_t>(), h->pos.as<uint32_t>(), h->clean.as<uint32_t>(),
                               h->rrec.as<uint4>(), h->seed_size, (uint32_t)h->bgate_claim,
                               (unsigned long long *)(ctr + CTR_WORK), ccl, h->xdrop, h->hspthresh,
                               B.cand.as<uint4>(), ctr + CTR_CAND, bs->cap_cand, B.slow.as<SurvRec>(),
                               ctr + CTR_SLOW, bs->cap_slow, ctr + CTR_ERR);
            if (binned) {  // the candidates in reference-position bins, for the plane reads
                uint32_t *sizes = B.cbins.as<uint32_t>(), *bases = sizes + CAND_BINS, *table = bases + CAND_BINS;
                hipLaunchKernelGGL(k_cand_hist, dim3(cblk), dim3(CAND_BINS), 0, h->stream, B.cand.as<uint4>(),
                                   ctr + CTR_CAND, bs->cap_cand, cand_shift, table);
                hipLaunchKernelGGL(k_cand_rows, dim3(CAND_BINS), dim3(256), 0, h->stream, table, cblk, sizes);
                hipLaunchKernelGGL(k_cand_scan, dim3(1), dim3(CAND_BINS), 0, h->stream, sizes, bases);
                hipLaunchKernelGGL(k_cand_scatter<8>, dim3(cblk), dim3(CAND_BINS), 0, h->stream, B.cand.as<uint4>(),
                                   ctr + CTR_CAND, bs->cap_cand, cand_shift, table, bases, B.cand2.as<uint4>());
            }
            hipLaunchKernelGGL(xgate_fn, dim3(xgrid), dim3(xthreads), 0, h->stream, h->ref2.as<uint32_t>(),
                               h->pairtab.as<uint32_t>() + (xplanes ? 0 : PAIRTAB_GRAY), h->xdrop, h->hspthresh,
                               B.qrec.as<QRec>(), B.kvalo.as<uint32_t>(), h->rrec.as<uint4>(),
                               binned ? B.cand2.as<uint4>() : B.cand.as<uint4>(),
                               ctr + CTR_CAND, bs->cap_cand, (uint32_t)h->bgate_claim,
                               (unsigned long long *)(ctr + CTR_WORK2), B.surv.as<SurvRec>(), ctr + CTR_SURV,
                               bs->cap_surv, B.slow.as<SurvRec>(), ctr + CTR_SLOW, bs->cap_slow, ctr + CTR_ERR,
                               B.plane.as<uint32_t>(), per, binned ? 1u : 0u);
        } else if (kmer_order)
            hipLaunchKernelGGL(kgate_fn, dim3(grid), dim3(h->bgate_threads), 0, h->stream, h->ref2.as<uint32_t>(),
                               h->pairtab.as<uint32_t>() + PAIRTAB_GRAY, h->xdrop, h->hspthresh, B.qrec.as<QRec>(),
                               B.skm.as<uint32_t>(), B.kvalo.as<uint32_t>(), B.gcnt.as<uint32_t>(), nkm,
                               h->index.as<uint32_t>(), h->pos.as<uint32_t>(), h->clean.as<uint32_t>(),
                               h->rrec.p ? h->rrec.as<uint4>() : nullptr, h->seed_size, (uint32_t)h->bgate_claim,
                               (unsigned long long *)(ctr + CTR_WORK),
                               B.surv.as<SurvRec>(), ctr + CTR_SURV, bs->cap_surv, B.slow.as<SurvRec>(),
                               ctr + CTR_SLOW, bs->cap_slow, ctr + CTR_ERR);
        else
            hipLaunchKernelGGL(gate_fn, dim3(grid), dim3(h->bgate_threads), 0, h->stream, h->ref2.as<uint32_t>(),
                               h->refx.as<uint32_t>(), h->ref_len, B.plane.as<uint32_t>(), B.coarse.as<uint32_t>(),
                               arena_len, h->pairtab.as<uint32_t>(), h->xdrop, h->hspthresh, B.seeds.as<uint64_t>(),
                               ctr + CTR_NVALID, per, h->index.as<uint32_t>(), h->pos.as<uint32_t>(), h->seed_size,
                               (uint32_t)h->bgate_claim, (unsigned long long *)(ctr + CTR_WORK),
                               B.surv.as<SurvRec>(), ctr + CTR_SURV, bs->cap_surv, B.slow.as<SurvRec>(),
                               ctr + CTR_SLOW, bs->cap_slow, ctr + CTR_ERR);
        // deferred hits: past the fixed windows on the planes, lane per hit; the
        // ones that meet an exception or a sequence end, exactly, wave per hit
        hipLaunchKernelGGL(k_bgate_long, dim3(2 * h->num_cus), dim3(1024), 0, h->stream, h->ref2.as<uint32_t>(),
                           h->refx.as<uint32_t>(), h->ref_len, B.plane.as<uint32_t>(), B.coarse.as<uint32_t>(),
                           arena_len, h->pairtab.as<uint32_t>(), h->xdrop, h->hspthresh, B.slow.as<SurvRec>(),
                           ctr + CTR_SLOW, bs->cap_slow, B.surv.as<SurvRec>(), ctr + CTR_SURV, bs->cap_surv,
                           B.exact.as<SurvRec>(), ctr + CTR_EXACT, ctr + CTR_ERR);
        hipLaunchKernelGGL(k_bgate_wave, dim3(2048), dim3(256), 0, h->stream, r, h->ref_len, ar, arena_len,
                           h->sep_drops ? 1 : 0, bo,
                           h->sub.as<int32_t>(), h->xdrop, h->hspthresh, B.exact.as<SurvRec>(), ctr + CTR_EXACT,
                           bs->cap_slow, B.surv.as<SurvRec>(), ctr + CTR_SURV, bs->cap_surv, ctr + CTR_ERR);
        };
        auto enqueue_post = [&]() {
        // materializer and compaction of kept records
        hipLaunchKernelGGL(k_fill_u32, dim3(std::min<uint32_t>(ceil_div(bs->cap_surv / 4 + 4, 256), 4096)), dim3(256), 0,
                           h->stream, (uint32_t *)B.keep.p, bs->cap_surv / 4 + 4, 0u);
        // one lane per survivor, then one wave per survivor for the ones a lane
        // handed on (B.kids is free until the compaction below)
        if (h->lane_materialize) {
            hipLaunchKernelGGL(k_bmaterialize_lane, dim3(8 * h->num_cus), dim3(256), 0, h->stream, r, h->ref_len,
                               ar, bo, B.starts.as<uint32_t>(), B.nchunks.as<uint32_t>(), h->sub.as<int32_t>(),
                               h->xdrop, h->hspthresh, h->noentropy, 1.3862943649291992, B.surv.as<SurvRec>(),
                               ctr + CTR_SURV, bs->cap_surv, B.mhsp.as<SegmentPair>(), B.mscope.as<uint32_t>(),
                               B.keep.as<uint8_t>(), B.kids.as<uint32_t>(), ctr + CTR_LONG, bs->cap_surv,
                               ctr + CTR_MWORK, ctr + CTR_ERR);
            hipLaunchKernelGGL(k_bmaterialize, dim3(4096), dim3(256), 0, h->stream, r, h->ref_len, ar, bo,
                               B.starts.as<uint32_t>(), B.nchunks.as<uint32_t>(), h->sub.as<int32_t>(), h->xdrop,
                               h->hspthresh, h->noentropy, 1.3862943649291992, B.surv.as<SurvRec>(),
                               ctr + CTR_SURV, bs->cap_surv, B.mhsp.as<SegmentPair>(), B.mscope.as<uint32_t>(),
                               B.keep.as<uint8_t>(), B.kids.as<uint32_t>(), ctr + CTR_LONG, bs->cap_surv);
        } else {
            hipLaunchKernelGGL(k_bmaterialize, dim3(4096), dim3(256), 0, h->stream, r, h->ref_len, ar, bo,
                               B.starts.as<uint32_t>(), B.nchunks.as<uint32_t>(), h->sub.as<int32_t>(), h->xdrop,
                               h->hspthresh, h->noentropy, 1.3862943649291992, B.surv.as<SurvRec>(),
                               ctr + CTR_SURV, bs->cap_surv, B.mhsp.as<SegmentPair>(), B.mscope.as<uint32_t>(),
                               B.keep.as<uint8_t>(), nullptr, nullptr, 0u);
        }
        hipLaunchKernelGGL(k_tile_count_u8, dim3(kt), dim3(TILE_THREADS), 0, h->stream, B.keep.as<uint8_t>(),
                           bs->cap_surv, h->tiles.as<uint32_t>());
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_THREADS), 0, h->stream, h->tiles.as<uint32_t>(), kt,
                           ctr + CTR_KEPT);
        hipLaunchKernelGGL(k_tile_emit_u8, dim3(kt), dim3(TILE_THREADS), 0, h->stream, B.keep.as<uint8_t>(),
                           bs->cap_surv, h->tiles.as<uint32_t>(), B.kids.as<uint32_t>());
        hipLaunchKernelGGL(k_gather_scoped, dim3(1024), dim3(256), 0, h->stream, B.mhsp.as<SegmentPair>(),
                           B.mscope.as<uint32_t>(), B.kids.as<uint32_t>(), ctr + CTR_KEPT, B.ohsp.as<SegmentPair>(),
                           B.oscope.as<uint32_t>());
        };
        const auto t_launch = std::chrono::steady_clock::now();
        if (h->batch_graph && use_graph) {
            GraphKey key;
            memset(&key, 0, sizeof key);
            const void *ptrs[] = {B.arena.p, B.plane.p, B.coarse.p, B.desc.p, B.poff.p, B.blk_off.p, B.blk_len.p,
                                  B.valid.p, B.ids.p, B.voff.p, B.seeds.p, B.counts.p, B.prefix.p, B.tiles.p,
                                  B.nchunks.p, B.starts.p, B.bhits.p, B.surv.p, B.slow.p, B.mhsp.p, B.mscope.p,
                                  B.keep.p, B.kids.p, B.ohsp.p, B.oscope.p, B.ctr.p, h->tiles.p, h->ref2.p,
                                  h->refx.p, h->pairtab.p, h->index.p, h->pos.p, B.gcnt.p, B.sq.p, B.skm.p,
                                  h->clean.p, B.qrec.p, h->rrec.p, B.kvalo.p, B.exact.p, B.kkey.p, B.kgaps.p,
                                  B.cand.p, binned ? B.cand2.p : nullptr, binned ? B.cbins.p : nullptr};
            static_assert(sizeof ptrs <= sizeof key.ptrs, "graph key pointer slots");
            memcpy(key.ptrs, ptrs, sizeof ptrs);
            uint64_t vals[] = {nb, npos, arena_len, bs->cap_surv, bs->cap_slow | (uint64_t)bs->cap_cand << 32,
                               cap_seeds | (uint64_t)coarse << 40 | (uint64_t)xgrid << 41, per, h->max_hits,
                               (uint64_t)grid | (uint64_t)h->bgate_claim << 32,
                               (uint64_t)p->transitions | (uint64_t)h->lane_materialize << 8 | (uint64_t)h->coarse_long << 9, nt, kt,
                               cap_tiles, (uint64_t)h->bgate_step | (uint64_t)h->bgate_threads << 32,
                               (uint64_t)(uintptr_t)h->sub.p, s.kmer_size, (uint64_t)(uintptr_t)h->ref.p,
                               (uint64_t)kmer_order, ntk, (uint64_t)(uintptr_t)(const void *)kgate_fn,
                               (uint64_t)(uintptr_t)(const void *)cgate_fn,
                               (uint64_t)(uintptr_t)(const void *)xgate_fn | (uint64_t)xthreads << 48};
            static_assert(sizeof vals <= sizeof key.vals, "graph key value slots");
            memcpy(key.vals, vals, sizeof vals);
            memcpy(key.shape, s.pos, sizeof s.pos);
            key.shape[32] = s.size;
            if (!bs->have_graph || memcmp(&key, &bs->key, sizeof key) != 0) {
                std::lock_guard<std::mutex> lk(capture_mutex());
                for (int g = 0; g < 4; g++) {
                    if (bs->exec[g]) (void)hipGraphExecDestroy(bs->exec[g]);
                    if (bs->graph[g]) (void)hipGraphDestroy(bs->graph[g]);
                    bs->exec[g] = nullptr;
                    bs->graph[g] = nullptr;
                }
                bs->have_graph = false;
                for (int g = 0; g < 4; g++) {
                    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
                    if (g == 0) enqueue_pre();
                    else if (g == 1) enqueue_group();
                    else if (g == 2) enqueue_gate();
                    else enqueue_post();
                    hipError_t ce = hipStreamEndCapture(h->stream, &bs->graph[g]);
                    if (ce != hipSuccess) return set_error(-(int)ce, "graph capture: %s", hipGetErrorString(ce));
                    HIP_TRY(hipGraphInstantiate(&bs->exec[g], bs->graph[g], nullptr, nullptr, 0));
                }
                bs->key = key;
                bs->have_graph = true;
                bs->graph_captures++;
            } else {
                bs->graph_replays++;
            }
            // timing events sit between the graph launches, on the stream
            HIP_TRY(hipGraphLaunch(bs->exec[0], h->stream));
            if ((rc = enqueue_sort())) return rc;
            HIP_TRY(hipGraphLaunch(bs->exec[1], h->stream));
            HIP_TRY(hipEventRecord(h->ev_gate0, h->stream));
            HIP_TRY(hipGraphLaunch(bs->exec[2], h->stream));
            HIP_TRY(hipEventRecord(h->ev_gate1, h->stream));
            HIP_TRY(hipGraphLaunch(bs->exec[3], h->stream));
            HIP_TRY(hipEventRecord(h->ev_post1, h->stream));
        } else {
            enqueue_pre();
            if ((rc = enqueue_sort())) return rc;
            enqueue_group();
            HIP_TRY(hipEventRecord(h->ev_gate0, h->stream));
            enqueue_gate();
            HIP_TRY(hipEventRecord(h->ev_gate1, h->stream));
            enqueue_post();
            HIP_TRY(hipEventRecord(h->ev_post1, h->stream));
        }
        h->launches += 22;
        HIP_TRY(hipGetLastError());
        uint32_t hc[CTR_N];
        HIP_TRY(hipMemcpyAsync(h->h_small, ctr, CTR_N * 4, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        const auto t_synced = std::chrono::steady_clock::now();
        h->pipeline_syncs++;
        memcpy(hc, h->h_small, sizeof hc);
        float ms = 0, post_ms = 0;
        if (hipEventElapsedTime(&ms, h->ev_gate0, h->ev_gate1) != hipSuccess) {
            (void)hipGetLastError();
            ms = 0;
        }
        if (hipEventElapsedTime(&post_ms, h->ev_gate1, h->ev_post1) != hipSuccess) {
            (void)hipGetLastError();
            post_ms = 0;
        }
        h->last_gate_ms = ms;
        h->last_post_ms = post_ms;
        h->last_surv = hc[CTR_SURV];
        h->last_deferred = hc[CTR_SLOW];
        h->last_cand = coarse ? hc[CTR_CAND] : 0;
        // fatal conditions first: the gate still fills the lists on a failing
        // batch, and an overflow rerun would only repeat the failure
        if (hc[CTR_ERR] & 1u) return set_error(HSPZ_E_CHUNK, "a single seed exceeds max_hits %u", h->max_hits);
        if (hc[CTR_ERR] & 2u) return set_error(HSPZ_E_CHUNK, "a batch needs more than %u max_hits chunks", MAX_CHUNKS);
        if (hc[CTR_ERR] & 4u) return set_error(HSPZ_E_ARG, "a batch has more than 2^32 hits");
        if (hc[CTR_ERR] & 8u) {  // survivor or slow list overflowed: grow and rerun
            uint32_t ns = 0, nd = 0, nc = bs->cap_cand;
            if (!grow_list_cap(bs->cap_surv, hc[CTR_SURV], &ns) || !grow_list_cap(bs->cap_slow, hc[CTR_SLOW], &nd) ||
                (coarse && !grow_list_cap(bs->cap_cand, hc[CTR_CAND], &nc)))
                return set_error(HSPZ_E_OOM,
                                 "batch needs more than %u survivor, deferred-hit or candidate records (%u, %u, %u)",
                                 LIST_CAP_MAX, hc[CTR_SURV], hc[CTR_SLOW], hc[CTR_CAND]);
            bs->cap_surv = ns;
            bs->cap_slow = nd;
            bs->cap_cand = nc;
            h->batch_reruns++;
            // the failed attempt's time is device time, not the rerun's preparation
            const auto t_now = std::chrono::steady_clock::now();
            h->batch_device_ns += (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(t_now - t_launch)
                                      .count();
            t_begin = t_now;
            continue;
        }
        uint32_t nk = hc[CTR_KEPT], nout = nk;
        const void *src_hsp = B.ohsp.p, *src_scope = B.oscope.p;
        if (nk && h->device_dedup) {
            if ((rc = device_dedup(h, B, nk, (uint64_t)nb * MAX_CHUNKS, &nout))) return rc;
            src_hsp = B.mhsp.p;
            src_scope = B.mscope.p;
            *deduped = true;
        }
        if (dout && h->device_dedup) {
            // straight into the call's output: each batch's count from the sorted
            // scopes (B.starts is free once the pass has run), then the records
            dout->batch_counts.assign(nb, 0);
            if (nout) {
                *deduped = true;
                uint32_t *bstart = B.starts.as<uint32_t>();
                hipLaunchKernelGGL(k_scope_starts, dim3(ceil_div(nb + 1, 256)), dim3(256), 0, h->stream,
                                   (const uint32_t *)src_scope, nout, nb, MAX_CHUNKS, bstart);
                HIP_TRY(hipGetLastError());
                std::vector<uint32_t> bs0(nb + 1);
                HIP_TRY(hipMemcpyAsync(bs0.data(), bstart, (size_t)(nb + 1) * 4, hipMemcpyDeviceToHost, h->stream));
                if (dout->len + nout > dout->cap) {
                    const size_t cap = std::max(dout->len + nout, dout->cap + dout->cap / 2);
                    void *nbuf = realloc(dout->buf, cap * 16);
                    if (!nbuf) return set_error(HSPZ_E_OOM, "host output allocation of %zu records", cap);
                    dout->buf = (hspz_segment_pair *)nbuf;
                    dout->cap = cap;
                }
                HIP_TRY(hipStreamSynchronize(h->stream));
                if ((rc = read_to_host(h, dout->buf + dout->len, src_hsp, (size_t)nout * 16))) return rc;
                for (uint32_t b = 0; b < nb; b++) dout->batch_counts[b] = bs0[b + 1] - bs0[b];
                dout->len += nout;
            }
            hsps.clear();
            scopes.clear();
        } else {
            hsps.resize(nout);
            scopes.resize(nout);
            if (nout) {
                HIP_TRY(hipMemcpyAsync(hsps.data(), src_hsp, (size_t)nout * 16, hipMemcpyDeviceToHost, h->stream));
                HIP_TRY(hipMemcpyAsync(scopes.data(), src_scope, (size_t)nout * 4, hipMemcpyDeviceToHost, h->stream));
            }
        }
        // the batches' hit counts, behind the dedup on the stream (a copy to
        // pageable memory returns once it has landed)
        std::vector<uint64_t> bh(nb);
        HIP_TRY(hipMemcpyAsync(bh.data(), B.bhits.p, (size_t)nb * 8, hipMemcpyDeviceToHost, h->stream));
        uint64_t hits = 0;
        for (uint64_t x : bh) hits += x;
        HIP_TRY(hipStreamSynchronize(h->stream));
        const auto t_read = std::chrono::steady_clock::now();
        auto ns = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
            return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(b - a).count();
        };
        h->batch_prep_ns += ns(t_begin, t_launch);
        h->batch_device_ns += ns(t_launch, t_synced);
        h->batch_read_ns += ns(t_synced, t_read);
        h->gate_ns += (uint64_t)((double)ms * 1e6);
        h->gate_hits += hits;
        h->gate_launches++;
        h->gate_deferred += hc[CTR_SLOW];
        if (st) {
            st->seeds += (uint64_t)hc[CTR_NVALID] * per;
            st->hits += hits;
            st->raw_hsps += nk;
            st->batches += nb;
        }
        return 0;
    }
}

// Device bytes one query position of a batch pass needs (k-mer-major order,
// up to per = 16 seed slots): valid flag and rank 5; per slot the seed 8, its
// count 4 and prefix 8, the grouping keys and values in and out 16, the sort's
// temporary storage ~8, the survivor and deferred-hit lists' share ~12; the
// query record 64; the arena and its planes ~3.
constexpr uint64_t PASS_BYTES_PER_POSITION = 5 + 16 * 56 + 64 + 3;

// Hits one device pass is sized for: its survivor and deferred-hit lists grow
// with the hits (about 0.1-0.5% of them are deferred), so a pass against a
// dense index holds fewer positions.
constexpr double PASS_HITS = 32e9;

// Query positions per device pass: the engine's HSPZ_FLAG_BATCH_MAX_POSITIONS,
// or (0) as many as half of the device memory that is free or already held by
// this engine's pass buffers allows, within [2^20, 2^26], and at most
// PASS_HITS expected hits.
uint64_t pass_positions(hspz_engine *h, uint32_t per) {
    if (h->batch_max_positions) return h->batch_max_positions;
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) {
        (void)hipGetLastError();
        return 1ull << 24;
    }
    uint64_t held = 0;
    const BatchBufs &B = h->batch->b;
    for (const DevBuf *b : {&B.arena, &B.plane, &B.coarse, &B.valid, &B.ids, &B.seeds, &B.counts, &B.prefix,
                            &B.surv, &B.slow, &B.exact, &B.mhsp, &B.ohsp, &B.sq, &B.skm, &B.qrec, &B.kvalo,
                            &B.ktmp, &B.kkey})
        held += b->cap;
    const uint64_t avail = (uint64_t)f + held;
    uint64_t want = avail > (4ull << 30) ? (avail - (4ull << 30)) / 2 / PASS_BYTES_PER_POSITION : 0;
    want = std::min<uint64_t>(want, (uint64_t)(PASS_HITS / std::max(hits_per_position(h, per), 1.0)));
    return std::min<uint64_t>(std::max<uint64_t>(want, 1ull << 20), 1ull << 26);
}

}  // namespace

extern "C" int hspz_filter_batch(hspz_engine *h, const hspz_pass_params *p, const hspz_query_chunk *chunks,
                                 size_t n_chunks, hspz_batch_output *out, hspz_pass_stats *stats) {
    clear_error();
    if (!h || !p || !out || (n_chunks && !chunks)) return set_error(HSPZ_E_ARG, "null argument");
    memset(out, 0, sizeof *out);
    if (!h->fast_gate) return set_error(HSPZ_E_ARG, "batched path needs the packed gate (xdrop in [0, 4000], ACGT scores <= 32767)");
    Shape s;
    if (int rc = parse_shape(p->pattern ? p->pattern : "12of19", &s)) return rc;
    if (s.size != h->seed_size) return set_error(HSPZ_E_SHAPE, "seed shape span %u != engine seed size %u", s.size, h->seed_size);
    if (p->wga_chunk == 0 || p->interval_size == 0) return set_error(HSPZ_E_ARG, "zero chunk or interval size");
    for (size_t c = 0; c < n_chunks; c++) {
        if (chunks[c].len && !chunks[c].fwd) return set_error(HSPZ_E_ARG, "chunk %zu has null data", c);
        if (chunks[c].len >= 0x7FFFFFFFull - 2 * QPAD) return set_error(HSPZ_E_ARG, "chunk %zu is longer than 2^31 bases", c);
    }
    HIP_TRY(hipSetDevice(h->device));
    if (!h->batch) h->batch = new hspz_batch_state();

    // ---- every batch of every chunk in the reference's order (run.rs:278-299)
    std::vector<HostBatch> all;
    std::vector<size_t> civ(n_chunks, 0);
    for (size_t c = 0; c < n_chunks; c++) {
        uint32_t bl = (uint32_t)chunks[c].len;
        if (bl <= s.size) continue;
        uint32_t end_pos = bl - s.size, qbl = bl - s.size;
        for (uint64_t x = 0; x < end_pos; x += p->interval_size) civ[c]++;
        for (uint32_t i = 0; i < civ[c]; i++) {
            uint32_t a = (uint32_t)((uint64_t)i * p->interval_size);
            uint32_t e = (uint32_t)std::min<uint64_t>(end_pos, (uint64_t)a + p->interval_size);
            for (int rev = 0; rev < 2; rev++) {
                if ((rev == 0 && !p->plus) || (rev == 1 && !p->minus)) continue;
                uint32_t lo = rev ? qbl - e : a, hi = rev ? qbl - a : e;
                for (uint64_t x = lo; x < hi; x += p->wga_chunk)
                    all.push_back({(uint32_t)c, i, rev, (uint32_t)x,
                                   (uint32_t)std::min<uint64_t>(x + p->wga_chunk, (uint64_t)hi + 1)});
            }
        }
    }
    std::vector<size_t> ibase(n_chunks + 1, 0);
    for (size_t c = 0; c < n_chunks; c++) ibase[c + 1] = ibase[c] + civ[c];
    const size_t nlists = 2 * ibase[n_chunks];
    std::vector<std::vector<hspz_segment_pair>> lists(nlists);

    // ---- device passes over consecutive runs of batches: each holds at most
    // pass_positions() query positions (a batch larger than that runs alone)
    // and 2^20 batches. Every batch is its own dedup scope, so splitting the
    // call between batches changes no list.
    const uint64_t pmax = pass_positions(h, p->transitions ? 1 + s.kmer_size : 1);
    std::vector<std::pair<size_t, size_t>> passes;
    for (size_t b0 = 0; b0 < all.size();) {
        size_t b1 = b0;
        uint64_t np = 0;
        while (b1 < all.size() && b1 - b0 < (1u << 20)) {
            uint64_t w = all[b1].hi - all[b1].lo;
            if (b1 > b0 && np + w > pmax) break;
            np += w;
            b1++;
        }
        passes.push_back({b0, b1});
        b0 = b1;
    }
    h->batch_passes_last = passes.size();
    hspz_pass_stats st = {0, 0, 0, 0, 0};
    double post_ns = 0;
    // device dedup: every pass appends its lists, in output order, to one array
    DirectOut dout;
    std::vector<size_t> dcount(nlists, 0);
    struct FreeOnExit {
        DirectOut &d;
        ~FreeOnExit() { free(d.buf); }
    } free_on_exit{dout};
    const bool direct = h->device_dedup;
    for (const auto &ps : passes) {
        const size_t cf = all[ps.first].chunk, cl = all[ps.second - 1].chunk;
        std::vector<HostBatch> hb(all.begin() + ps.first, all.begin() + ps.second);
        for (HostBatch &b : hb) b.chunk -= (uint32_t)cf;
        std::vector<hspz_segment_pair> raw;
        std::vector<uint32_t> scopes;
        bool deduped = false;
        dout.batch_counts.clear();
        // one pass per call replays a captured graph; several run eagerly (their shapes differ)
        if (int rc = batch_pass(h, s, p, chunks + cf, cl - cf + 1, hb, raw, scopes, &st, &deduped,
                                passes.size() == 1, direct ? &dout : nullptr))
            return rc;
        const auto t_post = std::chrono::steady_clock::now();
        if (direct) {
            for (size_t b = 0; b < dout.batch_counts.size(); b++) {
                const HostBatch &x = all[ps.first + b];
                dcount[2 * (ibase[x.chunk] + x.interval) + x.rev] += dout.batch_counts[b];
            }
            post_ns += (double)std::chrono::duration_cast<std::chrono::nanoseconds>(
                           std::chrono::steady_clock::now() - t_post)
                           .count();
            continue;
        }
        // group by scope (local batch * MAX_CHUNKS + chunk): ascending scope = output order
        std::vector<uint32_t> order(raw.size());
        std::vector<hspz_segment_pair> grouped;
        if (deduped) {  // already grouped by ascending scope and ordered within each
            for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
            grouped.swap(raw);
        } else {  // counting sort by scope, stable
            uint32_t smax = 0;
            for (uint32_t sc : scopes) smax = std::max(smax, sc);
            std::vector<uint32_t> cnt(raw.empty() ? 1 : (size_t)smax + 2, 0);
            for (uint32_t sc : scopes) cnt[sc + 1]++;
            for (size_t i = 1; i < cnt.size(); i++) cnt[i] += cnt[i - 1];
            for (uint32_t i = 0; i < order.size(); i++) order[cnt[scopes[i]]++] = i;
            grouped.resize(raw.size());
            for (size_t i = 0; i < order.size(); i++) grouped[i] = raw[order[i]];
        }
        std::vector<size_t> gstart;
        for (size_t i = 0; i < order.size(); i++)
            if (i == 0 || scopes[order[i]] != scopes[order[i - 1]]) gstart.push_back(i);
        gstart.push_back(order.size());
        size_t ng = gstart.size() - 1;
        std::vector<size_t> kept(ng);
        if (deduped) {
            for (size_t g = 0; g < ng; g++) kept[g] = gstart[g + 1] - gstart[g];
        } else {
            unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
            std::vector<std::thread> th;
            for (unsigned t = 0; t < nth; t++)
                th.emplace_back([&, t] {
                    for (size_t g = t; g < ng; g += nth)
                        kept[g] = dedup_and_order(grouped.data() + gstart[g], gstart[g + 1] - gstart[g]);
                });
            for (auto &x : th) x.join();
        }
        // append per chunk -> interval -> fw, rc (scopes ascending = batch order)
        for (size_t g = 0; g < ng; g++) {
            const HostBatch &x = all[ps.first + scopes[order[gstart[g]]] / MAX_CHUNKS];
            auto &dst = lists[2 * (ibase[x.chunk] + x.interval) + x.rev];
            dst.insert(dst.end(), grouped.begin() + gstart[g], grouped.begin() + gstart[g] + kept[g]);
        }
        post_ns += (double)std::chrono::duration_cast<std::chrono::nanoseconds>(
                       std::chrono::steady_clock::now() - t_post)
                       .count();
    }
    const auto t_out = std::chrono::steady_clock::now();
    if (direct) {
        st.hsps = dout.len;
        out->counts = (size_t *)malloc(std::max<size_t>(nlists, 1) * sizeof(size_t));
        out->chunk_intervals = (size_t *)malloc(std::max<size_t>(n_chunks, 1) * sizeof(size_t));
        if (!out->counts || !out->chunk_intervals) {
            hspz_batch_output_free(out);
            return set_error(HSPZ_E_OOM, "host output allocation");
        }
        if (!dout.buf) dout.buf = (hspz_segment_pair *)malloc(16);  // an empty call still returns an array
        if (!dout.buf) {
            hspz_batch_output_free(out);
            return set_error(HSPZ_E_OOM, "host output allocation");
        }
        out->hsps = dout.buf;  // malloc / realloc: 16-byte aligned, released by free
        dout.buf = nullptr;
        for (size_t i = 0; i < nlists; i++) out->counts[i] = dcount[i];
        for (size_t c = 0; c < n_chunks; c++) out->chunk_intervals[c] = civ[c];
        out->n_hsps = dout.len;
        out->n_chunks = n_chunks;
        h->batch_post_ns += (uint64_t)post_ns + (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                                                    std::chrono::steady_clock::now() - t_out)
                                                    .count();
        if (stats) *stats = st;
        return 0;
    }
    size_t total = 0;
    for (auto &l : lists) total += l.size();
    st.hsps = total;
    out->hsps = (hspz_segment_pair *)aligned_alloc(16, std::max<size_t>(total, 1) * 16);
    out->counts = (size_t *)malloc(std::max<size_t>(nlists, 1) * sizeof(size_t));
    out->chunk_intervals = (size_t *)malloc(std::max<size_t>(n_chunks, 1) * sizeof(size_t));
    if (!out->hsps || !out->counts || !out->chunk_intervals) {
        hspz_batch_output_free(out);
        return set_error(HSPZ_E_OOM, "host output allocation");
    }
    size_t at = 0;
    for (size_t i = 0; i < nlists; i++) {
        memcpy(out->hsps + at, lists[i].data(), lists[i].size() * 16);
        at += lists[i].size();
        out->counts[i] = lists[i].size();
    }
    for (size_t c = 0; c < n_chunks; c++) out->chunk_intervals[c] = civ[c];
    out->n_hsps = total;
    out->n_chunks = n_chunks;
    h->batch_post_ns += (uint64_t)post_ns + (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                                                std::chrono::steady_clock::now() - t_out)
                                                .count();
    if (stats) *stats = st;
    return 0;
}

extern "C" void hspz_batch_output_free(hspz_batch_output *out) {
    if (!out) return;
    free(out->hsps);
    free(out->counts);
    free(out->chunk_intervals);
    memset(out, 0, sizeof *out);
}

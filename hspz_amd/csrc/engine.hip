// engine.hip — the reference-resident seed-and-filter engine behind the C ABI
// (include/hspz.h). Restates hspZ's `Engine` (src/gpu/mod.rs:817-2307) on HIP:
// one stream per engine, grow-only device buffers, on-device scans, and three
// blocking reads per max_hits chunk (total hits, survivor count, HSP count)
// where the reference does four to five.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/hspz.h"
#include "common.hpp"
#include "kernels.hpp"
#include "kernels_batch.hpp"
#include "encode_dev.hpp"

using namespace hspz;

namespace hspz {
size_t dedup_and_order(hspz_segment_pair *h, size_t n);  // host.cpp
}

#define HIP_TRY(x)                                                                       \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            return set_error(-(int)e_, "%s: %s", #x, hipGetErrorString(e_));             \
    } while (0)

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    template <class T>
    T *as() const { return (T *)p; }
};

// run.rs:278-299 batch descriptor
struct Batch {
    uint32_t interval;
    int rev;
    uint32_t lo, hi;
};

}  // namespace

struct hspz_batch_state;  // batch_host.hpp

// The target side of engines that share it (hspz_engine_share): the device
// buffers of the reference, index, pos table, planes, clean bits and reference
// records, freed when the last engine holding them is.
struct ResidentTarget {
    int device = 0;
    std::vector<void *> bufs;
    ~ResidentTarget() {
        if (bufs.empty()) return;
        (void)hipSetDevice(device);
        for (void *p : bufs) (void)hipFree(p);
    }
};

namespace {
constexpr size_t H_SMALL_BYTES = 4096;  // pinned host words for counters and chunk limits
}

struct hspz_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_gate0 = nullptr, ev_gate1 = nullptr, ev_post1 = nullptr;
    hipEvent_t ev_stage[2] = {nullptr, nullptr};  // read_to_host's staging halves
    // reference side (resident)
    DevBuf ref, index, pos, sub;
    uint32_t ref_len = 0;
    size_t index_len = 0, pos_len = 0;
    uint64_t index_build_ns = 0;  // device index build (hspz_engine_new_device_index)
    int32_t sub_host[64];
    // query side
    DevBuf qf, qr;
    uint32_t query_len = 0;
    // device encoding of raw query bytes (k_encode_segments): the encode and
    // encode-of-complement tables, the raw staging buffer, the segment list
    DevBuf enc_tabs, rawq, encseg;
    // packed planes + coarse exception bitmaps for the fast gate (k_gate_packed)
    DevBuf ref2, refx, qf2, qfx, qr2, qrx, pairtab, slow, slow_n;
    bool fast_gate = false;
    int sum_mode = 0;  // score-sum form of the merged gates (SUM_*): raw when every ACGT score is within
                       // +-G2_RAW_SCORE, 16-cell trees when xdrop and the scores are <= G2_NARROW_SCORE
    uint32_t unsafe_codes = 0xFFFFFFFFu;  // perm table: byte k = 0xFF when code 4 + k needs the exact gate
    // seed slots (gpu/mod.rs:933-945)
    DevBuf seeds[2];
    uint32_t seed_len[2] = {0, 0};
    // work buffers, grow-only (gpu/mod.rs:1539-1565)
    DevBuf hit_num, tiles, total, anchors, flags, surv, hsp, done, ids2, reduced, valid;
    uint32_t *h_small = nullptr;  // pinned scratch for counts (H_SMALL_BYTES)
    // parameters
    uint32_t seed_size = 19;
    int32_t xdrop = 910, hspthresh = 3000, noentropy = 0;
    uint32_t max_hits = 0;
    // stats
    uint64_t launches = 0, pipeline_syncs = 0, reference_uploads = 0, query_swaps = 0;
    uint64_t dev_bytes = 0, peak_dev_bytes = 0;
    uint64_t gate_ns = 0, gate_hits = 0, gate_launches = 0, gate_deferred = 0;
    // wall time of hspz_filter_batch: host layout and uploads, device pass
    // (launch to drained stream), read-back, host dedup and assembly
    uint64_t batch_prep_ns = 0, batch_device_ns = 0, batch_read_ns = 0, batch_post_ns = 0;
    uint64_t batch_reruns = 0;  // batched passes rerun after a list overflow
    uint32_t list_cap = 0;      // HSPZ_FLAG_BATCH_LIST_CAP; 0 = sized from the batch
    uint32_t batch_max_positions = 0;  // HSPZ_FLAG_BATCH_MAX_POSITIONS; 0 = from free device memory
    uint64_t batch_passes_last = 0;    // device passes of the last hspz_filter_batch call
    int coarse_gate = 1;               // HSPZ_FLAG_BATCH_COARSE: class-score bound ahead of the two-cell maps
    bool cand_planes = true;           // HSPZ_FLAG_BATCH_CAND_PLANES: candidates' windows from the planes
    bool lane_materialize = true;      // HSPZ_FLAG_BATCH_LANE_MATERIALIZE: a survivor per lane first
    int cand_bins = 1;                 // HSPZ_FLAG_BATCH_CAND_BINS: 0 off, 1 small passes, 2 every pass
    int cand_binned = 0;               // whether the last batched pass binned them
                                       // (0 off, 1 paired form where the class scores fit, 2 single-hit form)
    int coarse_used = 0;               // the form the last batched pass ran (0: none)
    bool coarse_fits16 = false;        // every |class score| <= COARSE_PAIR_MAX_CLASS
    bool coarse_pv = false;            // the set-bit count adds to the coarse bound's positive part
    CoarseCoef coarse_cc = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    bool coarse_long = true;           // HSPZ_FLAG_BATCH_COARSE_LONG
    uint64_t last_cand = 0;            // candidates of the last batched pass
    int gate_blocks = 2048;
    // batched gate configuration (hspz_set_flag, hspz_autotune_batch)
    int bgate_blocks = 0;  // persistent grid; 0 = as many blocks as are resident at once
    int bgate_step = 96, bgate_threads = 1024, bgate_claim = 256;  // the autotuner's usual pick
    int batch_order = 2;                    // 0 seed-major, 1 k-mer-major, 2 by batch size
    int batch_order_used = 0;               // order of the last batched pass
    DevBuf clean;                           // pos_table-aligned clean-window bits (k-mer-major gate)
    DevBuf rrec;                            // pos_table-ordered reference records (k_ref_records)
    int rrec_mode = 0;  // HSPZ_FLAG_REF_RECORDS: -1 never, 0 merged steps, 1 every k-mer-major gate
    bool rrec_declined = false;
    bool lookahead = true;                  // k_bgate_ahead for merged steps (needs rrec)
    bool sep_drops = false;                 // code 7 scores <= -(xdrop+1) against every code
    int bgate_grid_used = 0;                // grid of the last batched launch
    int num_cus = 0;
    double last_gate_ms = 0, last_post_ms = 0;  // event times of the last batched pass
    uint64_t last_surv = 0, last_deferred = 0;   // its survivor and deferred-hit counts
    hspz_batch_state *batch = nullptr;
    bool batch_graph = true;  // capture the batched device pass in a hipGraph and replay it
    bool device_dedup = true; // dedup_and_order of the batched path on the device
    // hspz_seed_and_filter_noalloc: this device's address of the scratch array
    // (looked up once per device at engine creation) and the test cap on its tile
    uint8_t *scratch = nullptr;
    uint32_t noalloc_tile_hits = 0;
    // asynchronous seed uploads (hspz_upload_seeds_async): a copy stream, per
    // slot the copy's completion and the last compute reading the slot
    hipStream_t copy_stream = nullptr;
    hipEvent_t copy_ready[2] = {nullptr, nullptr}, compute_done[2] = {nullptr, nullptr};
    bool copy_pending[2] = {false, false};
    uint64_t seed_uploads = 0;
    // pinned host staging pool (hspz_pinned_take / hspz_pinned_give)
    std::mutex pinned_mu;
    std::vector<std::pair<uint64_t *, size_t>> pinned_free;  // returned buffers and their capacity (u64)
    std::unordered_map<uint64_t *, size_t> pinned_out;        // handed-out buffers
    bool pinned_ok = false;                                   // a pinned buffer was handed out
    bool dump_raw = false;  // HSPZ_FLAG_DUMP_RAW: hspz_seed_and_filter keeps its pre-dedup records
    // pinned staging for reading large batched outputs (two halves, used in turn)
    uint8_t *read_stage = nullptr;
    size_t read_piece = 64u << 20, read_stage_piece = 0;  // piece size; the staging's piece size
    // the resident target when shared between engines (hspz_engine_share): it
    // owns ref, index, pos, ref2, refx, clean and rrec, freed with its last holder
    std::shared_ptr<ResidentTarget> resident;
};

namespace {

// Serializes hipGraph captures with the calls that must not overlap another
// thread's capture (allocation and release of device memory): engines of a
// serving loop capture and grow their buffers from different host threads.
std::mutex &capture_mutex() {
    static std::mutex mu;
    return mu;
}

// Grow-only device buffer: at least `bytes`, with 1/8 slack so that a
// slightly larger next call reuses it (exact = no slack, for buffers sized
// once, such as the resident reference records).
int reserve(hspz_engine *h, DevBuf &b, size_t bytes, bool exact = false) {
    if (bytes <= b.cap && b.p) return 0;
    std::lock_guard<std::mutex> lk(capture_mutex());
    size_t want = std::max<size_t>(exact ? bytes : bytes + bytes / 8, 256);
    if (b.p) {
        HIP_TRY(hipStreamSynchronize(h->stream));
        HIP_TRY(hipFree(b.p));
        h->dev_bytes -= b.cap;
        b.p = nullptr;
        b.cap = 0;
    }
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        b.p = nullptr;
        return set_error(HSPZ_E_OOM, "hipMalloc(%zu) failed: %s (the engine holds %llu bytes)", want,
                         hipGetErrorString(e), (unsigned long long)h->dev_bytes);
    }
    b.cap = want;
    h->dev_bytes += want;
    h->peak_dev_bytes = std::max(h->peak_dev_bytes, h->dev_bytes);
    return 0;
}

void release(DevBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

// release() with the engine's device-byte count kept.
void release(hspz_engine *h, DevBuf &b) {
    h->dev_bytes -= b.cap;
    release(b);
}

// Uploads a code-7-padded copy of an encoded sequence: PAD bytes either side.
int upload_seq(hspz_engine *h, DevBuf &b, const uint8_t *src, size_t len) {
    if (int rc = reserve(h, b, len + 2 * PAD)) return rc;
    HIP_TRY(hipMemsetAsync(b.p, 7, len + 2 * PAD, h->stream));
    if (len) HIP_TRY(hipMemcpyAsync(b.as<uint8_t>() + PAD, src, len, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return 0;
}

uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// 2-bit plane (16 bases per word, two spare words) and coarse exception bitmap
// (one bit per 64 bases, two spare words) of an uploaded, padded sequence.
int build_planes(hspz_engine *h, const DevBuf &seq, uint32_t len, DevBuf &plane, DevBuf &coarse) {
    uint32_t nw = ceil_div(len, 16) + 2, nc = ceil_div(len, 2048) + 2;
    if (int rc = reserve(h, plane, (size_t)nw * 4)) return rc;
    if (int rc = reserve(h, coarse, (size_t)nc * 4)) return rc;
    const uint8_t *codes = seq.as<uint8_t>() + PAD;
    hipLaunchKernelGGL(k_pack2, dim3(std::min<uint32_t>(ceil_div(nw, 256), 8192)), dim3(256), 0, h->stream, codes, len,
                       plane.as<uint32_t>(), nw);
    hipLaunchKernelGGL(k_coarse_exceptions, dim3(std::min<uint32_t>(ceil_div(nc, 8), 8192)), dim3(256), 0, h->stream,
                       codes, len, coarse.as<uint32_t>(), nc, h->unsafe_codes);
    h->launches += 2;
    HIP_TRY(hipGetLastError());
    return 0;
}

// Inclusive scan of u32[n] in place; the total lands in h->total[0].
int scan_inclusive(hspz_engine *h, uint32_t *data, uint32_t n) {
    uint32_t nt = std::max<uint32_t>(ceil_div(n, TILE), 1);
    if (int rc = reserve(h, h->tiles, (size_t)nt * 4)) return rc;
    hipLaunchKernelGGL(k_tile_sum_u32, dim3(nt), dim3(TILE_THREADS), 0, h->stream, data, n, h->tiles.as<uint32_t>());
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_THREADS), 0, h->stream, h->tiles.as<uint32_t>(), nt, h->total.as<uint32_t>());
    hipLaunchKernelGGL(k_tile_apply_u32, dim3(nt), dim3(TILE_THREADS), 0, h->stream, data, n, h->tiles.as<uint32_t>());
    h->launches += 3;
    HIP_TRY(hipGetLastError());
    return 0;
}

// Reads h->total[0] to the host (one blocking round trip).
int read_total(hspz_engine *h, uint32_t *out) {
    HIP_TRY(hipMemcpyAsync(h->h_small, h->total.p, 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->pipeline_syncs++;
    *out = h->h_small[0];
    return 0;
}

// Stable compaction of nonzero u8 flags into ids (count first, then emit).
int compact_flags(hspz_engine *h, const uint8_t *flags, uint32_t n, DevBuf &ids, uint32_t *count) {
    uint32_t nt = std::max<uint32_t>(ceil_div(n, TILE), 1);
    if (int rc = reserve(h, h->tiles, (size_t)nt * 4)) return rc;
    hipLaunchKernelGGL(k_tile_count_u8, dim3(nt), dim3(TILE_THREADS), 0, h->stream, flags, n, h->tiles.as<uint32_t>());
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_THREADS), 0, h->stream, h->tiles.as<uint32_t>(), nt, h->total.as<uint32_t>());
    h->launches += 2;
    HIP_TRY(hipGetLastError());
    if (int rc = read_total(h, count)) return rc;
    if (*count == 0) return 0;
    if (int rc = reserve(h, ids, (size_t)*count * 4)) return rc;
    hipLaunchKernelGGL(k_tile_emit_u8, dim3(nt), dim3(TILE_THREADS), 0, h->stream, flags, n, h->tiles.as<uint32_t>(), ids.as<uint32_t>());
    h->launches++;
    HIP_TRY(hipGetLastError());
    return 0;
}

// gpu/mod.rs:2397-2420, the KegAlign lower_bound walk. Returns -1 on the
// reference's underflow (a seed alone beyond max_hits at index 0).
int chunk_limits(const std::vector<uint32_t> &hit_num, uint32_t max_hits,
                 std::vector<uint32_t> &out) {
    uint32_t n = (uint32_t)hit_num.size();
    uint32_t num_iter = hit_num.back() / max_hits + 1;
    std::vector<uint32_t> limits;
    uint32_t lim = max_hits;
    for (uint32_t it = 0; it + 1 < num_iter; it++) {
        uint32_t p = (uint32_t)(std::lower_bound(hit_num.begin(), hit_num.end(), lim) - hit_num.begin());
        if (p == 0) return -1;
        limits.push_back(p - 1);
        lim = hit_num[p - 1] + max_hits;
    }
    limits.push_back(n - 1);
    uint32_t s0 = 0, h0 = 0;
    for (uint32_t p : limits) {
        out.insert(out.end(), {s0, p, h0, hit_num[p]});
        s0 = p + 1;
        h0 = hit_num[p];
    }
    return 0;
}

// The slot's pending asynchronous seed copy, ordered before the compute
// stream's next work (the copy completes before any later kernel starts).
int join_seed_copy(hspz_engine *h, int slot) {
    if (!h->copy_pending[slot]) return 0;
    HIP_TRY(hipStreamWaitEvent(h->stream, h->copy_ready[slot], 0));
    h->copy_pending[slot] = false;
    return 0;
}

// The slot's pending asynchronous seed copy, waited for on the host.
int wait_seed_copy(hspz_engine *h, int slot) {
    if (!h->copy_pending[slot]) return 0;
    HIP_TRY(hipEventSynchronize(h->copy_ready[slot]));
    h->copy_pending[slot] = false;
    return 0;
}

// Marks the compute stream's last work on the slot (an asynchronous upload
// into the slot waits for it).
int mark_slot_computed(hspz_engine *h, int slot) {
    if (h->compute_done[slot]) HIP_TRY(hipEventRecord(h->compute_done[slot], h->stream));
    return 0;
}

int seed_and_filter_impl(hspz_engine *h, int slot, int rev, std::vector<hspz_segment_pair> &res,
                         uint32_t *num_hits_out, uint32_t *raw_out, std::vector<hspz_segment_pair> *raw_list) {
    HIP_TRY(hipSetDevice(h->device));
    if (int rc = join_seed_copy(h, slot)) return rc;
    uint32_t n = h->seed_len[slot];
    *num_hits_out = 0;
    *raw_out = 0;
    if (n == 0) return 0;
    const uint8_t *q = (rev ? h->qr : h->qf).as<uint8_t>() + PAD;
    const uint8_t *r = h->ref.as<uint8_t>() + PAD;
    const uint64_t *seeds = h->seeds[slot].as<uint64_t>();

    if (int rc = reserve(h, h->hit_num, (size_t)n * 4)) return rc;
    hipLaunchKernelGGL(k_find_num_hits, dim3(std::min<uint32_t>(ceil_div(n, 256), 4096)), dim3(256), 0, h->stream,
                       seeds, n, h->index.as<uint32_t>(), h->hit_num.as<uint32_t>());
    h->launches++;
    if (int rc = scan_inclusive(h, h->hit_num.as<uint32_t>(), n)) return rc;
    uint32_t num_hits;
    if (int rc = read_total(h, &num_hits)) return rc;
    *num_hits_out = num_hits;
    if (num_hits == 0) return 0;

    std::vector<uint32_t> chunks;
    if (num_hits <= h->max_hits) {
        chunks = {0, n - 1, 0, num_hits};
    } else {
        std::vector<uint32_t> cum(n);
        HIP_TRY(hipMemcpyAsync(cum.data(), h->hit_num.p, (size_t)n * 4, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        h->pipeline_syncs++;
        if (chunk_limits(cum, h->max_hits, chunks))
            return set_error(HSPZ_E_CHUNK, "a single seed exceeds max_hits %u", h->max_hits);
    }

    for (size_t c = 0; c < chunks.size(); c += 4) {
        uint32_t s0 = chunks[c], s1 = chunks[c + 1], h0 = chunks[c + 2], h1 = chunks[c + 3];
        uint32_t nseeds = s1 + 1 - s0, nh = h1 - h0;
        if (nh == 0) continue;
        if (int rc = reserve(h, h->anchors, (size_t)nh * 8)) return rc;
        if (int rc = reserve(h, h->flags, (size_t)nh + 16)) return rc;
        // find_hits selector (gpu/mod.rs:48-54): wave per seed when dense
        if ((uint64_t)nh >= 16ull * nseeds) {
            hipLaunchKernelGGL(k_find_hits_wave, dim3(ceil_div(nseeds, 4)), dim3(256), 0, h->stream,
                               h->index.as<uint32_t>(), h->pos.as<uint32_t>(), seeds, h->seed_size,
                               h->hit_num.as<uint32_t>(), h->anchors.as<uint64_t>(), s0, h0, nseeds);
        } else {
            hipLaunchKernelGGL(k_find_hits_thread, dim3(ceil_div(nseeds, 256)), dim3(256), 0, h->stream,
                               h->index.as<uint32_t>(), h->pos.as<uint32_t>(), seeds, h->seed_size,
                               h->hit_num.as<uint32_t>(), h->anchors.as<uint64_t>(), s0, h0, nseeds);
        }
        h->launches++;
        HIP_TRY(hipEventRecord(h->ev_gate0, h->stream));
        if (h->fast_gate) {
            if (int rc = reserve(h, h->slow, (size_t)nh * 4)) return rc;
            HIP_TRY(hipMemsetAsync(h->slow_n.p, 0, 4, h->stream));
            const uint32_t *q2 = (rev ? h->qr2 : h->qf2).as<uint32_t>();
            const uint32_t *qx = (rev ? h->qrx : h->qfx).as<uint32_t>();
            uint32_t gblocks = std::max<uint32_t>(1, std::min<uint32_t>(ceil_div(nh, GATE_CHUNK * 8), h->gate_blocks));
            hipLaunchKernelGGL(k_gate_packed, dim3(gblocks), dim3(512), 0, h->stream, h->ref2.as<uint32_t>(),
                               h->refx.as<uint32_t>(), h->ref_len, q2, qx, h->query_len, h->pairtab.as<uint32_t>(),
                               h->xdrop, h->hspthresh, h->anchors.as<uint64_t>(), nh, h->flags.as<uint8_t>(),
                               h->slow.as<uint32_t>(), h->slow_n.as<uint32_t>());
            uint32_t sblocks = std::max<uint32_t>(1, std::min<uint32_t>(ceil_div(nh, 64), 2048));
            hipLaunchKernelGGL(k_gate_wave, dim3(sblocks), dim3(256), 0, h->stream, r, h->ref_len, q, h->query_len,
                               h->sub.as<int32_t>(), h->xdrop, h->hspthresh, h->anchors.as<uint64_t>(),
                               h->flags.as<uint8_t>(), h->slow.as<uint32_t>(), h->slow_n.as<uint32_t>());
            h->launches += 2;
            HIP_TRY(hipMemcpyAsync(h->h_small + 1, h->slow_n.p, 4, hipMemcpyDeviceToHost, h->stream));
        } else {
            uint32_t gblocks = std::max<uint32_t>(1, std::min<uint32_t>(ceil_div(nh, GATE_CHUNK * 4), h->gate_blocks));
            hipLaunchKernelGGL(k_gate, dim3(gblocks), dim3(256), 0, h->stream, r, h->ref_len, q, h->query_len,
                               h->sub.as<int32_t>(), h->xdrop, h->hspthresh, h->anchors.as<uint64_t>(), nh,
                               h->flags.as<uint8_t>(), (const uint32_t *)nullptr, (const uint32_t *)nullptr);
            h->launches++;
        }
        HIP_TRY(hipEventRecord(h->ev_gate1, h->stream));
        HIP_TRY(hipGetLastError());
        uint32_t nsurv;
        if (int rc = compact_flags(h, h->flags.as<uint8_t>(), nh, h->surv, &nsurv)) return rc;
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, h->ev_gate0, h->ev_gate1));
        h->gate_ns += (uint64_t)((double)ms * 1e6);
        h->gate_hits += nh;
        h->gate_launches++;
        if (h->fast_gate) h->gate_deferred += h->h_small[1];
        if (nsurv == 0) continue;
        if (int rc = reserve(h, h->hsp, (size_t)nsurv * 16)) return rc;
        if (int rc = reserve(h, h->done, (size_t)nsurv + 16)) return rc;
        hipLaunchKernelGGL(k_materialize, dim3(std::min<uint32_t>(ceil_div(nsurv, 4), 16384)), dim3(256), 0, h->stream, r, h->ref_len, q,
                           h->query_len, h->sub.as<int32_t>(), h->xdrop, h->hspthresh, h->noentropy,
                           1.3862943649291992, h->anchors.as<uint64_t>(), h->surv.as<uint32_t>(), nsurv,
                           h->hsp.as<SegmentPair>(), h->done.as<uint8_t>());
        h->launches++;
        HIP_TRY(hipGetLastError());
        uint32_t nraw;
        if (int rc = compact_flags(h, h->done.as<uint8_t>(), nsurv, h->ids2, &nraw)) return rc;
        *raw_out += nraw;
        if (nraw == 0) continue;
        if (int rc = reserve(h, h->reduced, (size_t)nraw * 16)) return rc;
        hipLaunchKernelGGL(k_gather_hsp, dim3(ceil_div(nraw, 256)), dim3(256), 0, h->stream,
                           h->hsp.as<SegmentPair>(), h->ids2.as<uint32_t>(), nraw, h->reduced.as<SegmentPair>());
        h->launches++;
        size_t base = res.size();
        res.resize(base + nraw);
        HIP_TRY(hipMemcpyAsync(res.data() + base, h->reduced.p, (size_t)nraw * 16, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        h->pipeline_syncs++;
        if (raw_list) raw_list->insert(raw_list->end(), res.begin() + base, res.end());
        size_t kept = dedup_and_order(res.data() + base, nraw);
        res.resize(base + kept);
    }
    return 0;
}

// Uploads raw bytes and runs k_encode_segments over `segs` into fwd / rc
// (device pointers); the raw staging and segment buffers are the engine's.
int encode_on_device(hspz_engine *h, const uint8_t *raw, size_t n, const std::vector<EncSeg> &segs, uint8_t *fwd,
                     uint8_t *rc) {
    if (segs.empty()) return 0;
    if (int e = reserve(h, h->rawq, std::max<size_t>(n, 1))) return e;
    if (int e = reserve(h, h->encseg, segs.size() * sizeof(EncSeg))) return e;
    HIP_TRY(hipMemcpyAsync(h->rawq.p, raw, n, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->encseg.p, segs.data(), segs.size() * sizeof(EncSeg), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_encode_segments, dim3((uint32_t)segs.size()), dim3(256), 0, h->stream,
                       h->rawq.as<uint8_t>(), h->encseg.as<EncSeg>(), h->enc_tabs.as<uint8_t>(), fwd, rc);
    h->launches++;
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace

#include "noalloc.hpp"
#include "batch_host.hpp"
#include "server.hpp"
#include "tune_host.hpp"
#include "index_dev.hpp"

namespace {

// SeedTable::build on the device (index_dev.hpp): keys, one stable radix sort
// over the k-mer bits, bucket bounds, positions. Temporaries are released.
int device_index_build(hspz_engine *h, const Shape &s, uint32_t step) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint64_t st = step, len = h->ref_len;
    const uint64_t offset = (s.size + 1) % st, start = st - offset;
    const uint64_t nsteps = (len + offset) > s.size ? (len + offset - s.size) / st : 0;
    const uint32_t nk = 1u << (2 * s.kmer_size);
    if (nsteps >= 0xFFFFFFFFull) return set_error(HSPZ_E_ARG, "more than 2^32 indexed positions");
    int rc;
    if ((rc = reserve(h, h->index, (size_t)nk * 4))) return rc;
    h->index_len = nk;
    DevBuf keys, sorted, tmp;
    auto done = [&](int code) {
        (void)hipStreamSynchronize(h->stream);
        release(h, keys);
        release(h, sorted);
        release(h, tmp);
        return code;
    };
    hipLaunchKernelGGL(k_fill_u32, dim3(4096), dim3(256), 0, h->stream, h->index.as<uint32_t>(), nk, 0u);
    uint32_t nvalid = 0;
    if (nsteps) {
        if ((rc = reserve(h, keys, nsteps * 8, true)) || (rc = reserve(h, sorted, nsteps * 8, true))) return done(rc);
        ShapeArg sa;
        memset(&sa, 0, sizeof sa);
        memcpy(sa.pos, s.pos, s.kmer_size * 4);
        sa.kmer_size = s.kmer_size;
        const dim3 grid((uint32_t)std::min<uint64_t>((nsteps + 255) / 256, 65536));
        hipLaunchKernelGGL(k_index_keys, grid, dim3(256), 0, h->stream, h->ref.as<uint8_t>() + PAD, start, st,
                           nsteps, sa, s.size, nk, keys.as<uint64_t>());
        const unsigned end_bit = 32 + 2 * s.kmer_size + 1;
        size_t bytes = 0;
        hipError_t e = rocprim::radix_sort_keys(nullptr, bytes, keys.as<uint64_t>(), sorted.as<uint64_t>(),
                                                (size_t)nsteps, 32, end_bit, h->stream);
        if (e != hipSuccess) return done(set_error(-(int)e, "index sort: %s", hipGetErrorString(e)));
        if ((rc = reserve(h, tmp, bytes + 256, true))) return done(rc);
        e = rocprim::radix_sort_keys(tmp.p, bytes, keys.as<uint64_t>(), sorted.as<uint64_t>(), (size_t)nsteps, 32,
                                     end_bit, h->stream);
        if (e != hipSuccess) return done(set_error(-(int)e, "index sort: %s", hipGetErrorString(e)));
        hipLaunchKernelGGL(k_index_bounds, grid, dim3(256), 0, h->stream, sorted.as<uint64_t>(), nsteps, nk,
                           h->index.as<uint32_t>());
        if (hipGetLastError() != hipSuccess) return done(set_error(-1, "index bounds kernel failed"));
        HIP_TRY(hipMemcpyAsync(h->h_small, h->index.as<uint32_t>() + (nk - 1), 4, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        nvalid = h->h_small[0];
    }
    if ((rc = reserve(h, h->pos, std::max<size_t>(nvalid, 1) * 4, true))) return done(rc);
    if (nvalid) {
        const dim3 grid((uint32_t)std::min<uint64_t>((nvalid + 255) / 256, 65536));
        hipLaunchKernelGGL(k_index_pos, grid, dim3(256), 0, h->stream, sorted.as<uint64_t>(), (uint64_t)nvalid,
                           h->pos.as<uint32_t>());
        if (hipGetLastError() != hipSuccess) return done(set_error(-1, "index position kernel failed"));
    }
    h->pos_len = nvalid;
    h->launches += 5;
    rc = done(0);
    h->index_build_ns += (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                             std::chrono::steady_clock::now() - t0)
                             .count();
    return rc;
}

}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" const char *hspz_last_error(void) { return error_buffer(); }

extern "C" int hspz_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

extern "C" int hspz_device_memory(int device, uint64_t *free_bytes, uint64_t *total_bytes) {
    clear_error();
    HIP_TRY(hipSetDevice(device));
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return 0;
}

namespace {
int device_index_build(hspz_engine *h, const Shape &s, uint32_t step);

// Engine::new (gpu/mod.rs:1024-1128). With `build` set, the index is built
// on the device from the reference (cfg's index arrays are ignored). With
// `shared` set, the engine reads that engine's resident target (its
// ResidentTarget holder) instead of uploading or building one: cfg's
// reference and index arrays are ignored.
int engine_create(int device, const hspz_engine_config *cfg, const Shape *build, uint32_t build_step,
                  hspz_engine **out, const hspz_engine *shared = nullptr) {
    clear_error();
    if (!cfg || !out) return set_error(HSPZ_E_ARG, "null config or output");
    *out = nullptr;
    int ndev = hspz_device_count();
    if (ndev <= 0) return set_error(HSPZ_E_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= ndev) return set_error(HSPZ_E_ARG, "device %d out of range", device);
    if (shared) {
    } else if (!build) {
        if (!cfg->index_table || cfg->index_len == 0 || (cfg->index_len & (cfg->index_len - 1)))
            return set_error(HSPZ_E_ARG, "index_table must hold 4^k entries");
        if (cfg->pos_len && !cfg->pos_table) return set_error(HSPZ_E_ARG, "null pos_table");
        if (cfg->pos_len != cfg->index_table[cfg->index_len - 1])
            return set_error(HSPZ_E_ARG, "pos_table length %zu != last index entry %u", cfg->pos_len,
                             cfg->index_table[cfg->index_len - 1]);
    } else if (build->size != cfg->seed_size) {
        return set_error(HSPZ_E_SHAPE, "seed shape span %u != seed_size %u", build->size, cfg->seed_size);
    }
    if (!cfg->sub_mat) return set_error(HSPZ_E_ARG, "sub_mat is required (64 entries)");
    if (cfg->ref_len >= 0xFFFFFFFFull - 2 * PAD) return set_error(HSPZ_E_ARG, "reference longer than u32");
    if (cfg->ref_len && !cfg->ref_seq) return set_error(HSPZ_E_ARG, "null ref_seq");
    HIP_TRY(hipSetDevice(device));
    hspz_engine *h = new hspz_engine();
    h->device = device;
    auto fail = [&](int rc) {
        hspz_engine_free(h);
        return rc;
    };
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return fail(set_error(-1, "stream"));
    if (hipEventCreate(&h->ev_gate0) != hipSuccess || hipEventCreate(&h->ev_gate1) != hipSuccess ||
        hipEventCreate(&h->ev_post1) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_stage[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_stage[1], hipEventDisableTiming) != hipSuccess)
        return fail(set_error(-1, "events"));
    if (hipHostMalloc((void **)&h->h_small, H_SMALL_BYTES, hipHostMallocDefault) != hipSuccess)
        return fail(set_error(HSPZ_E_OOM, "pinned scratch"));
    if (hipDeviceGetAttribute(&h->num_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        h->num_cus < 1)
        return fail(set_error(-1, "device attributes"));
    if (int rc = scratch_ptr(device, &h->scratch)) return fail(rc);
    {
        uint8_t tabs[512];
        const uint8_t *enc = encoding_table(), *comp = complement_table();
        for (int i = 0; i < 256; i++) tabs[i] = enc[i], tabs[256 + i] = enc[comp[i]];
        if (int rc = reserve(h, h->enc_tabs, sizeof tabs)) return fail(rc);
        if (hipMemcpy(h->enc_tabs.p, tabs, sizeof tabs, hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_error(-1, "encoding table upload failed"));
    }
    h->seed_size = cfg->seed_size;
    h->xdrop = cfg->xdrop;
    h->hspthresh = cfg->hspthresh;
    h->noentropy = cfg->noentropy;
    memcpy(h->sub_host, cfg->sub_mat, sizeof h->sub_host);
    if (cfg->max_hits) {
        h->max_hits = cfg->max_hits;
    } else {
        size_t f = 0, t = 0;
        if (hipMemGetInfo(&f, &t) != hipSuccess) return fail(set_error(-1, "hipMemGetInfo"));
        h->max_hits = hspz_default_max_hits(t);
    }
    int rc;
    if ((rc = reserve(h, h->sub, 256))) return fail(rc);
    if ((rc = reserve(h, h->total, 256))) return fail(rc);
    if (hipMemcpy(h->sub.p, cfg->sub_mat, 256, hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_error(-1, "reference upload failed"));
    if (shared) {  // the target's buffers are the holder's, never released or counted here
        h->resident = shared->resident;
        h->ref = shared->ref;
        h->index = shared->index;
        h->pos = shared->pos;
        h->ref2 = shared->ref2;
        h->refx = shared->refx;
        h->clean = shared->clean;
        h->rrec = shared->rrec;
        h->rrec_declined = !shared->rrec.p;
        h->ref_len = shared->ref_len;
        h->index_len = shared->index_len;
        h->pos_len = shared->pos_len;
    } else {
    h->ref_len = (uint32_t)cfg->ref_len;
    if ((rc = upload_seq(h, h->ref, cfg->ref_seq, cfg->ref_len))) return fail(rc);
    h->reference_uploads++;
    }
    if (shared) {
    } else if (build) {
        if ((rc = device_index_build(h, *build, build_step))) return fail(rc);
    } else {
        if ((rc = reserve(h, h->index, cfg->index_len * 4))) return fail(rc);
        if ((rc = reserve(h, h->pos, std::max<size_t>(cfg->pos_len, 1) * 4, true))) return fail(rc);
        if (hipMemcpy(h->index.p, cfg->index_table, cfg->index_len * 4, hipMemcpyHostToDevice) != hipSuccess ||
            (cfg->pos_len &&
             hipMemcpy(h->pos.p, cfg->pos_table, cfg->pos_len * 4, hipMemcpyHostToDevice) != hipSuccess))
            return fail(set_error(-1, "index upload failed"));
        h->index_len = cfg->index_len;
        h->pos_len = cfg->pos_len;
    }
    // Fast gate: two-cell map table over ACGT x ACGT pairs (scores clamped at
    // -(xdrop+1)); entry = (a = s1 + s2, b = xdrop + min(s2, 0)) as int16
    // halves, s1 the cell met first: the lower cell of the key forward, the
    // upper one in reverse.
    {
        bool ok = cfg->xdrop >= 0 && cfg->xdrop <= G2_MAX_XDROP;
        for (int i = 0; i < 4 && ok; i++)
            for (int j = 0; j < 4; j++) ok = ok && cfg->sub_mat[i * 8 + j] <= G2_MAX_SCORE;
        h->fast_gate = ok;
        bool narrow = ok && cfg->xdrop <= G2_NARROW_SCORE;
        for (int i = 0; i < 4 && narrow; i++)
            for (int j = 0; j < 4; j++) narrow = narrow && cfg->sub_mat[i * 8 + j] <= G2_NARROW_SCORE;
        bool raw = ok;
        for (int i = 0; i < 4 && raw; i++)
            for (int j = 0; j < 4; j++)
                raw = raw && std::abs(std::max(cfg->sub_mat[i * 8 + j], -(cfg->xdrop + 1))) <= G2_RAW_SCORE;
        h->sum_mode = raw ? SUM_RAW : narrow ? SUM_TREE16 : SUM_TREE8;
        {
            // the coarse bound's class scores: the best ACGT score of each class
            // (match; transition, codes xor 2; transversion, xor odd)
            int32_t um = INT32_MIN, ut = INT32_MIN, uv = INT32_MIN;
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) {
                    const int32_t v = cfg->sub_mat[i * 8 + j];
                    int32_t &u = (i ^ j) == 0 ? um : (i ^ j) == 2 ? ut : uv;
                    u = std::max(u, v);
                }
            const int32_t pm = std::max(um, 0), pt = std::max(ut, 0), pv = std::max(uv, 0);
            // over the Gray-coded words: nd differing cells, np set xor bits,
            // nt = np - nd transitions, nv = 2 nd - np transversions, 16 - nd matches
            CoarseCoef &cc = h->coarse_cc;
            cc = {16 * um, 2 * uv - ut - um, ut - uv, 16 * pm, 2 * pv - pt - pm, pt - pv, 0, 0, 0, 0, 0};
            // ten best matches' worth of class-score sum on a side marks a likely long extension
            cc.lm = std::min(std::max(10 * um, 1), 32767);
            // paired form: biases folding the constant 16 um into the counts where they exist
            cc.c0 = cc.s0;
            bool bias = false;
            for (int32_t reach = 0; reach <= 256 && !bias; reach++)  // smallest |bd| + |bp| first
                for (int32_t bd = -reach; bd <= reach && !bias; bd++)
                    for (int32_t bp : {reach - std::abs(bd), std::abs(bd) - reach})
                        if (!bias && (int64_t)cc.sd * bd + (int64_t)cc.sp * bp == (int64_t)cc.s0) {
                            cc.bd = bd;
                            cc.bp = bp;
                            bias = true;
                        }
            if (bias) cc.c0 = 0;
            cc.pc = cc.p0 - cc.pd * cc.bd - cc.pp * cc.bp;
            const int32_t a = std::max({std::abs(um), std::abs(ut), std::abs(uv)});
            h->coarse_fits16 = a <= COARSE_PAIR_MAX_CLASS && 128 * a + std::abs(cc.pc) <= 32767;
            h->coarse_pv = cc.pp != 0;
        }
        if (ok) {
            uint32_t tab[1024];
            const int32_t X = cfg->xdrop;
            // the separator forces a drop against every code (as build_sub_mat's -10 * xdrop does)
            h->sep_drops = true;
            for (int j = 0; j < 8; j++)
                h->sep_drops = h->sep_drops && cfg->sub_mat[7 * 8 + j] <= -(X + 1) && cfg->sub_mat[j * 8 + 7] <= -(X + 1);
            // non-ACGT codes that can score above -(X+1) against some code
            h->unsafe_codes = 0;
            for (int c = 4; c < 8; c++) {
                bool unsafe = false;
                for (int j = 0; j < 8; j++)
                    unsafe = unsafe || cfg->sub_mat[c * 8 + j] > -(X + 1) || cfg->sub_mat[j * 8 + c] > -(X + 1);
                if (unsafe) h->unsafe_codes |= 0xFFu << (8 * (c - 4));
            }
            for (uint32_t key = 0; key < 256; key++) {
                uint32_t ra = key & 3, rb = (key >> 2) & 3, qa = (key >> 4) & 3, qb = (key >> 6) & 3;
                int32_t lo = std::max(cfg->sub_mat[ra * 8 + qa], -(X + 1));
                int32_t hi = std::max(cfg->sub_mat[rb * 8 + qb], -(X + 1));
                int32_t a = lo + hi;
                int32_t bf = X + std::min(hi, 0), br = X + std::min(lo, 0);
                tab[key] = ((uint32_t)a & 0xFFFFu) | ((uint32_t)bf << 16);
                tab[256 + key] = ((uint32_t)a & 0xFFFFu) | ((uint32_t)br << 16);
            }
            // the same maps keyed by Gray-coded cells, for the gates that read records
            for (uint32_t key = 0; key < 256; key++) {
                const uint32_t g = key ^ ((key >> 1) & 0x55u);  // each 2-bit field through its Gray code
                tab[512 + key] = tab[g];
                tab[768 + key] = tab[256 + g];
            }
            if ((rc = reserve(h, h->pairtab, sizeof tab))) return fail(rc);
            if ((rc = reserve(h, h->slow_n, 256))) return fail(rc);
            if (hipMemcpy(h->pairtab.p, tab, sizeof tab, hipMemcpyHostToDevice) != hipSuccess)
                return fail(set_error(-1, "pair table upload failed"));
            if (!shared && (rc = build_planes(h, h->ref, h->ref_len, h->ref2, h->refx))) return fail(rc);
            if (!shared && h->pos_len) {
                if ((rc = reserve(h, h->clean, ((h->pos_len + 31) / 32 + 2) * 4))) return fail(rc);
                hipLaunchKernelGGL(k_clean_bits, dim3(8192), dim3(256), 0, h->stream, h->pos.as<uint32_t>(),
                                   (uint32_t)h->pos_len, h->refx.as<uint32_t>(), h->ref_len, h->seed_size,
                                   h->clean.as<uint32_t>());
                if (hipGetLastError() != hipSuccess) return fail(set_error(-1, "clean-bit build failed"));
            }
            if (hipStreamSynchronize(h->stream) != hipSuccess) return fail(set_error(-1, "plane build failed"));
        }
    }
    *out = h;
    return 0;
}
}  // namespace

extern "C" int hspz_engine_new(int device, const hspz_engine_config *cfg, hspz_engine **out) {
    return engine_create(device, cfg, nullptr, 0, out);
}

extern "C" int hspz_engine_new_device_index(int device, const hspz_engine_config *cfg, const char *pattern,
                                            uint32_t step, hspz_engine **out) {
    clear_error();
    Shape s;
    if (int rc = parse_shape(pattern ? pattern : "12of19", &s)) return rc;
    return engine_create(device, cfg, &s, step ? step : 1, out);
}

extern "C" int hspz_engine_share(hspz_engine *base, hspz_engine **out) {
    clear_error();
    if (!base || !out) return set_error(HSPZ_E_ARG, "null argument");
    *out = nullptr;
    HIP_TRY(hipSetDevice(base->device));
    HIP_TRY(hipStreamSynchronize(base->stream));
    if (!base->resident) {
        // the reference records are built on first use; build them now so that
        // every engine reading this target reads the one copy
        if (int rc = ensure_ref_records(base, base->rrec_mode >= 0)) return rc;
        auto rt = std::make_shared<ResidentTarget>();
        rt->device = base->device;
        for (DevBuf *b : {&base->ref, &base->index, &base->pos, &base->ref2, &base->refx, &base->clean, &base->rrec})
            if (b->p) rt->bufs.push_back(b->p);
        base->resident = rt;
    }
    hspz_engine_config cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.sub_mat = base->sub_host;
    cfg.seed_size = base->seed_size;
    cfg.xdrop = base->xdrop;
    cfg.hspthresh = base->hspthresh;
    cfg.noentropy = base->noentropy;
    cfg.max_hits = base->max_hits;
    hspz_engine *h = nullptr;
    if (int rc = engine_create(base->device, &cfg, nullptr, 0, &h, base)) return rc;
    // the same gate configuration
    h->gate_blocks = base->gate_blocks;
    h->bgate_blocks = base->bgate_blocks;
    h->bgate_step = base->bgate_step;
    h->bgate_threads = base->bgate_threads;
    h->bgate_claim = base->bgate_claim;
    h->batch_order = base->batch_order;
    h->rrec_mode = base->rrec_mode;
    h->lookahead = base->lookahead;
    h->coarse_gate = base->coarse_gate;
    h->cand_planes = base->cand_planes;
    h->lane_materialize = base->lane_materialize;
    h->coarse_long = base->coarse_long;
    h->cand_bins = base->cand_bins;
    h->batch_graph = base->batch_graph;
    h->device_dedup = base->device_dedup;
    h->list_cap = base->list_cap;
    h->batch_max_positions = base->batch_max_positions;
    *out = h;
    return 0;
}

extern "C" int hspz_engine_index(const hspz_engine *h, uint32_t *index_table, size_t index_len,
                                 uint32_t **pos_table, size_t *pos_len) {
    clear_error();
    if (!h || !index_table || !pos_table || !pos_len) return set_error(HSPZ_E_ARG, "null argument");
    if (index_len != h->index_len) return set_error(HSPZ_E_ARG, "index_table must hold %zu entries", h->index_len);
    HIP_TRY(hipSetDevice(h->device));
    uint32_t *pos = (uint32_t *)malloc(std::max<size_t>(h->pos_len, 1) * 4);
    if (!pos) return set_error(HSPZ_E_OOM, "pos_table allocation failed");
    if (hipMemcpy(index_table, h->index.p, index_len * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        (h->pos_len && hipMemcpy(pos, h->pos.p, h->pos_len * 4, hipMemcpyDeviceToHost) != hipSuccess)) {
        free(pos);
        return set_error(-1, "index read-back failed");
    }
    *pos_table = pos;
    *pos_len = h->pos_len;
    return 0;
}

extern "C" void hspz_engine_free(hspz_engine *h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->resident) {  // the target's buffers go with their last holder
        for (DevBuf *b : {&h->ref, &h->index, &h->pos, &h->ref2, &h->refx, &h->clean, &h->rrec}) *b = DevBuf();
        h->resident.reset();
    }
    DevBuf *bufs[] = {&h->ref, &h->index, &h->pos, &h->sub, &h->qf, &h->qr, &h->enc_tabs, &h->rawq, &h->encseg,
                      &h->seeds[0], &h->seeds[1],
                      &h->ref2, &h->refx, &h->qf2, &h->qfx, &h->qr2, &h->qrx, &h->pairtab, &h->slow, &h->slow_n, &h->clean, &h->rrec,
                      &h->hit_num, &h->tiles, &h->total, &h->anchors, &h->flags, &h->surv, &h->hsp,
                      &h->done, &h->ids2, &h->reduced, &h->valid};
    for (DevBuf *b : bufs) release(*b);
    if (h->batch) {
        BatchBufs &B = h->batch->b;
        DevBuf *bb[] = {&B.arena, &B.plane, &B.coarse, &B.desc, &B.poff, &B.blk_off, &B.blk_len, &B.valid, &B.ids,
                        &B.voff, &B.seeds, &B.counts, &B.prefix, &B.tiles, &B.nchunks, &B.starts, &B.bhits, &B.surv,
                        &B.slow, &B.mhsp, &B.mscope, &B.keep, &B.kids, &B.ohsp, &B.oscope, &B.ctr, &B.gcnt,
                        &B.sq, &B.skm, &B.qrec, &B.kvalo, &B.ktmp, &B.exact, &B.kkey, &B.kgaps, &B.cand,
                        &B.cand2, &B.cbins};
        for (DevBuf *b : bb) release(*b);
        hspz_batch_state *bs = h->batch;
        DevBuf *dd[] = {&bs->dd_key_a, &bs->dd_key_b, &bs->dd_id_a, &bs->dd_id_b, &bs->dd_keep, &bs->dd_kids,
                        &bs->dd_tmp};
        for (DevBuf *b : dd) release(*b);
        for (int g = 0; g < 4; g++) {
            if (h->batch->exec[g]) (void)hipGraphExecDestroy(h->batch->exec[g]);
            if (h->batch->graph[g]) (void)hipGraphDestroy(h->batch->graph[g]);
        }
        delete h->batch;
    }
    for (int sl = 0; sl < 2; sl++) {
        if (h->copy_ready[sl]) (void)hipEventDestroy(h->copy_ready[sl]);
        if (h->compute_done[sl]) (void)hipEventDestroy(h->compute_done[sl]);
    }
    if (h->copy_stream) {
        (void)hipStreamSynchronize(h->copy_stream);
        (void)hipStreamDestroy(h->copy_stream);
    }
    if (h->read_stage) (void)hipHostFree(h->read_stage);
    for (auto &b : h->pinned_free) (void)hipHostFree(b.first);
    for (auto &b : h->pinned_out) (void)hipHostFree(b.first);
    if (h->h_small) (void)hipHostFree(h->h_small);
    if (h->ev_gate0) (void)hipEventDestroy(h->ev_gate0);
    if (h->ev_gate1) (void)hipEventDestroy(h->ev_gate1);
    if (h->ev_post1) (void)hipEventDestroy(h->ev_post1);
    for (hipEvent_t e : h->ev_stage)
        if (e) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

extern "C" int hspz_swap_query(hspz_engine *h, const uint8_t *fwd, const uint8_t *rc, size_t len) {
    clear_error();
    if (!h || (len && (!fwd || !rc))) return set_error(HSPZ_E_ARG, "null argument");
    if (len >= 0xFFFFFFFFull - 2 * PAD) return set_error(HSPZ_E_ARG, "query longer than u32");
    HIP_TRY(hipSetDevice(h->device));
    if (int e = upload_seq(h, h->qf, fwd, len)) return e;
    if (int e = upload_seq(h, h->qr, rc, len)) return e;
    if (h->fast_gate) {
        if (int e = build_planes(h, h->qf, (uint32_t)len, h->qf2, h->qfx)) return e;
        if (int e = build_planes(h, h->qr, (uint32_t)len, h->qr2, h->qrx)) return e;
    }
    h->query_len = (uint32_t)len;
    h->seed_len[0] = h->seed_len[1] = 0;
    h->query_swaps++;
    return 0;
}

extern "C" int hspz_swap_query_raw(hspz_engine *h, const uint8_t *raw, size_t len) {
    clear_error();
    if (!h || (len && !raw)) return set_error(HSPZ_E_ARG, "null argument");
    if (len >= 0xFFFFFFFFull - 2 * PAD) return set_error(HSPZ_E_ARG, "query longer than u32");
    HIP_TRY(hipSetDevice(h->device));
    for (DevBuf *b : {&h->qf, &h->qr}) {
        if (int e = reserve(h, *b, len + 2 * PAD)) return e;
        HIP_TRY(hipMemsetAsync(b->p, 7, len + 2 * PAD, h->stream));
    }
    std::vector<EncSeg> segs;
    encode_segments(segs, 0, (uint32_t)len, 0, 0);
    if (int e = encode_on_device(h, raw, len, segs, h->qf.as<uint8_t>() + PAD, h->qr.as<uint8_t>() + PAD)) return e;
    if (h->fast_gate) {
        if (int e = build_planes(h, h->qf, (uint32_t)len, h->qf2, h->qfx)) return e;
        if (int e = build_planes(h, h->qr, (uint32_t)len, h->qr2, h->qrx)) return e;
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->query_len = (uint32_t)len;
    h->seed_len[0] = h->seed_len[1] = 0;
    h->query_swaps++;
    return 0;
}

extern "C" int hspz_engine_query(const hspz_engine *h, uint8_t *fwd, uint8_t *rc, size_t len) {
    clear_error();
    if (!h || (len && (!fwd || !rc))) return set_error(HSPZ_E_ARG, "null argument");
    if (len != h->query_len) return set_error(HSPZ_E_ARG, "query holds %u bases, not %zu", h->query_len, len);
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (len) {
        HIP_TRY(hipMemcpy(fwd, h->qf.as<uint8_t>() + PAD, len, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(rc, h->qr.as<uint8_t>() + PAD, len, hipMemcpyDeviceToHost));
    }
    return 0;
}

extern "C" int hspz_upload_seeds(hspz_engine *h, int slot, const uint64_t *seeds, size_t n) {
    clear_error();
    if (!h || slot < 0 || slot > 1 || (n && !seeds)) return set_error(HSPZ_E_ARG, "bad slot or seeds");
    if (n > 0xFFFFFFFFull) return set_error(HSPZ_E_ARG, "more than 2^32 seeds");
    HIP_TRY(hipSetDevice(h->device));
    for (size_t i = 0; i < n; i++)
        if ((seeds[i] >> 32) >= h->index_len) return set_error(HSPZ_E_ARG, "seed %zu has k-mer beyond the index", i);
    if (int rc = wait_seed_copy(h, slot)) return rc;
    if (int rc = reserve(h, h->seeds[slot], std::max<size_t>(n, 1) * 8)) return rc;
    if (n) HIP_TRY(hipMemcpyAsync(h->seeds[slot].p, seeds, n * 8, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->seed_len[slot] = (uint32_t)n;
    h->seed_uploads++;
    return 0;
}

extern "C" int hspz_upload_seeds_async(hspz_engine *h, int slot, const uint64_t *seeds, size_t n) {
    clear_error();
    if (!h || slot < 0 || slot > 1 || (n && !seeds)) return set_error(HSPZ_E_ARG, "bad slot or seeds");
    if (n > 0xFFFFFFFFull) return set_error(HSPZ_E_ARG, "more than 2^32 seeds");
    HIP_TRY(hipSetDevice(h->device));
    for (size_t i = 0; i < n; i++)
        if ((seeds[i] >> 32) >= h->index_len) return set_error(HSPZ_E_ARG, "seed %zu has k-mer beyond the index", i);
    if (!h->copy_stream) {
        HIP_TRY(hipStreamCreateWithFlags(&h->copy_stream, hipStreamNonBlocking));
        for (int s = 0; s < 2; s++) {
            HIP_TRY(hipEventCreateWithFlags(&h->copy_ready[s], hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&h->compute_done[s], hipEventDisableTiming));
        }
    }
    // the slot's previous copy has landed (its host array is free again), and
    // a growing slot buffer is only released once no queued kernel reads it
    if (int rc = wait_seed_copy(h, slot)) return rc;
    if (h->seeds[slot].cap < std::max<size_t>(n, 1) * 8) HIP_TRY(hipStreamSynchronize(h->stream));
    if (int rc = reserve(h, h->seeds[slot], std::max<size_t>(n, 1) * 8)) return rc;
    h->seed_len[slot] = (uint32_t)n;
    if (n == 0) return 0;
    // ordered after the compute stream's last work on the slot, so the copy
    // cannot overwrite seeds a queued kernel still reads
    HIP_TRY(hipStreamWaitEvent(h->copy_stream, h->compute_done[slot], 0));
    HIP_TRY(hipMemcpyAsync(h->seeds[slot].p, seeds, n * 8, hipMemcpyHostToDevice, h->copy_stream));
    HIP_TRY(hipEventRecord(h->copy_ready[slot], h->copy_stream));
    h->copy_pending[slot] = true;
    h->seed_uploads++;
    return 0;
}

extern "C" int hspz_sync_seeds(hspz_engine *h, int slot) {
    clear_error();
    if (!h || slot < 0 || slot > 1) return set_error(HSPZ_E_ARG, "bad slot");
    HIP_TRY(hipSetDevice(h->device));
    return wait_seed_copy(h, slot);
}

extern "C" int hspz_pinned_take(hspz_engine *h, size_t n, uint64_t **buf) {
    clear_error();
    if (!h || !buf) return set_error(HSPZ_E_ARG, "null argument");
    std::lock_guard<std::mutex> lk(h->pinned_mu);
    long best = -1;
    for (size_t i = 0; i < h->pinned_free.size(); i++)
        if (h->pinned_free[i].second >= n && (best < 0 || h->pinned_free[i].second < h->pinned_free[best].second))
            best = (long)i;
    uint64_t *p = nullptr;
    size_t cap = 0;
    if (best >= 0) {
        p = h->pinned_free[best].first;
        cap = h->pinned_free[best].second;
        h->pinned_free.erase(h->pinned_free.begin() + best);
    } else {
        HIP_TRY(hipSetDevice(h->device));
        cap = std::max<size_t>(n, 1);
        if (hipHostMalloc((void **)&p, cap * 8, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(HSPZ_E_OOM, "pinned host allocation of %zu bytes failed", cap * 8);
        }
    }
    h->pinned_out[p] = cap;
    h->pinned_ok = true;
    *buf = p;
    return 0;
}

extern "C" int hspz_pinned_give(hspz_engine *h, uint64_t *buf) {
    clear_error();
    if (!h || !buf) return set_error(HSPZ_E_ARG, "null argument");
    std::lock_guard<std::mutex> lk(h->pinned_mu);
    auto it = h->pinned_out.find(buf);
    if (it == h->pinned_out.end()) return set_error(HSPZ_E_ARG, "buffer did not come from hspz_pinned_take");
    h->pinned_free.emplace_back(buf, it->second);
    h->pinned_out.erase(it);
    return 0;
}

extern "C" int hspz_check_seed_bytes(hspz_engine *h, int slot, const uint64_t *expected, size_t n) {
    clear_error();
    if (!h || slot < 0 || slot > 1 || (n && !expected)) return set_error(HSPZ_E_ARG, "bad slot or seeds");
    HIP_TRY(hipSetDevice(h->device));
    if (int rc = wait_seed_copy(h, slot)) return rc;
    if (h->seed_len[slot] != n)
        return set_error(HSPZ_E_ARG, "device seed count differs: %u != %zu", h->seed_len[slot], n);
    if (n == 0) return 0;
    std::vector<uint64_t> got(n);
    HIP_TRY(hipMemcpyAsync(got.data(), h->seeds[slot].p, n * 8, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (size_t i = 0; i < n; i++)
        if (got[i] != expected[i])
            return set_error(HSPZ_E_ARG, "device seed differs at %zu: %016llx != %016llx", i,
                             (unsigned long long)got[i], (unsigned long long)expected[i]);
    return 0;
}

extern "C" int hspz_launch_overhead_ms(hspz_engine *h, uint32_t iters, double *ms) {
    clear_error();
    if (!h || !ms || iters == 0) return set_error(HSPZ_E_ARG, "bad argument");
    HIP_TRY(hipSetDevice(h->device));
    const uint32_t *sink = h->total.as<uint32_t>();
    uint32_t *out = h->total.as<uint32_t>() + 1;
    for (int i = 0; i < 16; i++) hipLaunchKernelGGL(k_noop, dim3(1), dim3(64), 0, h->stream, sink, out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h->stream));
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 0; i < iters; i++) hipLaunchKernelGGL(k_noop, dim3(1), dim3(64), 0, h->stream, sink, out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h->stream));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    h->launches += 16 + iters;
    *ms = secs * 1000.0 / iters;
    return 0;
}

extern "C" int hspz_generate_seeds(hspz_engine *h, int slot, int rev, uint32_t start, uint32_t end,
                                   const uint32_t *shape_pos, uint32_t kmer_size, int transitions,
                                   uint32_t *n_out) {
    clear_error();
    if (!h || slot < 0 || slot > 1 || !shape_pos || !n_out) return set_error(HSPZ_E_ARG, "bad argument");
    if (kmer_size < 4 || kmer_size > 15) return set_error(HSPZ_E_SHAPE, "kmer_size %u outside 4..=15", kmer_size);
    if (((size_t)1 << (2 * kmer_size)) != h->index_len)
        return set_error(HSPZ_E_SHAPE, "device seed shape does not match the index");
    if (shape_pos[kmer_size - 1] + 1 > h->seed_size)
        return set_error(HSPZ_E_SHAPE, "device seed shape does not match Engine seed size");
    HIP_TRY(hipSetDevice(h->device));
    if (int rc = wait_seed_copy(h, slot)) return rc;  // the slot's buffer is rewritten (and may grow)
    *n_out = 0;
    h->seed_len[slot] = 0;
    uint32_t span = end > start ? end - start : 0;
    if (span == 0) return 0;
    if ((uint64_t)end + h->seed_size - 1 > h->query_len)
        return set_error(HSPZ_E_RANGE, "device seed range [%u, %u) exceeds query length %u", start, end, h->query_len);
    const uint8_t *q = (rev ? h->qr : h->qf).as<uint8_t>() + PAD;
    if (int rc = reserve(h, h->valid, (size_t)span + 16)) return rc;
    hipLaunchKernelGGL(k_seed_valid, dim3(std::min<uint32_t>(ceil_div(span, 256), 4096)), dim3(256), 0, h->stream,
                       q, start, span, h->seed_size, h->valid.as<uint8_t>());
    h->launches++;
    uint32_t nvalid;
    if (int rc = compact_flags(h, h->valid.as<uint8_t>(), span, h->ids2, &nvalid)) return rc;
    uint32_t per = transitions ? 1 + kmer_size : 1;
    uint64_t nseeds = (uint64_t)nvalid * per;
    if (nseeds > 0xFFFFFFFFull) return set_error(HSPZ_E_ARG, "more than 2^32 seeds");
    if (nvalid) {
        if (int rc = reserve(h, h->seeds[slot], nseeds * 8)) return rc;
        ShapeArg sa;
        memset(&sa, 0, sizeof sa);
        memcpy(sa.pos, shape_pos, kmer_size * 4);
        sa.kmer_size = kmer_size;
        hipLaunchKernelGGL(k_seed_scatter, dim3(ceil_div(nvalid, 256)), dim3(256), 0, h->stream, q, start,
                           h->ids2.as<uint32_t>(), nvalid, sa, transitions, h->seeds[slot].as<uint64_t>());
        h->launches++;
        HIP_TRY(hipGetLastError());
    }
    h->seed_len[slot] = (uint32_t)nseeds;
    *n_out = (uint32_t)nseeds;
    return mark_slot_computed(h, slot);
}

extern "C" int hspz_seed_and_filter(hspz_engine *h, int slot, int rev, hspz_filter_output *out) {
    clear_error();
    if (!h || !out || slot < 0 || slot > 1) return set_error(HSPZ_E_ARG, "bad argument");
    memset(out, 0, sizeof *out);
    if (h->query_len == 0)
        return set_error(HSPZ_E_NO_QUERY, "Engine has no query: swap_query() must be called before seed_and_filter()");
    std::vector<hspz_segment_pair> res, raw_list;
    uint32_t nh = 0, raw = 0;
    int rc = seed_and_filter_impl(h, slot, rev, res, &nh, &raw, h->dump_raw ? &raw_list : nullptr);
    if (rc == 0) rc = mark_slot_computed(h, slot);
    if (rc) return rc;
    out->num_hits = nh;
    out->raw_hsps = raw;
    out->n_hsps = res.size();
    if (!res.empty()) {
        out->hsps = (hspz_segment_pair *)aligned_alloc(16, ((res.size() * 16 + 15) / 16) * 16);
        if (!out->hsps) return set_error(HSPZ_E_OOM, "host output allocation");
        memcpy(out->hsps, res.data(), res.size() * 16);
    }
    if (!raw_list.empty()) {
        out->raw = (hspz_segment_pair *)aligned_alloc(16, raw_list.size() * 16);
        if (!out->raw) {
            hspz_filter_output_free(out);
            return set_error(HSPZ_E_OOM, "host output allocation");
        }
        memcpy(out->raw, raw_list.data(), raw_list.size() * 16);
        out->n_raw = raw_list.size();
    }
    return 0;
}

extern "C" void hspz_filter_output_free(hspz_filter_output *out) {
    if (!out) return;
    free(out->hsps);
    free(out->raw);
    out->hsps = out->raw = nullptr;
    out->n_hsps = out->n_raw = 0;
}

extern "C" int hspz_set_flag(hspz_engine *h, int which, int value) {
    clear_error();
    if (!h) return set_error(HSPZ_E_ARG, "null engine");
    switch (which) {
    case HSPZ_FLAG_BATCH_GRAPH: h->batch_graph = value != 0; break;
    case HSPZ_FLAG_BATCH_GATE_BLOCKS:
        if (value < 0 || value > 65536) return set_error(HSPZ_E_ARG, "gate blocks must be in [0, 65536]");
        h->bgate_blocks = value;
        break;
    case HSPZ_FLAG_BATCH_STEP_CELLS:
        if (value != 16 && value != 32 && value != 64 && value != 96 && value != 128)
            return set_error(HSPZ_E_ARG, "step cells must be 16, 32, 64, 96 or 128");
        h->bgate_step = value;
        break;
    case HSPZ_FLAG_BATCH_BLOCK_THREADS:
        if (value != 512 && value != 1024) return set_error(HSPZ_E_ARG, "block threads must be 512 or 1024");
        h->bgate_threads = value;
        break;
    case HSPZ_FLAG_BATCH_CLAIM_SEEDS:
        if (value < 1 || value > 4096) return set_error(HSPZ_E_ARG, "claim seeds must be in [1, 4096]");
        h->bgate_claim = value;
        break;
    case HSPZ_FLAG_BATCH_ORDER:
        if (value < 0 || value > 2) return set_error(HSPZ_E_ARG, "batch order must be 0, 1 or 2");
        h->batch_order = value;
        break;
    case HSPZ_FLAG_BATCH_DEVICE_DEDUP: h->device_dedup = value != 0; break;
    case HSPZ_FLAG_BATCH_LOOKAHEAD: h->lookahead = value != 0; break;
    case HSPZ_FLAG_BATCH_LIST_CAP:
        if (value != 0 && (value < 64 || value > (1 << 28)))
            return set_error(HSPZ_E_ARG, "list capacity must be 0 or in [64, 2^28]");
        h->list_cap = (uint32_t)value;
        break;
    case HSPZ_FLAG_DUMP_RAW: h->dump_raw = value != 0; break;
    case HSPZ_FLAG_BATCH_CAND_PLANES: h->cand_planes = value != 0; break;
    case HSPZ_FLAG_BATCH_LANE_MATERIALIZE: h->lane_materialize = value != 0; break;
    case HSPZ_FLAG_BATCH_COARSE_LONG: h->coarse_long = value != 0; break;
    case HSPZ_FLAG_BATCH_CAND_BINS:
        if (value < 0 || value > 2) return set_error(HSPZ_E_ARG, "candidate binning must be 0, 1 or 2");
        h->cand_bins = value;
        break;
    case HSPZ_FLAG_BATCH_COARSE:
        if (value < 0 || value > 3) return set_error(HSPZ_E_ARG, "coarse mode must be 0, 1, 2 or 3");
        h->coarse_gate = value;
        break;
    case HSPZ_FLAG_BATCH_MAX_POSITIONS:
        if (value < 0) return set_error(HSPZ_E_ARG, "max positions must be 0 (automatic) or positive");
        h->batch_max_positions = (uint32_t)value;
        break;
    case HSPZ_FLAG_NOALLOC_TILE_HITS:
        if (value != 0 && value < 64) return set_error(HSPZ_E_ARG, "noalloc tile hits must be 0 or at least 64");
        h->noalloc_tile_hits = (uint32_t)value;
        break;
    case HSPZ_FLAG_REF_RECORDS:
        if (value < -1 || value > 1) return set_error(HSPZ_E_ARG, "ref_records must be -1, 0 or 1");
        h->rrec_mode = value;
        if (value <= 0 && h->rrec.p && !h->resident) {  // a shared target keeps its records
            (void)hipStreamSynchronize(h->stream);
            std::lock_guard<std::mutex> lk(capture_mutex());
            release(h, h->rrec);
        }
        h->rrec_declined = false;
        break;
    default: return set_error(HSPZ_E_ARG, "unknown flag %d", which);
    }
    return 0;
}

extern "C" int hspz_get_flag(const hspz_engine *h, int which, int *value) {
    clear_error();
    if (!h || !value) return set_error(HSPZ_E_ARG, "null argument");
    switch (which) {
    case HSPZ_FLAG_BATCH_GRAPH: *value = h->batch_graph ? 1 : 0; break;
    case HSPZ_FLAG_BATCH_GATE_BLOCKS: *value = h->bgate_blocks ? h->bgate_blocks : h->bgate_grid_used; break;
    case HSPZ_FLAG_BATCH_STEP_CELLS: *value = h->bgate_step; break;
    case HSPZ_FLAG_BATCH_BLOCK_THREADS: *value = h->bgate_threads; break;
    case HSPZ_FLAG_BATCH_CLAIM_SEEDS: *value = h->bgate_claim; break;
    case HSPZ_FLAG_BATCH_ORDER: *value = h->batch_order; break;
    case HSPZ_FLAG_BATCH_ORDER_USED: *value = h->batch_order_used; break;
    case HSPZ_FLAG_BATCH_DEVICE_DEDUP: *value = h->device_dedup ? 1 : 0; break;
    case HSPZ_FLAG_REF_RECORDS: *value = h->rrec.p ? 2 : h->rrec_mode; break;
    case HSPZ_FLAG_BATCH_LOOKAHEAD: *value = h->lookahead ? 1 : 0; break;
    case HSPZ_FLAG_BATCH_LIST_CAP: *value = (int)h->list_cap; break;
    case HSPZ_FLAG_NOALLOC_TILE_HITS: *value = (int)h->noalloc_tile_hits; break;
    case HSPZ_FLAG_BATCH_MAX_POSITIONS: *value = (int)h->batch_max_positions; break;
    case HSPZ_FLAG_BATCH_COARSE: *value = h->coarse_gate; break;
    case HSPZ_FLAG_DUMP_RAW: *value = h->dump_raw ? 1 : 0; break;
    case HSPZ_FLAG_BATCH_CAND_PLANES: *value = h->cand_planes ? 1 : 0; break;
    case HSPZ_FLAG_BATCH_LANE_MATERIALIZE: *value = h->lane_materialize ? 1 : 0; break;
    case HSPZ_FLAG_BATCH_CAND_BINS: *value = h->cand_bins; break;
    case HSPZ_FLAG_BATCH_CAND_BINNED: *value = h->cand_binned; break;
    case HSPZ_FLAG_BATCH_COARSE_LONG: *value = h->coarse_long ? 1 : 0; break;
    case HSPZ_FLAG_BATCH_COARSE_USED: *value = h->coarse_used; break;
    default: return set_error(HSPZ_E_ARG, "unknown flag %d", which);
    }
    return 0;
}

extern "C" int hspz_get_stat(const hspz_engine *h, int which, uint64_t *value) {
    clear_error();
    if (!h || !value) return set_error(HSPZ_E_ARG, "null argument");
    switch (which) {
    case HSPZ_STAT_MAX_HITS: *value = h->max_hits; break;
    case HSPZ_STAT_LAUNCHES: *value = h->launches; break;
    case HSPZ_STAT_PIPELINE_SYNCS: *value = h->pipeline_syncs; break;
    case HSPZ_STAT_REFERENCE_UPLOADS: *value = h->reference_uploads; break;
    case HSPZ_STAT_QUERY_SWAPS: *value = h->query_swaps; break;
    case HSPZ_STAT_PEAK_DEVICE_BYTES: *value = h->peak_dev_bytes; break;
    case HSPZ_STAT_GATE_NS: *value = h->gate_ns; break;
    case HSPZ_STAT_GATE_HITS: *value = h->gate_hits; break;
    case HSPZ_STAT_GATE_LAUNCHES: *value = h->gate_launches; break;
    case HSPZ_STAT_GATE_DEFERRED: *value = h->gate_deferred; break;
    case HSPZ_STAT_GRAPH_CAPTURES: *value = h->batch ? h->batch->graph_captures : 0; break;
    case HSPZ_STAT_GRAPH_REPLAYS: *value = h->batch ? h->batch->graph_replays : 0; break;
    case HSPZ_STAT_BATCH_PREP_NS: *value = h->batch_prep_ns; break;
    case HSPZ_STAT_BATCH_DEVICE_NS: *value = h->batch_device_ns; break;
    case HSPZ_STAT_BATCH_READ_NS: *value = h->batch_read_ns; break;
    case HSPZ_STAT_BATCH_POST_NS: *value = h->batch_post_ns; break;
    case HSPZ_STAT_INDEX_BUILD_NS: *value = h->index_build_ns; break;
    case HSPZ_STAT_POS_LEN: *value = h->pos_len; break;
    case HSPZ_STAT_INDEX_LEN: *value = h->index_len; break;
    case HSPZ_STAT_BATCH_RERUNS: *value = h->batch_reruns; break;
    case HSPZ_STAT_BATCH_PASSES: *value = h->batch_passes_last; break;
    case HSPZ_STAT_DEVICE_BYTES: *value = h->dev_bytes; break;
    case HSPZ_STAT_BATCH_CANDIDATES: *value = h->last_cand; break;
    case HSPZ_STAT_SEED_UPLOADS: *value = h->seed_uploads; break;
    case HSPZ_STAT_PINNED_ACTIVE: *value = h->pinned_ok ? 1 : 0; break;
    default: return set_error(HSPZ_E_ARG, "unknown stat %d", which);
    }
    return 0;
}

// seed_and_filter_all (run.rs:266-495) over the engine's current query:
// batches per interval (plus chunks, then minus chunks over the mirrored
// range), device seeding, results concatenated per interval in batch order.
extern "C" int hspz_run_pass(hspz_engine *h, const hspz_pass_params *p, hspz_segment_pair **hsps,
                             size_t **counts, size_t *n_intervals, hspz_pass_stats *stats) {
    clear_error();
    if (!h || !p || !hsps || !counts || !n_intervals) return set_error(HSPZ_E_ARG, "null argument");
    *hsps = nullptr;
    *counts = nullptr;
    *n_intervals = 0;
    // no swap_query yet is an error; a swapped empty query has no intervals
    // and gives an empty pass, as seed_and_filter_all does (run.rs:278-300)
    if (h->query_swaps == 0) return set_error(HSPZ_E_NO_QUERY, "Engine has no query");
    Shape s;
    if (int rc = parse_shape(p->pattern ? p->pattern : "12of19", &s)) return rc;
    if (s.size != h->seed_size) return set_error(HSPZ_E_SHAPE, "seed shape span %u != engine seed size %u", s.size, h->seed_size);
    if (p->wga_chunk == 0 || p->interval_size == 0) return set_error(HSPZ_E_ARG, "zero chunk or interval size");
    uint32_t bl = h->query_len;
    std::vector<std::pair<uint32_t, uint32_t>> iv;
    if (bl > s.size) {
        uint32_t end_pos = bl - s.size;
        for (uint64_t c = 0; c < end_pos; c += p->interval_size)
            iv.push_back({(uint32_t)c, (uint32_t)std::min<uint64_t>(end_pos, c + p->interval_size)});
    }
    uint32_t qbl = bl > s.size ? bl - s.size : 0;
    std::vector<Batch> batches;
    auto add_chunks = [&](uint32_t i, int rev, uint32_t a, uint32_t b) {
        for (uint64_t x = a; x < b; x += p->wga_chunk)
            batches.push_back({i, rev, (uint32_t)x, (uint32_t)std::min<uint64_t>(x + p->wga_chunk, (uint64_t)b + 1)});
    };
    for (uint32_t i = 0; i < iv.size(); i++) {
        if (p->plus) add_chunks(i, 0, iv[i].first, iv[i].second);
        if (p->minus) add_chunks(i, 1, qbl - iv[i].second, qbl - iv[i].first);
    }
    std::vector<std::vector<hspz_segment_pair>> fw(iv.size()), rc(iv.size());
    hspz_pass_stats st = {0, 0, 0, 0, 0};
    for (const Batch &b : batches) {
        uint32_t ns = 0;
        if (int e = hspz_generate_seeds(h, 0, b.rev, b.lo, b.hi, s.pos, s.kmer_size, p->transitions, &ns)) return e;
        st.batches++;
        if (ns == 0) continue;
        st.seeds += ns;
        std::vector<hspz_segment_pair> res;
        uint32_t nh = 0, raw = 0;
        if (int e = seed_and_filter_impl(h, 0, b.rev, res, &nh, &raw, nullptr)) return e;
        st.hits += nh;
        st.raw_hsps += raw;
        auto &dst = b.rev ? rc[b.interval] : fw[b.interval];
        dst.insert(dst.end(), res.begin(), res.end());
    }
    size_t total = 0;
    for (size_t i = 0; i < iv.size(); i++) total += fw[i].size() + rc[i].size();
    st.hsps = total;
    hspz_segment_pair *all = (hspz_segment_pair *)aligned_alloc(16, std::max<size_t>(total, 1) * 16);
    size_t *cnt = (size_t *)malloc(std::max<size_t>(iv.size(), 1) * 2 * sizeof(size_t));
    if (!all || !cnt) {
        free(all);
        free(cnt);
        return set_error(HSPZ_E_OOM, "host output allocation");
    }
    size_t at = 0;
    for (size_t i = 0; i < iv.size(); i++) {
        memcpy(all + at, fw[i].data(), fw[i].size() * 16);
        at += fw[i].size();
        memcpy(all + at, rc[i].data(), rc[i].size() * 16);
        at += rc[i].size();
        cnt[2 * i] = fw[i].size();
        cnt[2 * i + 1] = rc[i].size();
    }
    *hsps = all;
    *counts = cnt;
    *n_intervals = iv.size();
    if (stats) *stats = st;
    return 0;
}

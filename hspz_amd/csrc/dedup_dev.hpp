// dedup_dev.hpp — dedup_and_order (hsp.rs:94-126) on the device, for the
// batched path's scope-tagged HSP list (included by batch_host.hpp).
//
// Per scope the reference sorts by (diagonal ref - query as u32, ref_start,
// len ascending, score descending), drops every record contained in the
// record before it in that order (same diagonal, nested [start, start + len]
// intervals, u32 arithmetic), and sorts the rest by (query_start, ref_start,
// len ascending, score descending). Here every scope is sorted at once: the
// scope is the most significant key, and each multi-field order is built from
// stable LSD radix sorts of a permutation (least significant field first).
// Records equal in every key field are identical, so the result equals the
// reference's stable sorts. The sorts run eagerly on the engine's stream after
// the captured pass (rocPRIM issues its own fills and needs the host count).
#pragma once

#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

namespace hspz {

// score descending as an ascending u32 key
__device__ __forceinline__ uint32_t desc_score_key(int32_t s) { return ~((uint32_t)s ^ 0x80000000u); }

// keys of the chosen order's least significant fields: (len << 32) | descending score
__global__ void k_dd_key_len_score(const SegmentPair *__restrict__ hsp, const uint32_t *__restrict__ perm,
                                   uint32_t n, uint64_t *__restrict__ key, uint32_t *__restrict__ ids) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t r = perm ? perm[i] : i;
        SegmentPair p = hsp[r];
        key[i] = ((uint64_t)p.len << 32) | desc_score_key(p.score);
        ids[i] = r;
    }
}

// key = ref_start of the permuted record
__global__ void k_dd_key_ref(const SegmentPair *__restrict__ hsp, const uint32_t *__restrict__ perm, uint32_t n,
                             uint32_t *__restrict__ key) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        key[i] = hsp[perm[i]].ref_start;
}

// key = (scope << 32) | diagonal (first order) or query_start (second order)
__global__ void k_dd_key_scope(const SegmentPair *__restrict__ hsp, const uint32_t *__restrict__ scope,
                               const uint32_t *__restrict__ perm, uint32_t n, int by_diagonal,
                               uint64_t *__restrict__ key) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t r = perm[i];
        SegmentPair p = hsp[r];
        uint32_t lo = by_diagonal ? p.ref_start - p.query_start : p.query_start;
        key[i] = ((uint64_t)scope[r] << 32) | lo;
    }
}

// keep[i] = 0 when record i (sorted order) is contained in record i - 1 of
// the same scope (the previous input record, kept or not).
__global__ void k_dd_mark(const SegmentPair *__restrict__ hsp, const uint32_t *__restrict__ scope,
                          const uint32_t *__restrict__ perm, uint32_t n, uint8_t *__restrict__ keep) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t r = perm[i];
        bool k = true;
        if (i > 0) {
            uint32_t q = perm[i - 1];
            if (scope[q] == scope[r]) {
                SegmentPair x = hsp[q], y = hsp[r];
                if (x.ref_start - x.query_start == y.ref_start - y.query_start) {
                    uint32_t xe = x.ref_start + x.len, ye = y.ref_start + y.len;
                    k = !((x.ref_start >= y.ref_start && xe <= ye) || (y.ref_start >= x.ref_start && ye <= xe));
                }
            }
        }
        keep[i] = k ? 1 : 0;
    }
}

// out[i] = perm[ids[i]] (compacted positions back to record indices)
__global__ void k_dd_compose(const uint32_t *__restrict__ perm, const uint32_t *__restrict__ ids, uint32_t n,
                             uint32_t *__restrict__ out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = perm[ids[i]];
}

__global__ void k_dd_gather(const SegmentPair *__restrict__ hsp, const uint32_t *__restrict__ scope,
                            const uint32_t *__restrict__ perm, uint32_t n, SegmentPair *__restrict__ out,
                            uint32_t *__restrict__ out_scope) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t r = perm[i];
        out[i] = hsp[r];
        out_scope[i] = scope[r];
    }
}

}  // namespace hspz

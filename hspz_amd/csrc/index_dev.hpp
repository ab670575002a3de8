// index_dev.hpp — SeedTable::build (seed.rs:126-373) on the device, for an
// engine whose reference index is built where it stays resident (included by
// engine.hip after the engine struct and buffer helpers).
//
// The reference's counting sort indexes positions p = start_offset + i * step
// (for step 1, p in [1, len - span]: position 0 is never indexed) whose span
// holds only uppercase ACGT, groups them by k-mer with positions ascending
// inside a bucket, and stores the inclusive cumulative bucket sizes. Here each
// position becomes the key (kmer << 32) | p, an invalid window taking the
// bucket 4^k; a stable radix sort over the k-mer bits only keeps the positions
// of a bucket in generation (ascending) order, so the sorted low words are the
// reference's pos_table, and the bucket ends give index_table. Works on the
// encoded sequence: uppercase ACGT are exactly the codes 0..3.
#pragma once

namespace hspz {

__global__ void k_index_keys(const uint8_t *__restrict__ ref, uint64_t start, uint64_t step, uint64_t nsteps,
                             ShapeArg shape, uint32_t span, uint32_t nk, uint64_t *__restrict__ keys) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nsteps;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t p = start + i * step;
        const uint8_t *w = ref + p;
        bool ok = true;
        for (uint32_t t = 0; t < span; t++) ok &= w[t] < 4;
        uint32_t kmer = 0;
        for (uint32_t t = 0; t < shape.kmer_size; t++) kmer = (kmer << 2) + (w[shape.pos[t]] & 3u);
        keys[i] = ((uint64_t)(ok ? kmer : nk) << 32) | (uint32_t)p;
    }
}

// index[k] = number of indexed positions with k-mer <= k, written by the last
// entry of each bucket for its k-mer and the empty k-mers after it (index is
// zero beforehand, which covers the k-mers before the first bucket).
__global__ void k_index_bounds(const uint64_t *__restrict__ sorted, uint64_t n, uint32_t nk,
                               uint32_t *__restrict__ index) {
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t kj = (uint32_t)(sorted[j] >> 32);
        if (kj >= nk) continue;
        uint32_t kn = j + 1 < n ? min((uint32_t)(sorted[j + 1] >> 32), nk) : nk;
        for (uint32_t k = kj; k < kn; k++) index[k] = (uint32_t)(j + 1);
    }
}

__global__ void k_index_pos(const uint64_t *__restrict__ sorted, uint64_t n, uint32_t *__restrict__ pos) {
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x)
        pos[j] = (uint32_t)sorted[j];
}

}  // namespace hspz

// encode_dev.hpp — raw query bytes to the engine's two encoded strands on the
// device (included by engine.hip): the host's per-bin "reverse complement +
// encode" (run.rs:92-97: sequence.rs:212-230 then :66-79) as one pass over one
// upload of the raw forward bytes.
//
// A segment maps raw bytes [src, src + n) of a block of length len to the
// forward codes at [dst_f, dst_f + n) and to the reverse-complement codes
// ending at dst_r (exclusive, = the rc block start + len - offset in block):
// rc[len - 1 - i] = encode(complement(raw[i])). The two 256-entry tables
// (encode, encode o complement) sit in LDS. HBM bytes per base: 1 read, 2
// written.
#pragma once

namespace hspz {

struct EncSeg {
    uint32_t src, n, dst_f, dst_r;
};
constexpr uint32_t ENC_SEG = 16384;  // bytes per segment (one block each)

__global__ void __launch_bounds__(256) k_encode_segments(const uint8_t *__restrict__ raw,
                                                         const EncSeg *__restrict__ segs,
                                                         const uint8_t *__restrict__ tabs,
                                                         uint8_t *__restrict__ fwd, uint8_t *__restrict__ rc) {
    __shared__ uint8_t tf[256], tr[256];
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) {
        tf[i] = tabs[i];
        tr[i] = tabs[256 + i];
    }
    __syncthreads();
    const EncSeg s = segs[blockIdx.x];
    for (uint32_t k = threadIdx.x; k < s.n; k += blockDim.x) {
        uint8_t x = raw[s.src + k];
        fwd[s.dst_f + k] = tf[x];
        rc[s.dst_r - 1 - k] = tr[x];
    }
}

// Segments of one raw block [src, src + len) whose strands start at f0 / r0.
inline void encode_segments(std::vector<EncSeg> &out, uint32_t src, uint32_t len, uint32_t f0, uint32_t r0) {
    for (uint32_t a = 0; a < len; a += ENC_SEG) {
        uint32_t n = std::min<uint32_t>(ENC_SEG, len - a);
        out.push_back({src + a, n, f0 + a, r0 + len - a});
    }
}

}  // namespace hspz

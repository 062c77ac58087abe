// Philox-4x32-10 counter-based RNG (Salmon, Moraes, Dror, Shaw, SC'11) for per-ray streams.
//
// Stream of ray r under 64-bit seed s: key = (s_lo, s_hi), counter = (block, r, lane, 0); each
// block yields four 32-bit words consumed in order. Any ray can be replayed from (s, r) alone,
// so a launch needs no RNG state in memory and any ray's path can be recomputed later.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace petch {

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
    uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// The four words of one Philox block, counter (blk, ray, lane, 0) under key `seed`. A stream
// that draws a fixed number of words per event can take event e's words from block e directly
// (no buffered state between events).
__device__ __forceinline__ void philox_block(uint64_t seed, uint32_t ray, uint32_t blk,
                                             uint32_t lane, uint32_t w[4]) {
  w[0] = blk;
  w[1] = ray;
  w[2] = lane;
  w[3] = 0u;
  philox4x32_10(w, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// uniform in [0, 1) with 24 random bits of the word
__device__ __forceinline__ float u24(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-8f; }

struct Rng {
  uint32_t k0, k1, ray, lane, blk;
  uint32_t w[4];
  int pos;

  __device__ __forceinline__ void init(uint64_t seed, uint32_t ray_id, uint32_t lane_id = 0) {
    k0 = (uint32_t)seed;
    k1 = (uint32_t)(seed >> 32);
    ray = ray_id;
    lane = lane_id;
    blk = 0;
    pos = 4;
  }
  __device__ __forceinline__ uint32_t next() {
    if (pos == 4) {
      w[0] = blk++;
      w[1] = ray;
      w[2] = lane;
      w[3] = 0u;
      philox4x32_10(w, k0, k1);
      pos = 0;
    }
    uint32_t v = (pos == 0) ? w[0] : (pos == 1) ? w[1] : (pos == 2) ? w[2] : w[3];
    ++pos;
    return v;
  }
  // uniform in [0, 1) with 24 random bits
  __device__ __forceinline__ float uniform() { return u24(next()); }
  // standard normal by Box-Muller from two uniforms (the first mapped to (0, 1])
  __device__ __forceinline__ float normal() {
    float u1 = (float)((next() >> 8) + 1u) * 5.9604644775390625e-8f;
    float u2 = uniform();
    return sqrtf(-2.f * logf(u1)) * cosf(6.28318530717958648f * u2);
  }
};

}  // namespace petch

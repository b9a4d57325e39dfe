// philox.hpp -- Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11), usable on host and device.
//
// Spike draws every random number from a (key, counter) pair instead of a per-thread
// state array (the reference keeps 8192 XORWOW states, CUDA RandomStateManager.cu:14-50),
// so a draw depends only on (seed, timestep, neuron id, stream) and never on the launch
// geometry or on which lane computes it.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define SPK_HD __host__ __device__ __forceinline__
#else
#define SPK_HD inline
#endif

namespace spk {

enum : uint32_t { kStreamInput = 1u, kStreamExternal = 2u };

struct u32x4 { uint32_t x, y, z, w; };

SPK_HD uint32_t mulhi32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umulhi(a, b);
#else
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
#endif
}

SPK_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = mulhi32(M0, c.x), lo0 = M0 * c.x;
    const uint32_t hi1 = mulhi32(M1, c.z), lo1 = M1 * c.z;
    u32x4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// One call yields four draws: draw n of (id, stream) is output n mod 4 of the block with
// counter (n / 4 as 64 bits, id, stream) and key = seed.
SPK_HD u32x4 draw_block(uint64_t seed, int64_t block, uint32_t id, uint32_t stream) {
  u32x4 c;
  c.x = (uint32_t)(uint64_t)block;
  c.y = (uint32_t)((uint64_t)block >> 32);
  c.z = id;
  c.w = stream;
  return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

SPK_HD uint32_t pick(const u32x4& r, int c) { return c == 0 ? r.x : (c == 1 ? r.y : (c == 2 ? r.z : r.w)); }

// Draw n (n >= 0) of (seed, id, stream).
SPK_HD uint32_t draw_u32(uint64_t seed, int64_t n, uint32_t id, uint32_t stream) {
  return pick(draw_block(seed, n >> 2, id, stream), (int)(n & 3));
}

}  // namespace spk

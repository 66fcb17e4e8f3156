// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror & Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11), usable on host and device, and
// the Box-Muller transform used for the stochastic model error.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpo {

struct U4 {
  uint32_t x, y, z, w;
};

// One call: counter (c0..c3), key (k0, k1), ten rounds with key schedule.
__host__ __device__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    // one 32 x 32 -> 64-bit product gives both halves
    const uint64_t p0 = (uint64_t)M0 * (uint64_t)c.x, p1 = (uint64_t)M1 * (uint64_t)c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    U4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
  }
  return c;
}

// Standard normal pair from two 32-bit words: uniforms at bin centres of a
// 2^-24 grid (never 0 or 1), r = sqrt(-2 ln u1), angle 2 pi u2.  The
// transcendentals are the hardware ones: v_log_f32 (log2, u1 is a normal
// number), v_sqrt_f32, and v_sin_f32 / v_cos_f32, whose argument is in
// revolutions, so u2 itself is the angle 2 pi u2.  Their errors are a few
// ulp, far below what the statistics of the field can show.
__device__ inline void box_muller(uint32_t a, uint32_t b, float* z0, float* z1) {
  const float u1 = (float)(a >> 8) * 0x1p-24f + 0x1p-25f;
  const float u2 = (float)(b >> 8) * 0x1p-24f + 0x1p-25f;
  constexpr float kMinus2Ln2 = -1.38629436f;  // -2 ln 2: ln u1 = ln 2 * log2 u1
  const float r = __builtin_amdgcn_sqrtf(kMinus2Ln2 * __builtin_amdgcn_logf(u1));
  *z0 = r * __builtin_amdgcn_cosf(u2);
  *z1 = r * __builtin_amdgcn_sinf(u2);
}

// The normal number of element `index` of draw `draw` in stream `stream`:
// one Philox call per group of four elements.
constexpr uint32_t kNoiseTag = 0x534F4152u;  // "SOAR"
__device__ inline float normal_at(uint32_t index, uint32_t draw, uint32_t stream, uint32_t k0, uint32_t k1) {
  const U4 r = philox4x32_10(U4{index >> 2, draw, stream, kNoiseTag}, k0, k1);
  float a, b;
  if ((index & 2u) == 0u) {
    box_muller(r.x, r.y, &a, &b);
  } else {
    box_muller(r.z, r.w, &a, &b);
  }
  return (index & 1u) ? b : a;
}

// The four normal numbers of elements 4q .. 4q+3 (one Philox call, two
// Box-Muller pairs): out[i] == normal_at(4q + i, ...).
__device__ inline void normals4(uint32_t q, uint32_t draw, uint32_t stream, uint32_t k0, uint32_t k1, float* out) {
  const U4 r = philox4x32_10(U4{q, draw, stream, kNoiseTag}, k0, k1);
  box_muller(r.x, r.y, &out[0], &out[1]);
  box_muller(r.z, r.w, &out[2], &out[3]);
}

}  // namespace gpo

// Counter-based Philox4x32-10 and the fp32 normal transform used for the
// noise term of both phase-field plugins.
//
// Replaces cuRAND XORWOW (src/random.cu:10-28, 79-82 of General-Simulator).
// Philox4x32-10 follows Salmon, Moraes, Dror and Shaw, "Parallel random
// numbers: as easy as 1, 2, 3" (SC'11). Counter layout:
//   ctr = (voxel/4, field, step, purpose)   purpose 0: noise, 1: initial field
//   key = (seed & 0xffffffff, seed >> 32)
// The normal transform is Box-Muller written with + - * and integer bit
// operations only (no library log/sqrt/sin/cos/division), evaluated with
// floating-point contraction disabled, so host and device produce identical
// bits. Constants are given as IEEE-754 binary32 bit patterns:
//   ln(1+f) = f*q(f), q a degree-9 least-squares fit on [sqrt(1/2)-1, sqrt(2)-1]
//   sqrt(x) = x * rsqrt(x), rsqrt by 3 Newton steps from the 0x5F3759DF seed
//   sin/cos on [0, pi/2) by their Taylor series to degree 13/14.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gsk {

__host__ __device__ __forceinline__ float bits_f(uint32_t b) { return __builtin_bit_cast(float, b); }
__host__ __device__ __forceinline__ uint32_t f_bits(float f) { return __builtin_bit_cast(uint32_t, f); }

__host__ __device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

#pragma clang fp contract(off)
__host__ __device__ __forceinline__ float nlog(float u) {
#pragma clang fp contract(off)
  const uint32_t i = f_bits(u);
  int e = (int)((i >> 23) & 0xFFu) - 127;
  float m = bits_f((i & 0x007FFFFFu) | 0x3F800000u);
  if (m > bits_f(0x3FB504F3u)) { m = m * 0.5f; e = e + 1; }
  const float f = m - 1.0f;
  float q = bits_f(0xBD9C70BBu);
  q = q * f + bits_f(0x3E043E25u);
  q = q * f + bits_f(0xBE07962Fu);
  q = q * f + bits_f(0x3E11342Cu);
  q = q * f + bits_f(0xBE2A13CFu);
  q = q * f + bits_f(0x3E4CD23Cu);
  q = q * f + bits_f(0xBE800214u);
  q = q * f + bits_f(0x3EAAAAA8u);
  q = q * f + bits_f(0xBEFFFFFCu);
  q = q * f + bits_f(0x3F800000u);
  const float p = f * q;
  const float t = (float)e;
  return t * bits_f(0x3F317200u) + (p + t * bits_f(0x35BFBE8Eu));
}

__host__ __device__ __forceinline__ float nsqrt(float x) {
#pragma clang fp contract(off)
  if (!(x > 0.0f)) return 0.0f;
  float y = bits_f(0x5F3759DFu - (f_bits(x) >> 1));
  const float h = 0.5f * x;
#pragma unroll
  for (int it = 0; it < 3; ++it) {
    float t = y * y;
    t = h * t;
    t = 1.5f - t;
    y = y * t;
  }
  return x * y;
}

__host__ __device__ __forceinline__ void nsincos(float th, float& s, float& c) {
#pragma clang fp contract(off)
  const float z = th * th;
  float ps = bits_f(0x2F309231u);
  ps = ps * z + bits_f(0xB2D7322Bu);
  ps = ps * z + bits_f(0x3638EF1Du);
  ps = ps * z + bits_f(0xB9500D01u);
  ps = ps * z + bits_f(0x3C088889u);
  ps = ps * z + bits_f(0xBE2AAAABu);
  ps = ps * z + 1.0f;
  s = th * ps;
  float pc = bits_f(0xAD49CBA5u);
  pc = pc * z + bits_f(0x310F76C7u);
  pc = pc * z + bits_f(0xB493F27Eu);
  pc = pc * z + bits_f(0x37D00D01u);
  pc = pc * z + bits_f(0xBAB60B61u);
  pc = pc * z + bits_f(0x3D2AAAABu);
  pc = pc * z + bits_f(0xBF000000u);
  c = pc * z + 1.0f;
}

__host__ __device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, float& z0, float& z1) {
#pragma clang fp contract(off)
  const float u1 = (float)((a >> 8) + 1u) * bits_f(0x33800000u);
  const float r = nsqrt(-2.0f * nlog(u1));
  const uint32_t q = b >> 30;
  const float fr = (float)((b >> 8) & 0x3FFFFFu) * bits_f(0x34800000u);
  const float th = fr * bits_f(0x3FC90FDBu);
  float s, c;
  nsincos(th, s, c);
  float cc = (q == 0u) ? c : (q == 1u) ? -s : (q == 2u) ? -c : s;
  float ss = (q == 0u) ? s : (q == 1u) ? c : (q == 2u) ? -s : -c;
  z0 = r * cc;
  z1 = r * ss;
}
#pragma clang fp contract(on)

// Four standard normals for voxels 4g..4g+3 of `field` at `step`.
__host__ __device__ __forceinline__ void normal4(uint64_t seed, uint32_t g, uint32_t field, uint32_t step,
                                                 float out[4]) {
  uint32_t c[4] = {g, field, step, 0u};
  philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  box_muller(c[0], c[1], out[0], out[1]);
  box_muller(c[2], c[3], out[2], out[3]);
}

// The two normals of normal4 for voxels 4g + 2h, 4g + 2h + 1 (h = 0 or 1), bit-identical to
// out[2h], out[2h + 1] of normal4.
__host__ __device__ __forceinline__ void normal2(uint64_t seed, uint32_t g, uint32_t field, uint32_t step, uint32_t h,
                                                 float out[2]) {
  uint32_t c[4] = {g, field, step, 0u};
  philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  box_muller(h ? c[2] : c[0], h ? c[3] : c[1], out[0], out[1]);
}

// Four uniforms in [0,1) (24-bit) for the initial-field generator.
__host__ __device__ __forceinline__ void uniform4_bits(uint64_t seed, uint32_t g, uint32_t field, uint32_t out[4]) {
  uint32_t c[4] = {g, field, 0u, 1u};
  philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  out[0] = c[0] >> 8; out[1] = c[1] >> 8; out[2] = c[2] >> 8; out[3] = c[3] >> 8;
}

}  // namespace gsk

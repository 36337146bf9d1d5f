// Counter-based random numbers for particle loading and inflow.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies with the
// published multipliers and Weyl key increments.  Every macro-particle owns
// the counter (index lo, index hi, event, 4*species + draw), so the stream of
// a particle does not depend on launch shape, and the host and the device
// draw the same integers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pic {

struct u32x4 {
  uint32_t v[4];
};

__host__ __device__ inline u32x4 philox4x32_10(uint32_t c0, uint32_t c1,
                                               uint32_t c2, uint32_t c3,
                                               uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  u32x4 o;
  o.v[0] = c0;
  o.v[1] = c1;
  o.v[2] = c2;
  o.v[3] = c3;
  return o;
}

// 53-bit uniform double in [0,1) from two words.
__host__ __device__ inline double uniform53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) *
         (1.0 / 9007199254740992.0);
}

// The seven uniforms one macro-particle consumes: three for the position
// inside its cell, four for three Box-Muller normals.
struct ParticleDraw {
  double ux, uy, uz;  // position fractions
  double n0, n1, n2;  // standard normals
};

__host__ __device__ inline ParticleDraw draw_particle(uint64_t seed,
                                                      int species,
                                                      uint32_t event,
                                                      uint64_t idx) {
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  const uint32_t i0 = (uint32_t)idx, i1 = (uint32_t)(idx >> 32);
  const uint32_t s = 4u * (uint32_t)species;
  u32x4 a = philox4x32_10(i0, i1, event, s + 0, k0, k1);
  u32x4 b = philox4x32_10(i0, i1, event, s + 1, k0, k1);
  u32x4 c = philox4x32_10(i0, i1, event, s + 2, k0, k1);
  u32x4 d = philox4x32_10(i0, i1, event, s + 3, k0, k1);
  ParticleDraw r;
  r.ux = uniform53(a.v[0], a.v[1]);
  r.uy = uniform53(a.v[2], a.v[3]);
  r.uz = uniform53(b.v[0], b.v[1]);
  const double g0 = uniform53(b.v[2], b.v[3]);
  const double g1 = uniform53(c.v[0], c.v[1]);
  const double g2 = uniform53(c.v[2], c.v[3]);
  const double g3 = uniform53(d.v[0], d.v[1]);
  const double two_pi = 6.283185307179586;
  const double r0 = sqrt(-2.0 * log(1.0 - g0));
  const double r1 = sqrt(-2.0 * log(1.0 - g2));
  r.n0 = r0 * cos(two_pi * g1);
  r.n1 = r0 * sin(two_pi * g1);
  r.n2 = r1 * cos(two_pi * g3);
  return r;
}

}  // namespace pic

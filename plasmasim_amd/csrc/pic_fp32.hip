// FP32 chosen form of the step on gfx950.
//
// Particles: packed cell (k<<20 | l<<10 | m, guard-inclusive indices, -1 for
// a dead slot), fp32 in-cell offsets in [0,1) and fp32 velocities: 28 bytes,
// positions uniformly resolved to 2^-24 cell anywhere in the box.  Fields:
// fp32 planes; the B planes hold the induced field B - B0 (B0 is static), so
// the curl never forms B - B0 in single precision next to the dipole.
//
// The particle pass runs one workgroup per sort tile.  Particles of a tile
// are contiguous after the sort; the workgroup accumulates their current and
// density in an LDS window (tile +- 3 cells) of 64-bit fixed-point integers
// (on gfx950 integer LDS atomics are an order of magnitude faster than fp32
// ones, and their sums do not depend on the order of the additions), then
// adds the non-zero window nodes to the global planes.  A particle that
// drifted more than one cell out of its tile since the sort, or a particle of
// the inflow tail outside its chunk's window, deposits with global atomics
// directly.  Atomic mode: global fp32 atomics into the fp32 planes.
// Deterministic mode: global 64-bit integer atomics into fixed-point planes,
// converted to fp32 after the pass, so results are bit-reproducible.
//
// Sorting: the tile sort of pic_tsort.h.  Called on its own it copies the
// particles into their new slots; inside pic_step the species' pass of the
// sorting step reads them through the permutation instead and writes the
// other buffer, and nothing else touches the particles in between.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>

#include "philox.h"
#include "pic_internal.h"
#include "pic_tiles.h"
#include "pic_tsort.h"

namespace pic {
namespace f32 {

#ifndef PIC_KWIN32
#define PIC_KWIN32 1536
#endif
#ifndef PIC_MINWAVES32
#define PIC_MINWAVES32 1
#endif
#ifndef PIC_PPT32
#define PIC_PPT32 4  // particles per thread and vector width of the pass
#endif
#ifndef PIC_PASS_BLOCK
#define PIC_PASS_BLOCK 256  // threads per workgroup of the particle pass
#endif
#if PIC_MINWAVES32 > 1
#define PIC_WAVES32 __attribute__((amdgpu_waves_per_eu(PIC_MINWAVES32, 8)))
#else
#define PIC_WAVES32
#endif
constexpr int kWin = PIC_KWIN32;  // LDS window capacity in nodes (4 x int64 each)
constexpr int kPassBlock = PIC_PASS_BLOCK;
constexpr int kPackShift = 10;
constexpr int kPackMask = 1023;

__device__ __host__ inline int32_t pack(int k, int l, int m) {
  return (k << 20) | (l << 10) | m;
}
__device__ inline void unpack(int32_t c, int &k, int &l, int &m) {
  k = c >> 20;
  l = (c >> kPackShift) & kPackMask;
  m = c & kPackMask;
}

// constants of the particle pass for one species, fp32
struct Push32 {
  float T;          // q dt / (2 m), m signed
  float hdt;        // dt / (2 dx): offset change per unit velocity per half step
  float c, c2, inv_c;
  float qw;         // q w / dx^3: current per unit velocity
  float w;          // w / dx^3: density
  float eimf_x, eimf_y;  // V * IMF terms added to the gathered E (:2817-2818)
  float blo[3], bhi[3];  // ballistic zone bounds in cell units
  int periodic[2];
  float fx;         // fixed-point units per unit value (a power of two)
  double fx_inv;    // 1 / fx
  float box_lo[3], box_hi[3];  // thrust_fave's central box, cell units (open)
  int box_margin;   // cells a particle can be from its tile's shifted anchor
  uint32_t bprev;   // byte offset of the end-of-step total B from the bt planes
};

// grid of planes used by the pass
struct Grid32 {
  Geom g;
  const float *ex, *ey, *ez;
  const float *btx, *bty, *btz;  // total B = B0 + B1, maintained for the gather
};

// ---------------------------------------------------------------------------
// small helpers

struct AxisW32 {
  int bi, bh;
  float i0, i1, h0, h1;
};

__device__ inline AxisW32 axis_w(int k, float o) {
  AxisW32 a;
  a.bi = k;
  a.i1 = o;
  a.i0 = 1.0f - o;
  if (o >= 0.5f) {
    a.bh = k;
    a.h1 = o - 0.5f;
  } else {
    a.bh = k - 1;
    a.h1 = o + 0.5f;
  }
  a.h0 = 1.0f - a.h1;
  return a;
}

__device__ inline float lerp8(const float *__restrict__ F, int64_t i, int64_t dl,
                              int64_t dk, float a0, float a1, float b0, float b1,
                              float c0, float c1) {
  const float v00 = c0 * F[i] + c1 * F[i + 1];
  const float v01 = c0 * F[i + dl] + c1 * F[i + dl + 1];
  const float v10 = c0 * F[i + dk] + c1 * F[i + dk + 1];
  const float v11 = c0 * F[i + dk + dl] + c1 * F[i + dk + dl + 1];
  return a0 * (b0 * v00 + b1 * v01) + a1 * (b0 * v10 + b1 * v11);
}

// trilinear gather on the six staggered lattices (shape_func_v2
// :3017-3326); B is the total field B0 + B1
__device__ inline void gather32(const Grid32 &G, const AxisW32 &X,
                                const AxisW32 &Y, const AxisW32 &Z, float f[6]) {
  const Geom &g = G.g;
  const int64_t dl = g.SZ, dk = (int64_t)g.SY * g.SZ;
  f[0] = lerp8(G.ex, g.idx(X.bi, Y.bh, Z.bh), dl, dk, X.i0, X.i1, Y.h0, Y.h1, Z.h0, Z.h1);
  f[1] = lerp8(G.ey, g.idx(X.bh, Y.bi, Z.bh), dl, dk, X.h0, X.h1, Y.i0, Y.i1, Z.h0, Z.h1);
  f[2] = lerp8(G.ez, g.idx(X.bh, Y.bh, Z.bi), dl, dk, X.h0, X.h1, Y.h0, Y.h1, Z.i0, Z.i1);
  f[3] = lerp8(G.btx, g.idx(X.bh, Y.bi, Z.bi), dl, dk, X.h0, X.h1, Y.i0, Y.i1, Z.i0, Z.i1);
  f[4] = lerp8(G.bty, g.idx(X.bi, Y.bh, Z.bi), dl, dk, X.i0, X.i1, Y.h0, Y.h1, Z.i0, Z.i1);
  f[5] = lerp8(G.btz, g.idx(X.bi, Y.bi, Z.bh), dl, dk, X.i0, X.i1, Y.i0, Y.i1, Z.h0, Z.h1);
}

// cal_track4 (:2804-2870) in fp32: B scaled by 1/gamma(u^{n-1/2}), gamma1
// set to 0 when |v| >= c
__device__ inline void boris32(const Push32 &p, const float f[6], float &vx,
                               float &vy, float &vz) {
  const float T = p.T;
  const float ex = f[0] + p.eimf_x, ey = f[1] + p.eimf_y, ez = f[2];
  const float d = p.c2 - (vx * vx + vy * vy + vz * vz);
  // gamma1 = c / sqrt(c^2 - v^2); with u = gamma1 v, c^2 + u^2 = (gamma1 c)^2,
  // so gamma2 = c / sqrt(c^2 + u^2) = 1 / gamma1 = sqrt(d) / c (1 when the
  // reference's isnan clamp zeroes u)
  const float r = rsqrtf(d);
  const float g1 = d > 0.0f ? p.c * r : 0.0f;
  float ux = g1 * vx, uy = g1 * vy, uz = g1 * vz;
  const float g2 = d > 0.0f ? (d * p.inv_c) * r : 1.0f;
  const float bx = g2 * f[3], by = g2 * f[4], bz = g2 * f[5];
  const float S = __fdividef(2.0f * T, 1.0f + T * T * (bx * bx + by * by + bz * bz));
  ux += ex * T;
  uy += ey * T;
  uz += ez * T;
  const float ox = ux + T * (uy * bz - uz * by);
  const float oy = uy + T * (uz * bx - ux * bz);
  const float oz = uz + T * (ux * by - uy * bx);
  ux += S * (oy * bz - oz * by) + ex * T;
  uy += S * (oz * bx - ox * bz) + ey * T;
  uz += S * (ox * by - oy * bx) + ez * T;
  const float g = p.c * rsqrtf(p.c2 + ux * ux + uy * uy + uz * uz);
  vx = g * ux;
  vy = g * uy;
  vz = g * uz;
}

// half-step drift of one axis: offset += v dt/(2dx), at most one cell
// crossed (|v dt/(2dx)| < 1): the crossing is floor(offset) in {-1, 0, 1};
// o - 1 is exact, o + 1 can round up to 1 and is kept below it
__device__ inline void drift_axis(int &k, float &o, float v, float hdt) {
  o += v * hdt;
  const float f = floorf(o);
  o = fminf(o - f, 0.99999994f);
  k += (int)f;
}

// ---------------------------------------------------------------------------
// accumulation targets

// Atomics with explicit address spaces: the LDS adds must stay ds_add_* and
// the global ones global_atomic_*; through generic pointers the compiler may
// merge both paths into flat atomics.
#define PIC_LDS __attribute__((address_space(3)))
#define PIC_GLB __attribute__((address_space(1)))

typedef unsigned long long u64;

// Fixed-point value of one contribution already multiplied by the scale.
// The scale is a power of two chosen so that every scaled contribution is
// below 2^30 in magnitude (multiplying by it is exact); the conversion
// truncates to an int32 (an error below one unit, 2^-30 of the largest
// contribution) that sign-extends to the 64-bit integer.
__device__ inline u64 to_fx(float v) { return (u64)(long long)(int)v; }

// global target: fp32 planes (atomic mode) or fixed-point planes
template <bool DET>
struct GlobT;
template <>
struct GlobT<false> {
  using T = float;
};
template <>
struct GlobT<true> {
  using T = u64;
};

__device__ inline void glob_add(float *p, float v) {
  __hip_atomic_fetch_add((PIC_GLB float *)p, v, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void glob_add(u64 *p, float v) {
  __hip_atomic_fetch_add((PIC_GLB u64 *)p, to_fx(v), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}
// flush of one LDS node
__device__ inline void glob_flush(float *p, u64 v, double inv_s) {
  __hip_atomic_fetch_add((PIC_GLB float *)p, (float)((double)(long long)v * inv_s),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void glob_flush(u64 *p, u64 v, double) {
  __hip_atomic_fetch_add((PIC_GLB u64 *)p, v, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// the fused particle pass
//
// Per-particle work is written for a low VALU count, the limit of this pass:
// 32-bit byte offsets from uniform plane bases for the gather (SGPR-base
// loads), one window index per particle for the LDS deposit, and a lane-
// constant walk of each 8-node stencil.  Lanes of a wave mostly hold
// particles of the same cell (the particles are cell-sorted) and would hit
// the same LDS node in the same instruction; lane j walks its stencil
// mirrored along the axes set in j & 7, so one instruction's lanes spread
// over the eight nodes.  The mirror costs two selects per axis and quantity;
// the eight node offsets of the walk are per-lane constants.

// the lane's walk over the 8 nodes of a stencil in the LDS window
struct Walk {
  int first;        // byte offset of the first visited node from the stencil base
  int off[8];       // byte offsets of the visits from the first node
  // an axis's two weights in walk order are s x + b and -s x + c for the
  // weight x of its upper node: (1 - x, x), or (x, 1 - x) walked downwards
  float s[3], b[3], c[3];
};

__device__ inline Walk make_walk(int WY, int WZ) {
  Walk k;
  const int sw = threadIdx.x & 7;
  const bool mx = sw & 4, my = sw & 2, mz = sw & 1;  // axes walked downwards
  const int sz = 8, sy = 8 * WZ, sx = 8 * WY * WZ;
  const int dx = mx ? -sx : sx, dy = my ? -sy : sy, dz = mz ? -sz : sz;
  k.first = (mx ? sx : 0) + (my ? sy : 0) + (mz ? sz : 0);
  k.off[0] = 0;
  k.off[1] = dz;
  k.off[2] = dy;
  k.off[3] = dy + dz;
  k.off[4] = dx;
  k.off[5] = dx + dz;
  k.off[6] = dx + dy;
  k.off[7] = dx + dy + dz;
  const bool mm[3] = {mx, my, mz};
  for (int a = 0; a < 3; ++a) {
    k.s[a] = mm[a] ? 1.0f : -1.0f;
    k.b[a] = mm[a] ? 0.0f : 1.0f;
    k.c[a] = mm[a] ? 1.0f : 0.0f;
  }
  return k;
}

// the weights (1 - x, x) of an axis in the lane's walk order; each is one
// rounding of the same value as the plain 1 - x and x
__device__ __forceinline__ void walk_pair(const Walk &k, int a, float x, float &w0,
                                          float &w1) {
  w0 = __builtin_fmaf(k.s[a], x, k.b[a]);
  w1 = __builtin_fmaf(-k.s[a], x, k.c[a]);
}

// a contribution known to be non-negative (the density: weight and shape
// factors are >= 0): its truncation is >= 0 and the high word is zero
__device__ __forceinline__ void lds_fx_add_pos(char *acc, int byte, float v) {
  __hip_atomic_fetch_add((PIC_LDS u64 *)(acc + byte), (u64)(uint32_t)(int)v,
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void lds_fx_add(char *acc, int byte, float v) {
  __hip_atomic_fetch_add((PIC_LDS u64 *)(acc + byte), to_fx(v), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
}

// the same stencil straight into the global planes (particles outside the
// window); q carries the fixed-point scale in deterministic mode only
template <typename G>
__device__ __noinline__ void glob_dep8(G *glob, const Geom &g, int a, int b, int c,
                                       float q, float a0, float a1, float b0,
                                       float b1, float c0, float c1) {
  const int64_t dl = g.SZ, dk = (int64_t)g.SY * g.SZ;
  const int64_t i = g.idx(a, b, c);
  const float qa0 = q * a0, qa1 = q * a1;
  glob_add(glob + i, qa0 * (b0 * c0));
  glob_add(glob + i + 1, qa0 * (b0 * c1));
  glob_add(glob + i + dl, qa0 * (b1 * c0));
  glob_add(glob + i + dl + 1, qa0 * (b1 * c1));
  glob_add(glob + i + dk, qa1 * (b0 * c0));
  glob_add(glob + i + dk + 1, qa1 * (b0 * c1));
  glob_add(glob + i + dk + dl, qa1 * (b1 * c0));
  glob_add(glob + i + dk + dl + 1, qa1 * (b1 * c1));
}

__device__ __forceinline__ const float *at(const float *base, uint32_t byte) {
  return (const float *)((const char *)base + byte);
}

// trilinear interpolation of one plane around the byte offset o of its
// stencil base; sl, sk: byte strides of l and k
__device__ __forceinline__ float lerp8o(const float *F, uint32_t o, uint32_t sl,
                                        uint32_t sk, float a0, float a1, float b0,
                                        float b1, float c0, float c1) {
  const float *p00 = at(F, o), *p01 = at(F, o + sl), *p10 = at(F, o + sk),
              *p11 = at(F, o + sk + sl);
  const float v00 = c0 * p00[0] + c1 * p00[1];
  const float v01 = c0 * p01[0] + c1 * p01[1];
  const float v10 = c0 * p10[0] + c1 * p10[1];
  const float v11 = c0 * p11[0] + c1 * p11[1];
  return a0 * (b0 * v00 + b1 * v01) + a1 * (b0 * v10 + b1 * v11);
}

// gather32 with 32-bit offsets: the cell's byte offset o and the half-node
// shifts (byte offset of one step down along each axis, or 0)
__device__ __forceinline__ void gather_o(const Grid32 &G, uint32_t o, uint32_t sl,
                                         uint32_t sk, uint32_t hx, uint32_t hy,
                                         uint32_t hz, const AxisW32 &X,
                                         const AxisW32 &Y, const AxisW32 &Z,
                                         float f[6]) {
  const uint32_t ox = o - hx, oy = o - hy, oz = o - hz;
  f[0] = lerp8o(G.ex, oy - hz, sl, sk, X.i0, X.i1, Y.h0, Y.h1, Z.h0, Z.h1);
  f[1] = lerp8o(G.ey, ox - hz, sl, sk, X.h0, X.h1, Y.i0, Y.i1, Z.h0, Z.h1);
  f[2] = lerp8o(G.ez, ox - hy, sl, sk, X.h0, X.h1, Y.h0, Y.h1, Z.i0, Z.i1);
  f[3] = lerp8o(G.btx, ox, sl, sk, X.h0, X.h1, Y.i0, Y.i1, Z.i0, Z.i1);
  f[4] = lerp8o(G.bty, oy, sl, sk, X.i0, X.i1, Y.h0, Y.h1, Z.i0, Z.i1);
  f[5] = lerp8o(G.btz, oz, sl, sk, X.i0, X.i1, Y.i0, Y.i1, Z.h0, Z.h1);
}

// window of one workgroup, uniform
struct WinS {
  int k0, l0, m0, WY, WZ;
  int vk0, vk1, vl0, vl1, vm0, vm1;
  bool on;
};

__device__ inline bool in_win_s(const WinS &w, int k, int l, int m) {
  return w.on && (unsigned)(k - w.vk0) <= (unsigned)(w.vk1 - w.vk0) &&
         (unsigned)(l - w.vl0) <= (unsigned)(w.vl1 - w.vl0) &&
         (unsigned)(m - w.vm0) <= (unsigned)(w.vm1 - w.vm0);
}

// J (current) or n deposit of one particle at cell (k,l,m), offsets
// (ox,oy,oz), into the workgroup's LDS window
// Deposit q * a_i b_j c_k at the 8 window nodes of a stencil whose lowest
// node is at byte offset base, visiting them in the lane's walk order; the
// weights come in that order (walk_pair) and q carries the fixed-point scale.
template <bool POS>
__device__ __forceinline__ void lds_dep8(char *acc, int base, const Walk &k, float q,
                                         float A0, float A1, float B0, float B1,
                                         float C0, float C1) {
  const float qa0 = q * A0, qa1 = q * A1;
  const float bc00 = B0 * C0, bc01 = B0 * C1, bc10 = B1 * C0, bc11 = B1 * C1;
  const int p = base + k.first;
  const auto add = [&](int byte, float v) {
    if (POS) lds_fx_add_pos(acc, byte, v); else lds_fx_add(acc, byte, v);
  };
  add(p + k.off[0], qa0 * bc00);
  add(p + k.off[1], qa0 * bc01);
  add(p + k.off[2], qa0 * bc10);
  add(p + k.off[3], qa0 * bc11);
  add(p + k.off[4], qa1 * bc00);
  add(p + k.off[5], qa1 * bc01);
  add(p + k.off[6], qa1 * bc10);
  add(p + k.off[7], qa1 * bc11);
}

__device__ __forceinline__ void deposit_lds(const WinS &w, const Walk &wk, char *acc,
                                            bool current, int k, int l, int m,
                                            float ox, float oy, float oz, float vx,
                                            float vy, float vz, const Push32 &kp) {
  const AxisW32 X = axis_w(k, ox), Y = axis_w(l, oy), Z = axis_w(m, oz);
  const int sl = 8 * w.WZ, sk = 8 * w.WY * w.WZ;
  const int c0 = (((k - w.k0) * w.WY + (l - w.l0)) * w.WZ + (m - w.m0)) * 8;
  const int hx = X.bh != k ? sk : 0, hy = Y.bh != l ? sl : 0, hz = Z.bh != m ? 8 : 0;
  float xh0, xh1, yh0, yh1, zh0, zh1;
  walk_pair(wk, 0, X.h1, xh0, xh1);
  walk_pair(wk, 1, Y.h1, yh0, yh1);
  walk_pair(wk, 2, Z.h1, zh0, zh1);
  if (current) {
    const float qs = kp.qw * kp.fx;
    float xi0, xi1, yi0, yi1, zi0, zi1;
    walk_pair(wk, 0, X.i1, xi0, xi1);
    walk_pair(wk, 1, Y.i1, yi0, yi1);
    walk_pair(wk, 2, Z.i1, zi0, zi1);
    lds_dep8<false>(acc + 0 * 8 * kWin, c0 - hy - hz, wk, qs * vx, xi0, xi1, yh0, yh1,
                    zh0, zh1);
    lds_dep8<false>(acc + 1 * 8 * kWin, c0 - hx - hz, wk, qs * vy, xh0, xh1, yi0, yi1,
                    zh0, zh1);
    lds_dep8<false>(acc + 2 * 8 * kWin, c0 - hx - hy, wk, qs * vz, xh0, xh1, yh0, yh1,
                    zi0, zi1);
  } else {
    lds_dep8<true>(acc + 3 * 8 * kWin, c0 - hx - hy - hz, wk, kp.w * kp.fx, xh0, xh1,
                   yh0, yh1, zh0, zh1);
  }
}

// lanes of a wave holding more out-of-window particles than this deposit
// their own particle each
constexpr int kCoopMax = 8;

// Deposit of the particles outside the window (pend), straight into the
// global planes.  They are few and scattered over the waves; a lane that
// issued its 8 (J: 24) atomics alone would hold the rest of its wave idle.
// Instead the wave takes them one at a time: the particle's cell, offsets
// and charges (q0..q2: J components, or q0: density) are read from its lane
// and each active lane adds one node of one component.  Must be reached by
// every active lane of the wave.  q carries the fixed-point scale in
// deterministic mode only.
template <typename A>
__device__ __forceinline__ void glob_coop(const Geom &g, bool pend, bool current,
                                          int k, int l, int m, float ox, float oy,
                                          float oz, float q0, float q1, float q2,
                                          A *jx, A *jy, A *jz, A *nd) {
  u64 todo = __builtin_amdgcn_ballot_w64(pend);
  if (todo == 0) return;
  if (__builtin_popcountll(todo) > kCoopMax) {
    if (pend) {
      const AxisW32 X = axis_w(k, ox), Y = axis_w(l, oy), Z = axis_w(m, oz);
      if (current) {
        glob_dep8(jx, g, X.bi, Y.bh, Z.bh, q0, X.i0, X.i1, Y.h0, Y.h1, Z.h0, Z.h1);
        glob_dep8(jy, g, X.bh, Y.bi, Z.bh, q1, X.h0, X.h1, Y.i0, Y.i1, Z.h0, Z.h1);
        glob_dep8(jz, g, X.bh, Y.bh, Z.bi, q2, X.h0, X.h1, Y.h0, Y.h1, Z.i0, Z.i1);
      } else {
        glob_dep8(nd, g, X.bh, Y.bh, Z.bh, q0, X.h0, X.h1, Y.h0, Y.h1, Z.h0, Z.h1);
      }
    }
    return;
  }
  const u64 ex = __builtin_amdgcn_read_exec();
  const int rank = (int)__builtin_amdgcn_mbcnt_hi(
      (uint32_t)(ex >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ex, 0u));
  const int nact = __builtin_popcountll(ex);
  const int nn = current ? 24 : 8;
  const int64_t dl = g.SZ, dk = (int64_t)g.SY * g.SZ;
  do {
    const int src = __builtin_ctzll(todo);
    todo &= todo - 1;
    const int sk = __builtin_amdgcn_readlane(k, src);
    const int sl = __builtin_amdgcn_readlane(l, src);
    const int sm = __builtin_amdgcn_readlane(m, src);
    const float sx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ox), src));
    const float sy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(oy), src));
    const float sz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(oz), src));
    const float s0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(q0), src));
    float s1 = 0.0f, s2 = 0.0f;
    if (current) {
      s1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(q1), src));
      s2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(q2), src));
    }
    const AxisW32 X = axis_w(sk, sx), Y = axis_w(sl, sy), Z = axis_w(sm, sz);
    for (int t = rank; t < nn; t += nact) {
      // node t: component t >> 3 (3: density), corner bits (x, y, z) of t & 7;
      // the component's own axis takes the integer-node weights
      const int c = current ? t >> 3 : 3;
      const bool bx = t & 4, by = t & 2, bz = t & 1;
      const float wx = c == 0 ? (bx ? X.i1 : X.i0) : (bx ? X.h1 : X.h0);
      const float wy = c == 1 ? (by ? Y.i1 : Y.i0) : (by ? Y.h1 : Y.h0);
      const float wz = c == 2 ? (bz ? Z.i1 : Z.i0) : (bz ? Z.h1 : Z.h0);
      const int kx = c == 0 ? X.bi : X.bh, ly = c == 1 ? Y.bi : Y.bh,
                mz = c == 2 ? Z.bi : Z.bh;
      const float q = c == 1 ? s1 : c == 2 ? s2 : s0;
      A *dst = c == 0 ? jx : c == 1 ? jy : c == 2 ? jz : nd;
      const int64_t i = g.idx(kx, ly, mz) + (bx ? dk : 0) + (by ? dl : 0) + (bz ? 1 : 0);
      glob_add(dst + i, (q * wx) * (wy * wz));
    }
  } while (todo);
}

// One particle of the fused pass.  PUSH_J: hot loop A (:731-812) - ballistic
// drift or gather + Boris, half drift, J deposit at x^{n+1/2}; DRIFT_N: hot
// loop B (:818-867) - half drift, density deposit at x^{n+1}; BOUNDARY:
// sr_particle (:2041-2531) on one rank.
template <bool PUSH_J, bool DRIFT_N, bool BOUNDARY, bool DET, bool THRUST, typename A>
__device__ __forceinline__ void particle_step(
    const Grid32 &G, const Push32 &kp, const WinS &w, const Walk &wk, char *acc,
    A *jx, A *jy, A *jz, A *nd, int32_t &cl, float &ox, float &oy, float &oz,
    float &vx, float &vy, float &vz, float (&th)[6], int box_class) {
  const Geom &g = G.g;
  if (cl < 0) return;
  int k, l, m;
  unpack(cl, k, l, m);
  // fixed-point scale of the global path; in deterministic mode both paths
  // form each node's contribution with the same products in the same order,
  // so the sums do not depend on which particles the windows hold
  const float fs = DET ? kp.fx : 1.0f;
  const auto inside = [&]() {
    return (unsigned)(k - 1) <= (unsigned)(g.SX - 3) &&
           (unsigned)(l - 1) <= (unsigned)(g.SY - 3) &&
           (unsigned)(m - 1) <= (unsigned)(g.SZ - 3);
  };
  if (PUSH_J) {
    const float xf = (float)k + ox, yf = (float)l + oy, zf = (float)m + oz;
    const bool in_arr = inside();
    const bool push = in_arr && !(xf < kp.blo[0] || xf > kp.bhi[0] || yf < kp.blo[1] ||
                                  yf > kp.bhi[1] || zf < kp.blo[2] || zf > kp.bhi[2]);
    // thrust_fave (:5124-5285) of the previous step: particles of the
    // central box at their end-of-step positions and velocities, E of that
    // step (unchanged since) and the total B of its end (bp planes)
    // box_class (uniform): 0 every particle of the workgroup outside the
    // box, 1 every one inside, 2 test each
    bool box = false;
    if (THRUST && box_class == 1) box = in_arr;
    if (THRUST && box_class == 2)
      box = in_arr && xf > kp.box_lo[0] && xf < kp.box_hi[0] && yf > kp.box_lo[1] &&
            yf < kp.box_hi[1] && zf > kp.box_lo[2] && zf < kp.box_hi[2];
    if (push || box) {
      const uint32_t sl = 4u * g.SZ, sk = 4u * g.SY * g.SZ;
      const uint32_t o = (((uint32_t)k * g.SY + l) * g.SZ + m) * 4u;
      const AxisW32 X = axis_w(k, ox), Y = axis_w(l, oy), Z = axis_w(m, oz);
      const uint32_t hx = X.bh != k ? sk : 0u, hy = Y.bh != l ? sl : 0u,
                     hz = Z.bh != m ? 4u : 0u;
      float f[6];
      if (push) gather_o(G, o, sl, sk, hx, hy, hz, X, Y, Z, f);
      if (THRUST && box) {
        const uint32_t pox = o - hx, poy = o - hy, poz = o - hz;
        float ex, ey, ez;
        if (push) {
          ex = f[0];
          ey = f[1];
          ez = f[2];
        } else {
          ex = lerp8o(G.ex, poy - hz, sl, sk, X.i0, X.i1, Y.h0, Y.h1, Z.h0, Z.h1);
          ey = lerp8o(G.ey, pox - hz, sl, sk, X.h0, X.h1, Y.i0, Y.i1, Z.h0, Z.h1);
          ez = lerp8o(G.ez, pox - hy, sl, sk, X.h0, X.h1, Y.h0, Y.h1, Z.i0, Z.i1);
        }
        // the kept planes follow the bt planes at a fixed byte offset
        const float bx = lerp8o(G.btx, pox + kp.bprev, sl, sk, X.h0, X.h1, Y.i0, Y.i1, Z.i0, Z.i1);
        const float by = lerp8o(G.bty, poy + kp.bprev, sl, sk, X.i0, X.i1, Y.h0, Y.h1, Z.i0, Z.i1);
        const float bz = lerp8o(G.btz, poz + kp.bprev, sl, sk, X.i0, X.i1, Y.i0, Y.i1, Z.h0, Z.h1);
        // q n dx^3 (E + v x B), split into the E and the v x B parts as
        // F[3..8] / F[9..14]; the factor q n dx^3 is applied in FP64
        th[0] += ex;
        th[1] += ey;
        th[2] += ez;
        th[3] += vy * bz - vz * by;
        th[4] += vz * bx - vx * bz;
        th[5] += vx * by - vy * bx;
      }
      if (push) boris32(kp, f, vx, vy, vz);
    }
    drift_axis(k, ox, vx, kp.hdt);
    drift_axis(l, oy, vy, kp.hdt);
    drift_axis(m, oz, vz, kp.hdt);
    const bool dep = inside();
    const bool lds_ok = in_win_s(w, k, l, m);
    if (dep && lds_ok)
      deposit_lds(w, wk, acc, true, k, l, m, ox, oy, oz, vx, vy, vz, kp);
    const float qs = kp.qw * fs;
    glob_coop(g, dep && !lds_ok, true, k, l, m, ox, oy, oz, qs * vx, qs * vy,
              qs * vz, jx, jy, jz, nd);
  }
  if (DRIFT_N) {
    drift_axis(k, ox, vx, kp.hdt);
    drift_axis(l, oy, vy, kp.hdt);
    drift_axis(m, oz, vz, kp.hdt);
    const bool dep = inside();
    const bool lds_ok = in_win_s(w, k, l, m);
    if (dep && lds_ok)
      deposit_lds(w, wk, acc, false, k, l, m, ox, oy, oz, vx, vy, vz, kp);
    glob_coop(g, dep && !lds_ok, false, k, l, m, ox, oy, oz, kp.w * fs, 0.0f, 0.0f,
              jx, jy, jz, nd);
  }
  int32_t out = pack(k, l, m);
  if (BOUNDARY) {
    // sr_particle (:2041-2531): z must stay in [2, NZ+2) cells; x and y
    // wrap periodically (shift by the box length) or drop.  Written with
    // selects: the lanes of a wave take different branches here.
    const bool inside_zs = (unsigned)(m - 2) < (unsigned)g.NZ &&
                           (unsigned)k < (unsigned)g.SX && (unsigned)l < (unsigned)g.SY;
    const int dk = k >= g.NX + 2 ? -g.NX : k < 2 ? g.NX : 0;
    const int dl = l >= g.NY + 2 ? -g.NY : l < 2 ? g.NY : 0;
    const bool dead = !inside_zs || (dk != 0 && !kp.periodic[0]) ||
                      (dl != 0 && !kp.periodic[1]);
    out = dead ? -1 : pack(k + dk, l + dl, m);
  } else if (k < 0 || l < 0 || m < 0 || k >= g.SX || l >= g.SY || m >= g.SZ) {
    out = -1;  // left the array: not representable
  }
  cl = out;
}

// thrust_fave partials of one workgroup: the six per-lane fp32 sums added
// in FP64 over the workgroup in a fixed order, stored at slot bid
__device__ inline void thrust_store(const float (&th)[6], double *out, int bid) {
  __shared__ double sh[6][kPassBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    double a = th[q];
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
    if (lane == 0) sh[q][w] = a;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double a = sh[threadIdx.x][0];
    for (int j = 1; j < kPassBlock / 64; ++j) a += sh[threadIdx.x][j];
    out[(int64_t)bid * 6 + threadIdx.x] = a;
  }
}

// PERM: the pass also applies a sort's permutation: slot i takes its
// particle from slot perm[i] of the arrays Q and stores the result in P (the
// other buffer of the species), so the sort needs no separate copy
template <bool PUSH_J, bool DRIFT_N, bool BOUNDARY, bool DET, bool THRUST, bool PERM>
__global__ __launch_bounds__(kPassBlock) PIC_WAVES32 void k_particles32(
    Grid32 G, TileGeom tg, Particles32 P, TileRun tr, Push32 kp,
    typename GlobT<DET>::T *jx, typename GlobT<DET>::T *jy,
    typename GlobT<DET>::T *jz, typename GlobT<DET>::T *nd, double *thrust_part,
    Particles32 Q, const uint32_t *__restrict__ perm) {
  using A = typename GlobT<DET>::T;
  __shared__ u64 acc_mem[4 * kWin];
  __shared__ Win sw;
  __shared__ int64_t s_range[2];
  __shared__ int s_box;
  const Geom &g = G.g;
  const int tid = threadIdx.x;
  const int bid = xcd_remap(blockIdx.x, tr.ntiles, (int)gridDim.x);

  if (tid == 0) {
    int64_t s, e;
    Win w;
    tile_setup(bid, tr, tg, kWin,
               [&](int64_t i, int &k, int &l, int &m) {
                 const int32_t cl = PERM ? Q.cell[perm[i]] : P.cell[i];
                 if (cl < 0) return false;
                 unpack(cl, k, l, m);
                 return true;
               },
               s, e, w);
    if (!(PUSH_J || DRIFT_N)) w.on = false;
    sw = w;
    s_range[0] = s;
    s_range[1] = e;
    if (THRUST && thrust_part) {
      // a tile's particles lie within box_margin cells of its shifted
      // anchor box (cells [a0, a1] per axis), so the whole tile is inside
      // the open box or outside it unless that range meets its faces;
      // tail chunks test each particle
      int cls = 2;
      if (bid < tr.ntiles) {
        const int tz = bid % tg.NTZ, ty = (bid / tg.NTZ) % tg.NTY,
                  tx = bid / (tg.NTZ * tg.NTY);
        const int T[3] = {tg.TX, tg.TY, tg.TZ}, t3[3] = {tx, ty, tz};
        bool in = true, out = false;
        for (int a = 0; a < 3; ++a) {
          const int a0 = 2 + t3[a] * T[a] + tg.shift[a] - kp.box_margin;
          const int a1 = 2 + t3[a] * T[a] + tg.shift[a] + T[a] - 1 + kp.box_margin;
          in = in && (float)a0 > kp.box_lo[a] && (float)(a1 + 1) <= kp.box_hi[a];
          out = out || (float)(a1 + 1) <= kp.box_lo[a] || (float)a0 >= kp.box_hi[a];
        }
        cls = out ? 0 : in ? 1 : 2;
      }
      s_box = cls;
    } else {
      s_box = 0;
    }
  }
  __syncthreads();
  const int64_t s = s_range[0], e = s_range[1];
  const int box_class = THRUST ? __builtin_amdgcn_readfirstlane(s_box) : 0;
  float th[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (s >= e) {
    if (THRUST && thrust_part && tid < 6) thrust_part[(int64_t)bid * 6 + tid] = 0.0;
    return;
  }
  // the window into scalar registers
  WinS w;
  w.on = __builtin_amdgcn_readfirstlane(sw.on ? 1 : 0) != 0;
  w.k0 = sw.k0;
  w.l0 = sw.l0;
  w.m0 = sw.m0;
  w.WY = __builtin_amdgcn_readfirstlane(sw.WY);
  w.WZ = __builtin_amdgcn_readfirstlane(sw.WZ);
  // this pass tests each deposit's own cell (after its drift), not the
  // cell at the start of the step: a deposit at cell c touches nodes c - 1
  // to c + 1, so every cell within the anchor +- (kMargin - 1) keeps its
  // nodes inside the window, one cell more than the start-of-step band
  w.vk0 = sw.vk0 - 1;
  w.vk1 = sw.vk1 + 1;
  w.vl0 = sw.vl0 - 1;
  w.vl1 = sw.vl1 + 1;
  w.vm0 = sw.vm0 - 1;
  w.vm1 = sw.vm1 + 1;
  const int WX = __builtin_amdgcn_readfirstlane(sw.WX);
  const int WN = w.on ? WX * w.WY * w.WZ : 0;
  for (int i = tid; i < 4 * WN; i += blockDim.x) acc_mem[(i / WN) * kWin + i % WN] = 0;
  __syncthreads();
  char *acc = (char *)acc_mem;
  const Walk wk = make_walk(w.WY, w.WZ);

  // each thread takes PIC_PPT32 consecutive particles with vector loads;
  // slots of a boundary group that belong to a neighbouring range are left
  // untouched
  typedef int vi __attribute__((ext_vector_type(PIC_PPT32)));
  typedef float vf __attribute__((ext_vector_type(PIC_PPT32)));
  constexpr int N = PIC_PPT32;
  constexpr int SH = N == 8 ? 3 : N == 4 ? 2 : N == 2 ? 1 : 0;
  // slot indices fit in int32 (the host bounds the stored count)
  const int s32 = (int)s, e32 = (int)e;
  const int q0 = s32 >> SH, q1 = (e32 + N - 1) >> SH;
  // 32-bit byte offsets from the uniform array bases (SGPR-base loads and
  // stores, no 64-bit address arithmetic per array)
  const auto vat = [](auto *base, uint32_t byte) {
    return (decltype(base))((char *)base + byte);
  };
  for (int q = q0 + tid; q < q1; q += blockDim.x) {
    const int i0 = q << SH;
    const uint32_t bo = (uint32_t)i0 * 4u;
    vi cs;
    vf xs, ys, zs, us, vs, ws;
    if (PERM) {
      // the capacity is a multiple of 256, so the group's perm entries are
      // allocated; slots outside the range read slot 0 and are not stored
      typedef unsigned vu __attribute__((ext_vector_type(PIC_PPT32)));
      const vu pi = *(const vu *)vat(perm, bo);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const int i = i0 + j;
        const uint32_t sb = i >= s32 && i < e32 ? pi[j] * 4u : 0u;
        cs[j] = *vat(Q.cell, sb);
        xs[j] = *vat(Q.ox, sb);
        ys[j] = *vat(Q.oy, sb);
        zs[j] = *vat(Q.oz, sb);
        us[j] = *vat(Q.vx, sb);
        vs[j] = *vat(Q.vy, sb);
        ws[j] = *vat(Q.vz, sb);
      }
    } else {
      cs = *(const vi *)vat(P.cell, bo);
      xs = *(const vf *)vat(P.ox, bo);
      ys = *(const vf *)vat(P.oy, bo);
      zs = *(const vf *)vat(P.oz, bo);
      us = *(const vf *)vat(P.vx, bo);
      vs = *(const vf *)vat(P.vy, bo);
      ws = *(const vf *)vat(P.vz, bo);
    }
    const bool full = i0 >= s32 && i0 + N - 1 < e32;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const int i = i0 + j;
      if (i < s32 || i >= e32) continue;
      int c = cs[j];
      float x = xs[j], y = ys[j], z = zs[j], u = us[j], v = vs[j], ww = ws[j];
      particle_step<PUSH_J, DRIFT_N, BOUNDARY, DET, THRUST>(G, kp, w, wk, acc, jx, jy,
                                                            jz, nd, c, x, y, z, u, v,
                                                            ww, th, box_class);
      cs[j] = c;
      xs[j] = x;
      ys[j] = y;
      zs[j] = z;
      us[j] = u;
      vs[j] = v;
      ws[j] = ww;
    }
    if (full) {
      *(vi *)vat(P.cell, bo) = cs;
      *(vf *)vat(P.ox, bo) = xs;
      *(vf *)vat(P.oy, bo) = ys;
      *(vf *)vat(P.oz, bo) = zs;
      *(vf *)vat(P.vx, bo) = us;
      *(vf *)vat(P.vy, bo) = vs;
      *(vf *)vat(P.vz, bo) = ws;
    } else {
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const int i = i0 + j;
        if (i < s32 || i >= e32) continue;
        P.cell[i] = cs[j];
        P.ox[i] = xs[j];
        P.oy[i] = ys[j];
        P.oz[i] = zs[j];
        P.vx[i] = us[j];
        P.vy[i] = vs[j];
        P.vz[i] = ws[j];
      }
    }
  }
  if (THRUST && thrust_part) thrust_store(th, thrust_part, bid);
  if (!w.on) return;
  __syncthreads();
  // flush the window: non-zero nodes inside the array go to the planes
  const int WYZ = w.WY * w.WZ;
  const double inv_s = kp.fx_inv;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    A *dst = c == 0 ? jx : c == 1 ? jy : c == 2 ? jz : nd;
    for (int j = tid; j < WN; j += blockDim.x) {
      const u64 v = acc_mem[c * kWin + j];
      if (v == 0) continue;
      const int kk = w.k0 + j / WYZ, ll = w.l0 + (j / w.WZ) % w.WY,
                mm = w.m0 + j % w.WZ;
      if (kk < 0 || ll < 0 || mm < 0 || kk >= g.SX || ll >= g.SY || mm >= g.SZ)
        continue;
      glob_flush(dst + g.idx(kk, ll, mm), v, inv_s);
    }
  }
}

// ---------------------------------------------------------------------------
// field kernels (fp32; B planes hold B1 = B - B0)

__global__ void k_zero32(float *__restrict__ p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0.0f;
}

__global__ void k_zero64(unsigned long long *__restrict__ p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0ull;
}

struct Sigma32 {
  const float *d4[3], *d12[3];
  float amp;
};

__device__ inline float sigma32(const Sigma32 &s, int k2, int l2, int m2) {
  const float a4 = s.d4[0][k2] + s.d4[1][l2] + s.d4[2][m2];
  if (a4 == 0.0f) return 0.0f;
  const float a12 = s.d12[0][k2] + s.d12[1][l2] + s.d12[2][m2];
  return s.amp * (0.1f * sqrtf(a4) + 0.01f * sqrtf(a12));
}

// renew_b4_2 (:2580-2611) on B1: B1 <- f B1 - h (27 D1 - D3)/24/dx; the
// total field B0 + B1 used by the gather is written alongside
__global__ __launch_bounds__(256) void k_renew_b32(Geom g, Fields<float> F,
                                                   float *bt, Sigma32 sg,
                                                   float dt, float dx,
                                                   float eps0) {
  const int m = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int l = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 2 + blockIdx.z;
  if (m >= g.NZ + 2 || l >= g.NY + 2) return;
  const float temp = 0.25f * dt / eps0, dth = 0.5f * dt;
  const float c24 = 1.0f / (24.0f * dx);
  const int64_t c = g.idx(k, l, m);
  const int64_t dl = g.SZ, dk = (int64_t)g.SY * g.SZ;
  const float *ex = F.f[PIC_EX], *ey = F.f[PIC_EY], *ez = F.f[PIC_EZ];
  float s, f, h;
  s = sigma32(sg, 2 * k + 1, 2 * l, 2 * m);
  f = s == 0.0f ? 1.0f : (1.f - s * temp) / (1.f + s * temp);
  h = s == 0.0f ? dth : dth / (1.f + s * temp);
  const int64_t pl = g.plane;
  float b1 = f * F.f[PIC_BX][c] -
             h * c24 * (27.f * (ez[c] - ez[c - dl] - ey[c] + ey[c - 1]) -
                        (ez[c + dl] - ez[c - 2 * dl] - ey[c + 1] + ey[c - 2]));
  F.f[PIC_BX][c] = b1;
  bt[c] = F.f[PIC_B0X][c] + b1;
  s = sigma32(sg, 2 * k, 2 * l + 1, 2 * m);
  f = s == 0.0f ? 1.0f : (1.f - s * temp) / (1.f + s * temp);
  h = s == 0.0f ? dth : dth / (1.f + s * temp);
  b1 = f * F.f[PIC_BY][c] -
       h * c24 * (27.f * (ex[c] - ex[c - 1] - ez[c] + ez[c - dk]) -
                  (ex[c + 1] - ex[c - 2] - ez[c + dk] + ez[c - 2 * dk]));
  F.f[PIC_BY][c] = b1;
  bt[pl + c] = F.f[PIC_B0Y][c] + b1;
  s = sigma32(sg, 2 * k, 2 * l, 2 * m + 1);
  f = s == 0.0f ? 1.0f : (1.f - s * temp) / (1.f + s * temp);
  h = s == 0.0f ? dth : dth / (1.f + s * temp);
  b1 = f * F.f[PIC_BZ][c] -
       h * c24 * (27.f * (ey[c] - ey[c - dk] - ex[c] + ex[c - dl]) -
                  (ey[c + dk] - ey[c - 2 * dk] - ex[c + dl] + ex[c - 2 * dl]));
  F.f[PIC_BZ][c] = b1;
  bt[2 * pl + c] = F.f[PIC_B0Z][c] + b1;
}

// renew_e4_2 (:2613-2652) with the curl of B1
__global__ __launch_bounds__(256) void k_renew_e32(Geom g, Fields<float> F,
                                                   Sigma32 sg, float dt,
                                                   float dx, float c_light,
                                                   float mu0) {
  const int m = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int l = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 2 + blockIdx.z;
  if (m >= g.NZ + 2 || l >= g.NY + 2) return;
  const float temp = c_light * c_light * dt;
  const float c24 = 1.0f / (24.0f * dx);
  const int64_t c = g.idx(k, l, m);
  const int64_t dl = g.SZ, dk = (int64_t)g.SY * g.SZ;
  const float *bx = F.f[PIC_BX], *by = F.f[PIC_BY], *bz = F.f[PIC_BZ];
  float s, f, h;
  s = sigma32(sg, 2 * k, 2 * l + 1, 2 * m + 1);
  f = s == 0.0f ? 1.0f : (1.f - 0.5f * mu0 * s * temp) / (1.f + 0.5f * mu0 * s * temp);
  h = s == 0.0f ? temp : temp / (1.f + 0.5f * mu0 * s * temp);
  F.f[PIC_EX][c] = f * F.f[PIC_EX][c] +
                   h * (c24 * (27.f * ((bz[c + dl] - bz[c]) - (by[c + 1] - by[c])) -
                               ((bz[c + 2 * dl] - bz[c - dl]) - (by[c + 2] - by[c - 1]))) -
                        mu0 * (F.f[PIC_JIX][c] - F.f[PIC_JEX][c]));
  s = sigma32(sg, 2 * k + 1, 2 * l, 2 * m + 1);
  f = s == 0.0f ? 1.0f : (1.f - 0.5f * mu0 * s * temp) / (1.f + 0.5f * mu0 * s * temp);
  h = s == 0.0f ? temp : temp / (1.f + 0.5f * mu0 * s * temp);
  F.f[PIC_EY][c] = f * F.f[PIC_EY][c] +
                   h * (c24 * (27.f * ((bx[c + 1] - bx[c]) - (bz[c + dk] - bz[c])) -
                               ((bx[c + 2] - bx[c - 1]) - (bz[c + 2 * dk] - bz[c - dk]))) -
                        mu0 * (F.f[PIC_JIY][c] - F.f[PIC_JEY][c]));
  s = sigma32(sg, 2 * k + 1, 2 * l + 1, 2 * m);
  f = s == 0.0f ? 1.0f : (1.f - 0.5f * mu0 * s * temp) / (1.f + 0.5f * mu0 * s * temp);
  h = s == 0.0f ? temp : temp / (1.f + 0.5f * mu0 * s * temp);
  F.f[PIC_EZ][c] = f * F.f[PIC_EZ][c] +
                   h * (c24 * (27.f * ((by[c + dk] - by[c]) - (bx[c + dl] - bx[c])) -
                               ((by[c + 2 * dk] - by[c - dk]) - (bx[c + 2 * dl] - bx[c - dl]))) -
                        mu0 * (F.f[PIC_JIZ][c] - F.f[PIC_JEZ][c]));
}

// sr_b / sr_e on one rank: the reference copies the total field into the
// guards.  For B this form stores B1 = B - B0 and B0 is not periodic, so the
// guard receives B1(src) + B0(src) - B0(guard); E planes pass b0 = nullptr.
__device__ inline void copy_total(float *f, const float *b0, float *bt,
                                  int64_t dst, int64_t src) {
  if (b0) {
    f[dst] = f[src] + (b0[src] - b0[dst]);
    bt[dst] = bt[src];
  } else {
    f[dst] = f[src];
  }
}

__global__ void k_halo_x32(Geom g, float *a, float *b, float *c,
                           const float *a0, const float *b0, const float *c0,
                           float *bt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)g.SY * g.SZ;
  if (i >= n) return;
  const int64_t dk = n;
  float *p[3] = {a, b, c};
  const float *p0[3] = {a0, b0, c0};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    float *t = bt ? bt + q * g.plane : nullptr;
    copy_total(p[q], p0[q], t, i, i + g.NX * dk);
    copy_total(p[q], p0[q], t, i + dk, i + (g.NX + 1) * dk);
    copy_total(p[q], p0[q], t, i + (g.NX + 2) * dk, i + 2 * dk);
    copy_total(p[q], p0[q], t, i + (g.NX + 3) * dk, i + 3 * dk);
  }
}

__global__ void k_halo_y32(Geom g, float *a, float *b, float *c,
                           const float *a0, const float *b0, const float *c0,
                           float *bt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)g.SX * g.SZ;
  if (i >= n) return;
  const int k = (int)(i / g.SZ), m = (int)(i % g.SZ);
  const int64_t dl = g.SZ;
  const int64_t o = g.idx(k, 0, m);
  float *p[3] = {a, b, c};
  const float *p0[3] = {a0, b0, c0};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    float *t = bt ? bt + q * g.plane : nullptr;
    copy_total(p[q], p0[q], t, o, o + g.NY * dl);
    copy_total(p[q], p0[q], t, o + dl, o + (g.NY + 1) * dl);
    copy_total(p[q], p0[q], t, o + (g.NY + 2) * dl, o + 2 * dl);
    copy_total(p[q], p0[q], t, o + (g.NY + 3) * dl, o + 3 * dl);
  }
}

__global__ void k_fold_x32(Geom g, Fields<float> F) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)g.SY * g.SZ;
  if (i >= n) return;
  const int64_t dk = n;
#pragma unroll
  for (int q = PIC_JIX; q <= PIC_NE; ++q) {
    float *f = F.f[q] + i;
    f[2 * dk] += f[(g.NX + 2) * dk];
    f[3 * dk] += f[(g.NX + 3) * dk];
    f[g.NX * dk] += f[0];
    f[(g.NX + 1) * dk] += f[dk];
  }
}

__global__ void k_fold_y32(Geom g, Fields<float> F) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)g.SX * g.SZ;
  if (i >= n) return;
  const int k = (int)(i / g.SZ), m = (int)(i % g.SZ);
  const int64_t dl = g.SZ;
#pragma unroll
  for (int q = PIC_JIX; q <= PIC_NE; ++q) {
    float *f = F.f[q] + g.idx(k, 0, m);
    f[2 * dl] += f[(g.NY + 2) * dl];
    f[3 * dl] += f[(g.NY + 3) * dl];
    f[g.NY * dl] += f[0];
    f[(g.NY + 1) * dl] += f[dl];
  }
}

// total B planes from B0 + B1 over the whole array
__global__ void k_bt_refresh(const float *__restrict__ b1,
                             const float *__restrict__ b0, float *bt,
                             int64_t plane, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride)
    for (int q = 0; q < 3; ++q) bt[q * plane + i] = b0[q * plane + i] + b1[q * plane + i];
}

// deterministic mode: fixed point -> fp32 planes
__global__ void k_fx_to_f32(const unsigned long long *__restrict__ a, float *b,
                            int64_t n, double inv_scale) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = (float)((double)(long long)a[i] * inv_scale);
}

// ---------------------------------------------------------------------------
// loading, inflow, sort, compaction

__device__ inline void store_draw(Particles32 P, int64_t o, int k, int l,
                                  int m, const ParticleDraw &d, double vth,
                                  double mu0, double mu1, double mu2) {
  float ox = (float)d.ux, oy = (float)d.uy, oz = (float)d.uz;
  if (ox >= 1.0f) ox = 0.99999994f;
  if (oy >= 1.0f) oy = 0.99999994f;
  if (oz >= 1.0f) oz = 0.99999994f;
  P.cell[o] = pack(k, l, m);
  P.ox[o] = ox;
  P.oy[o] = oy;
  P.oz[o] = oz;
  P.vx[o] = (float)(vth * d.n0 + mu0);
  P.vy[o] = (float)(vth * d.n1 + mu1);
  P.vz[o] = (float)(vth * d.n2 + mu2);
}

// init_particle_flat (:412-625) with the same Philox streams as the FP64 form
__global__ void k_load32(Geom g, Particles32 P, int64_t n, int ppc, double vth,
                         double mu0, double mu1, double mu2, uint64_t seed,
                         int sp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t cell = i / ppc;
  const int m = (int)(cell % g.NZ);
  const int l = (int)((cell / g.NZ) % g.NY);
  const int k = (int)(cell / ((int64_t)g.NZ * g.NY));
  store_draw(P, i, k + 2, l + 2, m + 2, draw_particle(seed, sp, 0u, (uint64_t)i),
             vth, mu0, mu1, mu2);
}

// add_*_flat refill (:6042-6104): same batch indexing as the FP64 form
// Inflow batch (add_*_flat, :5993-6817): particle b of the reference order
// (face, k, l, two z cells, ppc) is drawn from its own Philox counter, so
// the batch can be stored in any order.  One thread per stored slot; the
// slots take each face slab in blocks of 4 x 4 x 2 cells (ppc consecutive
// particles per cell) when NX and NY are multiples of 4, so that a tail
// chunk of the next passes covers a compact box of cells.
__global__ void k_refill32(Geom g, Particles32 P, int64_t base, int64_t count,
                           int ppc, double vth, double mu0, double mu1,
                           double mu2, uint64_t seed, int sp, uint32_t event) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int64_t per_face = (int64_t)g.NX * g.NY * 2 * ppc;
  const int face = (int)(t / per_face);
  const int64_t r = t - face * per_face;
  const int i = (int)(r % ppc);
  const int64_t c = r / ppc;  // cell of the slot order
  int k, l, m;
  if (g.NX % 4 == 0 && g.NY % 4 == 0) {
    const int w = (int)(c % 32);
    const int64_t blk = c / 32;
    const int nby = g.NY / 4;
    m = w & 1;
    l = (int)(blk % nby) * 4 + ((w >> 1) & 3);
    k = (int)(blk / nby) * 4 + (w >> 3);
  } else {
    m = (int)(c % 2);
    l = (int)((c / 2) % g.NY);
    k = (int)(c / (2 * (int64_t)g.NY));
  }
  const int64_t b = face * per_face + (((int64_t)k * g.NY + l) * 2 + m) * ppc + i;
  const int mc = face == 0 ? m + 2 : m + g.NZ;
  store_draw(P, base + t, k + 2, l + 2, mc,
             draw_particle(seed, sp, event, (uint64_t)b), vth, mu0, mu1, mu2);
}

// In-tile inflow.  The slab cells the inflow empties (two z layers at each
// open face) are refilled with as many particles per cell, and both lie in
// the first and last z tiles.  Instead of appending the batch after the
// sorted slots, each slab tile's share of it goes into that tile's dead
// slots (killed or lost particles since the sort), so the next passes find
// it inside the tile's window and no dead slots accumulate; what does not
// fit is appended.  Particle j of the share of tile (tx, ty) at face f is
// cell j / ppc of the tile's 2-layer slab (m fastest, then l, then k) and is
// drawn from the counter of its reference-order index, as in k_refill32.
struct SlabRefill {
  Geom g;
  TileGeom tg;
  const uint32_t *tile_start;
  int ppc, nface;
  double vth, mu0, mu1, mu2;
  uint64_t seed;
  int sp;
  uint32_t event;
};

__device__ inline int slab_tile(const SlabRefill &r, int blk, int &f, int &tx,
                                int &ty) {
  const int per_face = r.tg.NTX * r.tg.NTY;
  f = blk / per_face;
  const int t = blk - f * per_face;
  tx = t / r.tg.NTY;
  ty = t - tx * r.tg.NTY;
  return (tx * r.tg.NTY + ty) * r.tg.NTZ + (f == 0 ? 0 : r.tg.NTZ - 1);
}

// rank of each dead slot of [s, e) among the dead slots, 256 at a time;
// calls visit(slot, rank) for dead slots, returns the count
template <typename Visit>
__device__ inline int rank_dead(const int32_t *cell, int64_t s, int64_t e, Visit visit) {
  __shared__ int wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int running = 0;
  for (int64_t i0 = s; i0 < e; i0 += kBlock) {
    const int64_t i = i0 + threadIdx.x;
    const bool dead = i < e && cell[i] < 0;
    const u64 bal = __builtin_amdgcn_ballot_w64(dead);
    if (lane == 0) wsum[wv] = __builtin_popcountll(bal);
    __syncthreads();
    int base = running, total = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      if (w < wv) base += wsum[w];
      total += wsum[w];
    }
    if (dead) visit(i, base + __builtin_popcountll(bal & ((1ull << lane) - 1ull)));
    running += total;
    __syncthreads();
  }
  return running;
}

__device__ inline void slab_particle(const SlabRefill &r, int f, int tx, int ty,
                                     int j, Particles32 P, int64_t slot) {
  const int c = j / r.ppc, i = j - c * r.ppc;
  const int m = c & 1, rest = c >> 1;
  const int l = ty * r.tg.TY + rest % r.tg.TY, k = tx * r.tg.TX + rest / r.tg.TY;
  const int64_t per_face = (int64_t)r.g.NX * r.g.NY * 2 * r.ppc;
  const int64_t b = f * per_face + (((int64_t)k * r.g.NY + l) * 2 + m) * r.ppc + i;
  store_draw(P, slot, k + 2, l + 2, f == 0 ? m + 2 : m + r.g.NZ,
             draw_particle(r.seed, r.sp, r.event, (uint64_t)b), r.vth, r.mu0, r.mu1,
             r.mu2);
}

// pass 1: per slab tile, the particles that do not fit its dead slots
__global__ __launch_bounds__(256) void k_slab_count32(SlabRefill r, const int32_t *cell,
                                                      int *over) {
  int f, tx, ty;
  const int t = slab_tile(r, blockIdx.x, f, tx, ty);
  const int share = r.tg.TX * r.tg.TY * 2 * r.ppc;
  const int dead = rank_dead(cell, r.tile_start[t], r.tile_start[t + 1],
                             [](int64_t, int) {});
  if (threadIdx.x == 0) over[blockIdx.x] = share > dead ? share - dead : 0;
}

// exclusive scan of the overflow counts (one workgroup), total in over[nb]
__global__ __launch_bounds__(256) void k_slab_scan32(int *over, int nb) {
  __shared__ int part[kBlock];
  const int per = (nb + kBlock - 1) / kBlock;
  const int a = threadIdx.x * per, b = min(a + per, nb);
  int sum = 0;
  for (int i = a; i < b; ++i) sum += over[i];
  part[threadIdx.x] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int w = 0; w < kBlock; ++w) {
      const int v = part[w];
      part[w] = run;
      run += v;
    }
    over[nb] = run;
  }
  __syncthreads();
  int run = part[threadIdx.x];
  for (int i = a; i < b; ++i) {
    const int v = over[i];
    over[i] = run;
    run += v;
  }
}

// pass 2: the share of each slab tile into its dead slots (their indices
// collected in LDS first, so that every lane draws), the rest after the
// stored slots at tail + its offset; dynamic LDS of share ints
__global__ __launch_bounds__(256) void k_slab_refill32(SlabRefill r, Particles32 P,
                                                       const int *over, int64_t tail) {
  extern __shared__ int dead_slot[];
  int f, tx, ty;
  const int t = slab_tile(r, blockIdx.x, f, tx, ty);
  const int share = r.tg.TX * r.tg.TY * 2 * r.ppc;
  const int ndead =
      rank_dead(P.cell, r.tile_start[t], r.tile_start[t + 1], [&](int64_t slot, int rank) {
        if (rank < share) dead_slot[rank] = (int)slot;  // slots fit int32
      });
  __syncthreads();
  const int placed = min(ndead, share);
  const int64_t base = tail + over[blockIdx.x];
  for (int j = threadIdx.x; j < share; j += kBlock)
    slab_particle(r, f, tx, ty, j, P, j < placed ? (int64_t)dead_slot[j] : base + (j - placed));
}

// inflow deletion (:6019-6021): z < 4dx or z > NZ dx, on the cell+offset
// representation: m + oz < 4 <=> m <= 3; m + oz > NZ <=> m > NZ or
// (m == NZ and oz > 0)
__device__ inline bool inflow_killed(const Geom &g, const Particles32 &P, int64_t i,
                                     int32_t c) {
  if (c < 0) return false;
  const int m = c & kPackMask;
  return m <= 3 || m > g.NZ || (m == g.NZ && P.oz[i] > 0.0f);
}

// four slots per thread: one 16-byte load of the cells (the arrays start on
// 16-byte boundaries), a store only where a slot dies
__global__ void k_inflow_kill32(Geom g, Particles32 P, int64_t n) {
  const int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (i0 >= n) return;
  if (i0 + 4 <= n) {
    int4 c = *(const int4 *)(P.cell + i0);
    const bool k0 = inflow_killed(g, P, i0, c.x), k1 = inflow_killed(g, P, i0 + 1, c.y),
               k2 = inflow_killed(g, P, i0 + 2, c.z), k3 = inflow_killed(g, P, i0 + 3, c.w);
    if (k0 | k1 | k2 | k3) {
      c.x = k0 ? -1 : c.x;
      c.y = k1 ? -1 : c.y;
      c.z = k2 ? -1 : c.z;
      c.w = k3 ? -1 : c.w;
      *(int4 *)(P.cell + i0) = c;
    }
    return;
  }
  for (int64_t i = i0; i < n; ++i)
    if (inflow_killed(g, P, i, P.cell[i])) P.cell[i] = -1;
}

__global__ void k_live_flags32(const int32_t *__restrict__ cell, int64_t n,
                               uint8_t *flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flags[i] = cell[i] >= 0 ? 1 : 0;
}

// dst[i] = src[idx[i]], four slots per thread: the destination and the
// index are contiguous (16-byte loads and stores; the arrays start on
// 16-byte boundaries), the sources are gathered
template <typename T, typename V>
__device__ inline void gather4(T *__restrict__ d, const T *__restrict__ s,
                               const uint4 &j) {
  V v;
  v.x = s[j.x];
  v.y = s[j.y];
  v.z = s[j.z];
  v.w = s[j.w];
  *(V *)d = v;
}


// ---------------------------------------------------------------------------
// tile sort (the passes are in pic_tsort.h): the 3D destination map.  The
// destination of a stored particle: false when it dropped out, else its
// key's tile t and cell in the tile cin (key = t * cells per tile + cin) and
// h, its tile's home slot in the sort's 64-entry table: the tile coordinates
// modulo 4, so that the 27 tiles around any tile take distinct slots.
struct KeyMap3 {
  Geom g;
  TileGeom tg;
  __host__ __device__ uint32_t cells() const { return (uint32_t)(tg.TX * tg.TY * tg.TZ); }
  __device__ bool dest(int32_t c, uint32_t &t, uint32_t &cin, int &h) const {
    if (c < 0) return false;
    int k, l, m;
    unpack(c, k, l, m);
    k -= 2;
    l -= 2;
    m -= 2;
    if (k < 0 || k >= g.NX || l < 0 || l >= g.NY || m < 0 || m >= g.NZ) return false;
    int tx, ty, tz;
    if (tg.LX >= 0) {
      tx = k >> tg.LX;
      ty = l >> tg.LY;
      tz = m >> tg.LZ;
    } else {
      tx = k / tg.TX;
      ty = l / tg.TY;
      tz = m / tg.TZ;
    }
    const int ix = k - tx * tg.TX, iy = l - ty * tg.TY, iz = m - tz * tg.TZ;
    t = ((uint32_t)tx * tg.NTY + (uint32_t)ty) * tg.NTZ + (uint32_t)tz;
    cin = ((uint32_t)ix * tg.TY + (uint32_t)iy) * tg.TZ + (uint32_t)iz;
    h = ((tx & 3) << 4) | ((ty & 3) << 2) | (tz & 3);
    return true;
  }
};

__global__ void k_gather32(Particles32 dst, Particles32 src,
                           const uint32_t *__restrict__ idx, int64_t n) {
  const int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (i0 >= n) return;
  if (i0 + 4 <= n) {
    const uint4 j = *(const uint4 *)(idx + i0);
    gather4<int32_t, int4>(dst.cell + i0, src.cell, j);
    gather4<float, float4>(dst.ox + i0, src.ox, j);
    gather4<float, float4>(dst.oy + i0, src.oy, j);
    gather4<float, float4>(dst.oz + i0, src.oz, j);
    gather4<float, float4>(dst.vx + i0, src.vx, j);
    gather4<float, float4>(dst.vy + i0, src.vy, j);
    gather4<float, float4>(dst.vz + i0, src.vz, j);
    return;
  }
  for (int64_t i = i0; i < n; ++i) {
    const uint32_t j = idx[i];
    dst.cell[i] = src.cell[j];
    dst.ox[i] = src.ox[j];
    dst.oy[i] = src.oy[j];
    dst.oz[i] = src.oz[j];
    dst.vx[i] = src.vx[j];
    dst.vy[i] = src.vy[j];
    dst.vz[i] = src.vz[j];
  }
}

__global__ __launch_bounds__(256) void k_sort_keys32(
    Geom g, TileGeom tg, const int32_t *__restrict__ cell, int64_t n,
    uint32_t *keys) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t ncell = (uint32_t)g.NX * g.NY * g.NZ;
  uint32_t key = ncell;
  const int32_t c = cell[i];
  if (c >= 0) {
    int k, l, m;
    unpack(c, k, l, m);
    k -= 2;
    l -= 2;
    m -= 2;
    if (k >= 0 && k < g.NX && l >= 0 && l < g.NY && m >= 0 && m < g.NZ)
      key = tg.key(k, l, m);
  }
  keys[i] = key;
}

__global__ __launch_bounds__(256) void k_count_live32(
    const int32_t *__restrict__ cell, int64_t n, int64_t *count) {
  unsigned int t = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    t += cell[i] >= 0;
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
  __shared__ unsigned int part[kBlock / 64];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s = 0;
    for (int w = 0; w < kBlock / 64; ++w) s += part[w];
    if (s) atomicAdd((unsigned long long *)count, s);
  }
}

// ---------------------------------------------------------------------------
// diagnostics (fp64 reductions of fp32 state)

constexpr int kNQF = 9, kNQP = 21;

template <int NQ>
__device__ inline void block_reduce_store(double (&v)[NQ], double *out,
                                          int max_slot) {
  __shared__ double sh[NQ][kBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double a = v[q];
    for (int o = 32; o > 0; o >>= 1) {
      const double b = __shfl_xor(a, o);
      a = q == max_slot ? fmax(a, b) : a + b;
    }
    if (lane == 0) sh[q][w] = a;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < NQ; ++q) {
      double a = sh[q][0];
      for (int j = 1; j < kBlock / 64; ++j)
        a = q == max_slot ? fmax(a, sh[q][j]) : a + sh[q][j];
      out[(int64_t)blockIdx.x * NQ + q] = a;
    }
  }
}

__global__ __launch_bounds__(256) void k_diag_fields32(Geom g, Fields<float> F,
                                                       double dx, double eps0,
                                                       double mu0, double q,
                                                       double *out) {
  double v[kNQF] = {0};
  const int64_t ncell = (int64_t)g.NX * g.NY * g.NZ;
  const double dx3 = dx * dx * dx;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ncell;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int m = 2 + (int)(t % g.NZ);
    const int l = 2 + (int)((t / g.NZ) % g.NY);
    const int k = 2 + (int)(t / ((int64_t)g.NZ * g.NY));
    const int64_t c = g.idx(k, l, m);
    const int64_t dl = g.SZ, dk = (int64_t)g.SY * g.SZ;
    const double ex = F.f[PIC_EX][c], ey = F.f[PIC_EY][c], ez = F.f[PIC_EZ][c];
    const double b1x = F.f[PIC_BX][c], b1y = F.f[PIC_BY][c], b1z = F.f[PIC_BZ][c];
    const double bx = b1x + F.f[PIC_B0X][c], by = b1y + F.f[PIC_B0Y][c],
                 bz = b1z + F.f[PIC_B0Z][c];
    const double e2 = 0.5 * (ex * ex + ey * ey + ez * ez) * eps0;
    const double b2 = 0.5 * (bx * bx + by * by + bz * bz) / mu0;
    v[0] += e2 * dx3;
    v[1] += b2 * dx3;
    v[2] += 0.5 * (b1x * b1x + b1y * b1y + b1z * b1z) / mu0 * dx3;
    const double rho = (double)F.f[PIC_NI][c] - (double)F.f[PIC_NE][c];
    v[3] += q * rho * dx3;
    const float *Ex = F.f[PIC_EX], *Ey = F.f[PIC_EY], *Ez = F.f[PIC_EZ];
    const double r =
        -q * rho / eps0 +
        (-(double)Ex[c + 2 * dk] + 27.0 * Ex[c + dk] - 27.0 * Ex[c] + Ex[c - dk] -
         Ey[c + 2 * dl] + 27.0 * Ey[c + dl] - 27.0 * Ey[c] + Ey[c - dl] -
         Ez[c + 2] + 27.0 * Ez[c + 1] - 27.0 * Ez[c] + Ez[c - 1]) /
            (24. * dx);
    v[4] += r * r;
    v[5] = fmax(v[5], fabs(r));
    if (k - 2 < g.NX * 0.8 && k - 2 > g.NX * 0.2 && l - 2 < g.NY * 0.8 &&
        l - 2 > g.NY * 0.2 && m - 2 < g.NZ * 0.8 && m - 2 > g.NZ * 0.2) {
      v[7] += e2;
      v[8] += b2;
    }
  }
  block_reduce_store<kNQF>(v, out, 5);
}

__global__ __launch_bounds__(256) void k_diag_particles32(
    Grid32 G, Particles32 P, int64_t n, double dx, double c, double mass,
    double w, double qs, double *out) {
  double v[kNQP] = {0};
  const Geom &g = G.g;
  const double dx3 = dx * dx * dx;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t cl = P.cell[i];
    if (cl < 0) continue;
    int k, l, m;
    unpack(cl, k, l, m);
    const double x = (k + (double)P.ox[i]) * dx, y = (l + (double)P.oy[i]) * dx,
                 z = (m + (double)P.oz[i]) * dx;
    const double vx = P.vx[i], vy = P.vy[i], vz = P.vz[i];
    const double v2 = vx * vx + vy * vy + vz * vz;
    const double gm = 1.0 / sqrt(1.0 - v2 / (c * c));
    v[0] += (gm - 1.0) * mass * c * c * w * dx3;
    v[1] += 0.5 * mass * w * dx3 * v2;
    v[2] += gm * mass * vx * w * dx3;
    v[3] += gm * mass * vy * w * dx3;
    v[4] += gm * mass * vz * w * dx3;
    if (x < g.NX * dx * 0.8 && x > g.NX * dx * 0.2 && y < g.NY * dx * 0.8 &&
        y > g.NY * dx * 0.2 && z < g.NZ * dx * 0.8 && z > g.NZ * dx * 0.2 &&
        k >= 1 && k <= g.SX - 2 && l >= 1 && l <= g.SY - 2 && m >= 1 &&
        m <= g.SZ - 2) {
      float f[6];
      gather32(G, axis_w(k, P.ox[i]), axis_w(l, P.oy[i]), axis_w(m, P.oz[i]), f);
      const double qn = qs * w * dx3;
      v[5] += qn * (f[0] + vy * f[5] - vz * f[4]);
      v[6] += qn * (f[1] + vz * f[3] - vx * f[5]);
      v[7] += qn * (f[2] + vx * f[4] - vy * f[3]);
      v[8] += qn * f[0];
      v[9] += qn * f[1];
      v[10] += qn * f[2];
      v[11] += qn * (vy * f[5] - vz * f[4]);
      v[12] += qn * (vz * f[3] - vx * f[5]);
      v[13] += qn * (vx * f[4] - vy * f[3]);
      v[20] += 0.5 * mass * w * dx3 * v2;
    }
  }
  block_reduce_store<kNQP>(v, out, -1);
}

// ---------------------------------------------------------------------------
// host side

struct State32 {
  float *sig32 = nullptr;          // 6 x stride floats
  unsigned long long *fx = nullptr;  // deterministic accumulators, 8 planes
  float *bt = nullptr;               // total B = B0 + B1, 3 planes (+ 3 kept)
  bool fx_dirty = false;
  double fx_scale = 0;
  tsort::Bufs ts;  // the tile sort's lists (mover pool of capacity entries)
  // a tile sort of pic_step whose permutation the species' next fused pass
  // applies (set only between the sort and that pass of one step)
  bool pending[2] = {false, false};
};

static State32 *st(pic_ctx *ctx) { return (State32 *)ctx->form_state; }

// where a deferred permutation waits for its pass: the ions' in idx_alt,
// the electrons' in idx (the electron sort and everything up to the passes
// leave idx_alt alone)
static uint32_t *pass_perm(pic_ctx *ctx, int sp) {
  return sp == PIC_ION ? ctx->idx_alt : ctx->idx;
}

static Sigma32 sigma_f(pic_ctx *ctx) {
  Sigma32 s;
  State32 *S = st(ctx);
  for (int a = 0; a < 3; ++a) {
    s.d4[a] = S->sig32 + (2 * a) * ctx->sig_stride;
    s.d12[a] = S->sig32 + (2 * a + 1) * ctx->sig_stride;
  }
  s.amp = (float)ctx->cfg.sigma_amp;
  return s;
}

static Grid32 grid32(pic_ctx *ctx) {
  Fields<float> F = fields<float>(ctx);
  Grid32 G;
  G.g = ctx->geom;
  G.ex = F.f[PIC_EX];
  G.ey = F.f[PIC_EY];
  G.ez = F.f[PIC_EZ];
  const int64_t pl = ctx->geom.plane;
  G.btx = st(ctx)->bt;
  G.bty = st(ctx)->bt + pl;
  G.btz = st(ctx)->bt + 2 * pl;
  return G;
}

int state_restored(pic_ctx *ctx) {
  st(ctx)->fx_dirty = false;
  return refresh_total_b(ctx);
}

int refresh_total_b(pic_ctx *ctx) {
  Fields<float> F = fields<float>(ctx);
  const Geom &g = ctx->geom;
  const int64_t n = (int64_t)g.SX * g.SY * g.SZ;
  k_bt_refresh<<<1024, kBlock, 0, ctx->stream>>>(F.f[PIC_BX], F.f[PIC_B0X],
                                                 st(ctx)->bt, g.plane, n);
  PIC_LAUNCH_CHECK(ctx);
  PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int init_state(pic_ctx *ctx) {
  State32 *S = new (std::nothrow) State32();
  if (!S) return set_error(ctx, PIC_E_NOMEM, "fp32 state");
  ctx->form_state = S;
  const int stride = ctx->sig_stride;
  float *h = (float *)malloc((size_t)6 * stride * sizeof(float));
  double *hd = (double *)malloc((size_t)6 * stride * sizeof(double));
  if (!h || !hd) {
    free(h);
    free(hd);
    return set_error(ctx, PIC_E_NOMEM, "fp32 sigma");
  }
  PIC_HIP(ctx, hipMemcpy(hd, ctx->sig_mem, (size_t)6 * stride * sizeof(double),
                         hipMemcpyDeviceToHost));
  for (int i = 0; i < 6 * stride; ++i) h[i] = (float)hd[i];
  free(hd);
  hipError_t e = hipMalloc(&S->sig32, (size_t)6 * stride * sizeof(float));
  if (e == hipSuccess)
    e = hipMemcpy(S->sig32, h, (size_t)6 * stride * sizeof(float),
                  hipMemcpyHostToDevice);
  free(h);
  if (e != hipSuccess)
    return set_error(ctx, PIC_E_HIP, "fp32 sigma: %s", hipGetErrorString(e));
  {
    // fixed-point unit: 2^-30 of the largest single contribution q w c/dx^3
    // (per-node sums of millions of contributions stay far below 2^63)
    const pic_config &c = ctx->cfg;
    const double wmax = c.weight[0] > c.weight[1] ? c.weight[0] : c.weight[1];
    const double qmax = c.q > 1.0 ? c.q : 1.0;
    const double unit = qmax * wmax * (c.c > 1.0 ? c.c : 1.0) /
                        (c.dx * c.dx * c.dx);
    // power of two with |contribution| * scale <= 2^30
    S->fx_scale = ldexp(1.0, (int)floor(30.0 - log2(unit)));
  }
  // 6 planes: the total B of the gather, then the end-of-step copy that
  // thrust_fave's fused force reads at a fixed offset
  PIC_HIP(ctx, hipMalloc(&S->bt, (size_t)6 * ctx->geom.plane * sizeof(float)));
  {
    int rc = refresh_total_b(ctx);
    if (rc) return rc;
  }
  if (ctx->cfg.deposit == PIC_DEPOSIT_DETERMINISTIC) {
    PIC_HIP(ctx, hipMalloc(&S->fx, (size_t)8 * ctx->geom.plane * sizeof(unsigned long long)));
    PIC_HIP(ctx, hipMemset(S->fx, 0, (size_t)8 * ctx->geom.plane * sizeof(unsigned long long)));
  }
  {
    const TileGeom tg = tile_geom(ctx);
    // the tile sort's lists, at set-up so that no step allocates
    if (tg.ntiles() > 1 && tg.TX * tg.TY * tg.TZ <= tsort::kTileCells) {
      const int rc = tsort::reserve(ctx, S->ts, tg.ntiles(), ctx->cap);
      if (rc) return rc;
    }
  }
  return 0;
}

void free_state(pic_ctx *ctx) {
  State32 *S = st(ctx);
  if (!S) return;
  if (S->sig32) (void)hipFree(S->sig32);
  if (S->fx) (void)hipFree(S->fx);
  if (S->bt) (void)hipFree(S->bt);
  tsort::release(S->ts);
  delete S;
  ctx->form_state = nullptr;
}

static dim3 interior_grid(const Geom &g, dim3 &block) {
  block = dim3(64, 4, 1);
  return dim3((unsigned)div_up(g.NZ, 64), (unsigned)div_up(g.NY, 4),
              (unsigned)g.NX);
}

int renew_grid(pic_ctx *ctx) {
  Fields<float> F = fields<float>(ctx);
  k_zero32<<<2048, kBlock, 0, ctx->stream>>>(F.f[PIC_JIX], 8 * ctx->geom.plane);
  State32 *S = st(ctx);
  if (S->fx) {
    k_zero64<<<2048, kBlock, 0, ctx->stream>>>(S->fx, 8 * ctx->geom.plane);
    S->fx_dirty = false;
  }
  PIC_LAUNCH_CHECK(ctx);
  return 0;
}

int finish_deposit(pic_ctx *ctx) {
  State32 *S = st(ctx);
  if (!S || !S->fx || !S->fx_dirty) return 0;
  Fields<float> F = fields<float>(ctx);
  k_fx_to_f32<<<2048, kBlock, 0, ctx->stream>>>(S->fx, F.f[PIC_JIX],
                                                8 * ctx->geom.plane,
                                                1.0 / S->fx_scale);
  PIC_LAUNCH_CHECK(ctx);
  S->fx_dirty = false;
  return 0;
}

int renew_b_half(pic_ctx *ctx) {
  const pic_config &c = ctx->cfg;
  dim3 b, gr = interior_grid(ctx->geom, b);
  k_renew_b32<<<gr, b, 0, ctx->stream>>>(ctx->geom, fields<float>(ctx),
                                         st(ctx)->bt, sigma_f(ctx), (float)c.dt,
                                         (float)c.dx, (float)c.eps0);
  PIC_LAUNCH_CHECK(ctx);
  return 0;
}

int renew_e(pic_ctx *ctx) {
  const pic_config &c = ctx->cfg;
  dim3 b, gr = interior_grid(ctx->geom, b);
  k_renew_e32<<<gr, b, 0, ctx->stream>>>(ctx->geom, fields<float>(ctx),
                                         sigma_f(ctx), (float)c.dt,
                                         (float)c.dx, (float)c.c, (float)c.mu0);
  PIC_LAUNCH_CHECK(ctx);
  return 0;
}

int halo(pic_ctx *ctx, int first) {
  const Geom &g = ctx->geom;
  Fields<float> F = fields<float>(ctx);
  float *a = F.f[first], *b = F.f[first + 1], *c = F.f[first + 2];
  const bool isb = first == PIC_BX;
  const float *a0 = isb ? F.f[PIC_B0X] : nullptr, *b0 = isb ? F.f[PIC_B0Y] : nullptr,
              *c0 = isb ? F.f[PIC_B0Z] : nullptr;
  float *bt = isb ? st(ctx)->bt : nullptr;
  k_halo_x32<<<(unsigned)div_up((int64_t)g.SY * g.SZ, kBlock), kBlock, 0,
               ctx->stream>>>(g, a, b, c, a0, b0, c0, bt);
  k_halo_y32<<<(unsigned)div_up((int64_t)g.SX * g.SZ, kBlock), kBlock, 0,
               ctx->stream>>>(g, a, b, c, a0, b0, c0, bt);
  PIC_LAUNCH_CHECK(ctx);
  return 0;
}

int fold_current(pic_ctx *ctx) {
  const Geom &g = ctx->geom;
  Fields<float> F = fields<float>(ctx);
  k_fold_x32<<<(unsigned)div_up((int64_t)g.SY * g.SZ, kBlock), kBlock, 0,
               ctx->stream>>>(g, F);
  k_fold_y32<<<(unsigned)div_up((int64_t)g.SX * g.SZ, kBlock), kBlock, 0,
               ctx->stream>>>(g, F);
  PIC_LAUNCH_CHECK(ctx);
  return 0;
}

int load_particles(pic_ctx *ctx) {
  {
    int rc = settle_sorts(ctx);
    if (rc) return rc;
  }
  const pic_config &c = ctx->cfg;
  const int64_t n = (int64_t)c.nx * c.ny * c.nz * c.ppc;
  if (n > ctx->cap)
    return set_error(ctx, PIC_E_CAPACITY, "load: %lld > capacity %lld",
                     (long long)n, (long long)ctx->cap);
  const TileGeom tg = tile_geom(ctx);
  const bool ref_order = tg.ntiles() == 1;
  for (int sp = 0; sp < 2; ++sp) {
    Particles32 P = p32(ctx, sp, ctx->cur[sp]);
    if (n > 0)
      k_load32<<<(unsigned)div_up(n, kBlock), kBlock, 0, ctx->stream>>>(
          ctx->geom, P, n, c.ppc, c.vth[sp], c.vdrift[0], c.vdrift[1],
          c.vdrift[2], c.seed, sp);
    PIC_LAUNCH_CHECK(ctx);
    ctx->n[sp] = n;
    ctx->sorted_end[sp] = 0;
    (void)ref_order;
  }
  // the loaded order is the reference order; bring it to tile order
  for (int sp = 0; sp < 2; ++sp) {
    int rc = sort(ctx, sp);
    if (rc) return rc;
  }
  return 0;
}

static Push32 push_const(const pic_ctx *ctx, int sp) {
  const pic_config &c = ctx->cfg;
  Push32 p;
  const double m = sp == PIC_ION ? c.mass[0] : -c.mass[1];
  p.T = (float)(0.5 * c.q * c.dt / m);
  p.hdt = (float)(0.5 * c.dt / c.dx);
  p.c = (float)c.c;
  p.c2 = (float)(c.c * c.c);
  p.inv_c = (float)(1.0 / c.c);
  const double dx3 = c.dx * c.dx * c.dx;
  p.qw = (float)(c.q * c.weight[sp] / dx3);
  p.w = (float)(c.weight[sp] / dx3);
  p.eimf_x = (float)(c.vdrift[2] * c.imf[1]);
  p.eimf_y = (float)(-c.vdrift[2] * c.imf[0]);
  for (int a = 0; a < 3; ++a) {
    p.blo[a] = (float)(c.ballistic_lo[a] / c.dx);
    p.bhi[a] = (float)(c.ballistic_hi[a] / c.dx);
  }
  p.periodic[0] = c.periodic[0];
  p.periodic[1] = c.periodic[1];
  const double sc = ctx->form_state ? ((State32 *)ctx->form_state)->fx_scale : 1.0;
  p.fx = (float)sc;
  p.fx_inv = 1.0 / sc;
  // thrust_fave's box (:5148-5150): 0.2 N dx < x < 0.8 N dx on each axis;
  // the margin and the kept planes' offset are set when the sum is on
  p.box_margin = 0;
  p.bprev = 0;
  const int N[3] = {c.nx, c.ny, c.nz};
  for (int a = 0; a < 3; ++a) {
    p.box_lo[a] = (float)(N[a] * 0.2);
    p.box_hi[a] = (float)(N[a] * 0.8);
  }
  return p;
}

int push(pic_ctx *ctx, int sp, int do_push_j, int do_drift_n,
         int do_boundary) {
  // a pending sort permutation is applied by the fused pass only
  const bool fused = do_push_j && do_drift_n && do_boundary;
  if (!fused) {
    int rc = settle(ctx, sp);
    if (rc) return rc;
  }
  const int64_t n = ctx->n[sp];
  if (n == 0) return 0;
  // slot indices and byte offsets of the pass are 32-bit
  if (n > ((int64_t)1 << 30) - 8)
    return set_error(ctx, PIC_E_CAPACITY,
                     "fp32 particle pass: %lld stored particles of one species "
                     "exceed 2^30 (32-bit byte offsets)", (long long)n);
  State32 *S = st(ctx);
  TileGeom tg = tile_geom(ctx);
  add_drift_shift(tg, ctx->cfg, ctx->drift_halves[sp]);
  ctx->drift_halves[sp] += (do_push_j ? 1 : 0) + (do_drift_n ? 1 : 0);
  TileRun tr;
  tr.tile_start = ctx->tile_start[sp];
  tr.sorted_end = ctx->sorted_end[sp];
  tr.ntiles = tr.sorted_end > 0 ? tg.ntiles() : 0;
  if (tr.ntiles == 0) tr.sorted_end = 0;
  tr.n = n;
  tr.ntail = (int)div_up(n - tr.sorted_end, kTailChunk);
  const unsigned nb = (unsigned)(tr.ntiles + tr.ntail);
  if (nb == 0) return settle(ctx, sp);
  // with a pending permutation the pass reads the current buffer through it
  // and writes the other one
  const uint32_t *pm = S->pending[sp] ? pass_perm(ctx, sp) : nullptr;
  Particles32 Q = p32(ctx, sp, ctx->cur[sp]);
  Particles32 P = pm ? p32(ctx, sp, 1 - ctx->cur[sp]) : Q;
  Push32 kp = push_const(ctx, sp);
  Grid32 G = grid32(ctx);
  const bool det = S->fx != nullptr;
  const int sel = (do_push_j ? 4 : 0) | (do_drift_n ? 2 : 0) | (do_boundary ? 1 : 0);
  // The fused pass always runs the instantiation that can add the previous
  // step's thrust_fave (its push and deposit arithmetic is then the same
  // whether or not the sum is on: the unit contracts floating-point
  // operations freely, and two instantiations could round differently);
  // without a partials buffer it adds nothing.
  const bool thrust = ctx->thrust_now && sel == 7;
  double *tp = nullptr;
  if (thrust) {
    if ((int64_t)nb > ctx->thrust_part_slots)
      return set_error(ctx, PIC_E_CAPACITY, "thrust: %u workgroups, %lld partial slots",
                       nb, (long long)ctx->thrust_part_slots);
    tp = ctx->thrust_part;
    // the kept end-of-step B is the upper half of the 6-plane bt block
    kp.bprev = (uint32_t)(3 * ctx->geom.plane * sizeof(float));
    // farthest a particle moved from its tile since the sort, beyond the
    // window's drift shift: |v| < c per half step, h halves before this pass
    const pic_config &c = ctx->cfg;
    const int h = ctx->drift_halves[sp] - 2;  // incremented above for this pass
    int smax = 0;
    for (int a = 0; a < 3; ++a) smax = std::max(smax, std::abs(tg.shift[a]));
    kp.box_margin = (int)ceil(h * 0.5 * c.c * c.dt / c.dx) + smax + 1;
    // a particle that wrapped across a periodic face lies within the margin
    // of the opposite face; only while that stays outside the box may whole
    // tiles be classified
    if (kp.box_margin + 2 >= 0.2 * std::min(c.nx, c.ny)) kp.box_margin = 1 << 20;
  }
  if (det) {
    unsigned long long *base = S->fx;
    const int64_t pl = ctx->geom.plane;
    const int off = sp == PIC_ION ? 0 : 3;
    unsigned long long *jx = base + (off + 0) * pl, *jy = base + (off + 1) * pl,
                       *jz = base + (off + 2) * pl,
                       *nd = base + (6 + (sp == PIC_ION ? 0 : 1)) * pl;
#define L(A, B, C) \
  k_particles32<A, B, C, true, false, false><<<nb, kPassBlock, 0, ctx->stream>>>(G, tg, P, tr, kp, jx, jy, jz, nd, tp, Q, pm)
    switch (sel == 7 ? 8 : sel) {
      case 8:
        if (pm)
          k_particles32<true, true, true, true, true, true>
              <<<nb, kPassBlock, 0, ctx->stream>>>(G, tg, P, tr, kp, jx, jy, jz, nd, tp, Q, pm);
        else
          k_particles32<true, true, true, true, true, false>
              <<<nb, kPassBlock, 0, ctx->stream>>>(G, tg, P, tr, kp, jx, jy, jz, nd, tp, Q, pm);
        break;
      case 6: L(true, true, false); break;
      case 4: L(true, false, false); break;
      case 2: L(false, true, false); break;
      case 1: L(false, false, true); break;
      case 3: L(false, true, true); break;
      case 5: L(true, false, true); break;
      default: return 0;
    }
#undef L
    if (do_push_j || do_drift_n) S->fx_dirty = true;
  } else {
    Fields<float> F = fields<float>(ctx);
    float *jx = F.f[sp == PIC_ION ? PIC_JIX : PIC_JEX];
    float *jy = F.f[sp == PIC_ION ? PIC_JIY : PIC_JEY];
    float *jz = F.f[sp == PIC_ION ? PIC_JIZ : PIC_JEZ];
    float *nd = F.f[sp == PIC_ION ? PIC_NI : PIC_NE];
#define L(A, B, C) \
  k_particles32<A, B, C, false, false, false><<<nb, kPassBlock, 0, ctx->stream>>>(G, tg, P, tr, kp, jx, jy, jz, nd, tp, Q, pm)
    switch (sel == 7 ? 8 : sel) {
      case 8:
        if (pm)
          k_particles32<true, true, true, false, true, true>
              <<<nb, kPassBlock, 0, ctx->stream>>>(G, tg, P, tr, kp, jx, jy, jz, nd, tp, Q, pm);
        else
          k_particles32<true, true, true, false, true, false>
              <<<nb, kPassBlock, 0, ctx->stream>>>(G, tg, P, tr, kp, jx, jy, jz, nd, tp, Q, pm);
        break;
      case 6: L(true, true, false); break;
      case 4: L(true, false, false); break;
      case 2: L(false, true, false); break;
      case 1: L(false, false, true); break;
      case 3: L(false, true, true); break;
      case 5: L(true, false, true); break;
      default: return 0;
    }
#undef L
  }
  PIC_LAUNCH_CHECK(ctx);
  if (pm) {
    ctx->cur[sp] = 1 - ctx->cur[sp];
    S->pending[sp] = false;
  }
  if (thrust) {
    const pic_config &c = ctx->cfg;
    const double qs = sp == PIC_ION ? c.q : -c.q;
    return thrust_add(ctx, sp, (int)nb, qs * c.weight[sp] * c.dx * c.dx * c.dx);
  }
  return 0;
}

int particle_boundary(pic_ctx *ctx, int sp) { return push(ctx, sp, 0, 0, 1); }

float *total_b(pic_ctx *ctx) { return st(ctx)->bt; }

// stable compaction of live particles (tile ranges become invalid)
static int compact(pic_ctx *ctx, int sp) {
  const int64_t n = ctx->n[sp];
  Particles32 src = p32(ctx, sp, ctx->cur[sp]);
  k_live_flags32<<<(unsigned)div_up(n, kBlock), kBlock, 0, ctx->stream>>>(
      src.cell, n, ctx->flags);
  PIC_LAUNCH_CHECK(ctx);
  int64_t kept = 0;
  int rc = select_flagged(ctx, ctx->flags, ctx->idx, n, &kept);
  if (rc) return rc;
  Particles32 dst = p32(ctx, sp, 1 - ctx->cur[sp]);
  if (kept > 0)
    k_gather32<<<(unsigned)div_up(kept, 4 * kBlock), kBlock, 0, ctx->stream>>>(
        dst, src, ctx->idx, kept);
  PIC_LAUNCH_CHECK(ctx);
  ctx->cur[sp] = 1 - ctx->cur[sp];
  ctx->n[sp] = kept;
  ctx->sorted_end[sp] = 0;
  return 0;
}

// add_*_flat: slab particles become dead slots in place, new particles are
// appended after the stored ones (the sort compacts)
int inflow(pic_ctx *ctx, int sp, uint32_t event) {
  {
    int rc = settle(ctx, sp);
    if (rc) return rc;
  }
  const pic_config &c = ctx->cfg;
  const Geom &g = ctx->geom;
  if (ctx->n[sp] > 0) {
    Particles32 P = p32(ctx, sp, ctx->cur[sp]);
    k_inflow_kill32<<<(unsigned)div_up(ctx->n[sp], 4 * kBlock), kBlock, 0,
                      ctx->stream>>>(g, P, ctx->n[sp]);
    PIC_LAUNCH_CHECK(ctx);
  }
  const int nface = c.refill_far ? 2 : 1;
  const int64_t add = (int64_t)g.NX * g.NY * 2 * c.ppc * nface;
  if (ctx->n[sp] + add > ctx->cap) {
    int rc = compact(ctx, sp);
    if (rc) return rc;
  }
  if (ctx->n[sp] + add > ctx->cap)
    return set_error(ctx, PIC_E_CAPACITY,
                     "inflow: %lld + %lld particles exceed capacity %lld",
                     (long long)ctx->n[sp], (long long)add, (long long)ctx->cap);
  const TileGeom tg = tile_geom(ctx);
  const int nslab = nface * tg.NTX * tg.NTY;
  // in-tile refill: tiles current, the slab layers inside the end z tiles,
  // slots kept by the adjoint's forward (it needs the batch contiguous).
  // With one z tile layer both faces' slabs lie in the same tiles, and two
  // workgroups would fill the same dead slots: such tilings append.
  const bool in_tile = add > 0 && !ctx->adj_mode && ctx->sorted_end[sp] > 0 &&
                       ctx->sorted_end[sp] <= ctx->n[sp] && tg.ntiles() > 1 &&
                       (nface == 1 || tg.NTZ >= 2) &&
                       tg.TZ >= 2 && g.NZ % tg.TZ == 0 && g.NX % tg.TX == 0 &&
                       g.NY % tg.TY == 0 && (int64_t)nslab + 1 <= ctx->cap &&
                       (int64_t)tg.TX * tg.TY * 2 * c.ppc * sizeof(int) <= 65536;
  if (in_tile) {
    Particles32 P = p32(ctx, sp, ctx->cur[sp]);
    SlabRefill r;
    r.g = g;
    r.tg = tg;
    r.tile_start = ctx->tile_start[sp];
    r.ppc = c.ppc;
    r.nface = nface;
    r.vth = c.vth[sp];
    r.mu0 = c.vdrift[0];
    r.mu1 = c.vdrift[1];
    r.mu2 = c.vdrift[2];
    r.seed = c.seed;
    r.sp = sp;
    r.event = event;
    int *over = (int *)ctx->keys;  // free between sorts
    k_slab_count32<<<(unsigned)nslab, kBlock, 0, ctx->stream>>>(r, P.cell, over);
    k_slab_scan32<<<1, kBlock, 0, ctx->stream>>>(over, nslab);
    const size_t lds = (size_t)tg.TX * tg.TY * 2 * c.ppc * sizeof(int);
    k_slab_refill32<<<(unsigned)nslab, kBlock, lds, ctx->stream>>>(r, P, over, ctx->n[sp]);
    PIC_LAUNCH_CHECK(ctx);
    // the refill takes its offsets from the scan on the device; the host
    // reads the appended count after launching it, so the GPU stays busy
    PIC_HIP(ctx, hipMemcpyAsync(ctx->host_counter, over + nslab, sizeof(int),
                                hipMemcpyDeviceToHost, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int64_t appended = *(int *)ctx->host_counter;
    ctx->n[sp] += appended;
    return 0;
  }
  if (add > 0) {
    Particles32 P = p32(ctx, sp, ctx->cur[sp]);
    k_refill32<<<(unsigned)div_up(add, kBlock), kBlock, 0, ctx->stream>>>(
        g, P, ctx->n[sp], add, c.ppc, c.vth[sp], c.vdrift[0], c.vdrift[1],
        c.vdrift[2], c.seed, sp, event);
    PIC_LAUNCH_CHECK(ctx);
  }
  ctx->n[sp] += add;
  return 0;
}

// apply a deferred sort permutation with the permutation gather (an error
// path of pic_step, or a call other than the fused pass)
int settle(pic_ctx *ctx, int sp) {
  State32 *S = st(ctx);
  if (!S || !S->pending[sp]) return 0;
  const int64_t n = ctx->n[sp];
  Particles32 src = p32(ctx, sp, ctx->cur[sp]);
  Particles32 dst = p32(ctx, sp, 1 - ctx->cur[sp]);
  k_gather32<<<(unsigned)div_up(n, 4 * kBlock), kBlock, 0, ctx->stream>>>(
      dst, src, pass_perm(ctx, sp), n);
  PIC_LAUNCH_CHECK(ctx);  // a failed launch leaves the permutation waiting
  ctx->cur[sp] = 1 - ctx->cur[sp];
  S->pending[sp] = false;
  return 0;
}

bool pending(const pic_ctx *ctx, int sp) {
  const State32 *S = (const State32 *)ctx->form_state;
  return S && S->pending[sp];
}

int settle_sorts(pic_ctx *ctx) {
  int rc = settle(ctx, PIC_ION);
  const int re = settle(ctx, PIC_ELECTRON);
  return rc ? rc : re;
}

// defer: a tile sort leaves its permutation for the species' next fused
// pass instead of copying the particles (pic_step, FP32 form)
static int sort_impl(pic_ctx *ctx, int sp, bool defer) {
  const Geom &g = ctx->geom;
  State32 *S = st(ctx);
  {
    int rc = settle(ctx, sp);
    if (rc) return rc;
  }
  ctx->drift_halves[sp] = 0;
  const TileGeom tg = tile_geom(ctx);
  const int nt = tg.ntiles();
  const int64_t n = ctx->n[sp];
  if (n == 0) {
    ctx->sorted_end[sp] = 0;
    return 0;
  }
  Particles32 src = p32(ctx, sp, ctx->cur[sp]);
  if (nt > 1 && tg.TX * tg.TY * tg.TZ <= tsort::kTileCells && ctx->sorted_end[sp] > 0 &&
      ctx->sorted_end[sp] <= n) {
    bool done = false;
    int64_t valid = 0;
    const KeyMap3 km{g, tg};
    uint32_t *perm = defer ? pass_perm(ctx, sp) : ctx->idx;
    int rc = tsort::reserve(ctx, S->ts, nt, ctx->cap);
    if (!rc)
      rc = tsort::run(ctx, km, nt, ctx->tile_start[sp], ctx->sorted_end[sp], n, src.cell,
                      S->ts, perm, &done, &valid);
    if (rc) return rc;
    if (done && defer && valid > 0) {
      ctx->n[sp] = valid;
      ctx->sorted_end[sp] = valid;
      ctx->sort_tile_count++;
      S->pending[sp] = true;
      return 0;
    }
    if (done) {
      Particles32 dst = p32(ctx, sp, 1 - ctx->cur[sp]);
      if (valid > 0)
        k_gather32<<<(unsigned)div_up(valid, 4 * kBlock), kBlock, 0, ctx->stream>>>(
            dst, src, perm, valid);
      PIC_LAUNCH_CHECK(ctx);
      ctx->cur[sp] = 1 - ctx->cur[sp];
      ctx->n[sp] = valid;
      ctx->sorted_end[sp] = valid;
      ctx->sort_tile_count++;
      return 0;
    }
  }
  ctx->sort_radix_count++;
  k_sort_keys32<<<(unsigned)div_up(n, kBlock), kBlock, 0, ctx->stream>>>(
      g, tg, src.cell, n, ctx->keys);
  PIC_LAUNCH_CHECK(ctx);
  const uint32_t ncell = (uint32_t)g.NX * g.NY * g.NZ;
  int bits = 1;
  while (bits < 32 && ((uint64_t)1 << bits) <= ncell) ++bits;
  int rc = sort_keys(ctx, ctx->keys, ctx->keys_alt, ctx->idx, n, bits);
  if (rc) return rc;
  rc = tile_bounds(ctx, ctx->keys_alt, n, (uint32_t)(tg.TX * tg.TY * tg.TZ), nt,
                   ctx->tile_start[sp]);
  if (rc) return rc;
  // the kept count is the start of the drop tile
  PIC_HIP(ctx, hipMemcpyAsync(ctx->host_counter, ctx->tile_start[sp] + nt,
                              sizeof(uint32_t), hipMemcpyDeviceToHost,
                              ctx->stream));
  PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const int64_t valid = (int64_t)(*(uint32_t *)ctx->host_counter);
  Particles32 dst = p32(ctx, sp, 1 - ctx->cur[sp]);
  if (valid > 0)
    k_gather32<<<(unsigned)div_up(valid, 4 * kBlock), kBlock, 0, ctx->stream>>>(
        dst, src, ctx->idx, valid);
  PIC_LAUNCH_CHECK(ctx);
  ctx->cur[sp] = 1 - ctx->cur[sp];
  ctx->n[sp] = valid;
  ctx->sorted_end[sp] = valid;
  return 0;
}

int sort(pic_ctx *ctx, int sp) { return sort_impl(ctx, sp, false); }

int sort_for_pass(pic_ctx *ctx, int sp) { return sort_impl(ctx, sp, true); }

int count_live(pic_ctx *ctx, int sp, int64_t *out) {
  {
    int rc = settle(ctx, sp);
    if (rc) return rc;
  }
  const int64_t n = ctx->n[sp];
  *out = 0;
  if (n == 0) return 0;
  Particles32 P = p32(ctx, sp, ctx->cur[sp]);
  PIC_HIP(ctx, hipMemsetAsync(ctx->dev_counter, 0, sizeof(int64_t), ctx->stream));
  k_count_live32<<<(unsigned)std::min<int64_t>(2048, div_up(n, kBlock)), kBlock,
                   0, ctx->stream>>>(P.cell, n, ctx->dev_counter);
  PIC_LAUNCH_CHECK(ctx);
  PIC_HIP(ctx, hipMemcpyAsync(ctx->host_counter, ctx->dev_counter,
                              sizeof(int64_t), hipMemcpyDeviceToHost,
                              ctx->stream));
  PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *out = ctx->host_counter[0];
  return 0;
}

static_assert(kNQF == kDiagNQF && kNQP == kDiagNQP, "diagnostics layout");

int diagnose_partials(pic_ctx *ctx, double *pf, unsigned long long *counts) {
  {
    int rc = settle_sorts(ctx);
    if (rc) return rc;
  }
  const pic_config &c = ctx->cfg;
  const Geom &g = ctx->geom;
  double *pp[2] = {pf + (int64_t)kDiagNBF * kNQF,
                   pf + (int64_t)kDiagNBF * kNQF + (int64_t)kDiagNBP * kNQP};
  k_diag_fields32<<<kDiagNBF, kBlock, 0, ctx->stream>>>(g, fields<float>(ctx), c.dx,
                                                        c.eps0, c.mu0, c.q, pf);
  PIC_LAUNCH_CHECK(ctx);
  Grid32 G = grid32(ctx);
  for (int sp = 0; sp < 2; ++sp) {
    Particles32 P = p32(ctx, sp, ctx->cur[sp]);
    k_diag_particles32<<<kDiagNBP, kBlock, 0, ctx->stream>>>(
        G, P, ctx->n[sp], c.dx, c.c, c.mass[sp], c.weight[sp],
        sp == PIC_ION ? c.q : -c.q, pp[sp]);
    PIC_LAUNCH_CHECK(ctx);
    if (counts && ctx->n[sp] > 0) {
      k_count_live32<<<(unsigned)std::min<int64_t>(2048, div_up(ctx->n[sp], kBlock)),
                       kBlock, 0, ctx->stream>>>(P.cell, ctx->n[sp],
                                                 (int64_t *)(counts + sp));
      PIC_LAUNCH_CHECK(ctx);
    }
  }
  return 0;
}

int diagnose(pic_ctx *ctx, pic_diag *out) {
  const int nbf = kDiagNBF, nbp = kDiagNBP;
  const int64_t need = kDiagPartials;
  if (ctx->diag_partial_cap < need) {
    if (ctx->diag_partial) (void)hipFree(ctx->diag_partial);
    ctx->diag_partial = nullptr;
    PIC_HIP(ctx, hipMalloc(&ctx->diag_partial, need * sizeof(double)));
    ctx->diag_partial_cap = need;
  }
  double *pf = ctx->diag_partial;
  double *pp[2] = {pf + (int64_t)nbf * kNQF,
                   pf + (int64_t)nbf * kNQF + (int64_t)nbp * kNQP};
  int rc0 = diagnose_partials(ctx, pf, nullptr);
  if (rc0) return rc0;
  double *h = (double *)malloc(need * sizeof(double));
  if (!h) return set_error(ctx, PIC_E_NOMEM, "diagnose: host buffer");
  hipError_t e = hipMemcpyAsync(h, pf, need * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    free(h);
    return set_error(ctx, PIC_E_HIP, "diagnose: %s", hipGetErrorString(e));
  }
  memset(out, 0, sizeof(*out));
  double s[kNQF] = {0};
  for (int b = 0; b < nbf; ++b)
    for (int q = 0; q < kNQF; ++q) {
      const double v = h[(int64_t)b * kNQF + q];
      s[q] = q == 5 ? fmax(s[q], v) : s[q] + v;
    }
  out->field_energy_e = s[0];
  out->field_energy_b = s[1];
  out->field_energy_b1 = s[2];
  out->charge = s[3];
  out->gauss_l2 = sqrt(s[4]);
  out->gauss_max = s[5];
  out->energy_box[0] = s[7];
  out->energy_box[1] = s[8];
  for (int sp = 0; sp < 2; ++sp) {
    double t[kNQP] = {0};
    const double *hp = h + (pp[sp] - pf);
    for (int b = 0; b < nbp; ++b)
      for (int q = 0; q < kNQP; ++q) t[q] += hp[(int64_t)b * kNQP + q];
    out->kinetic[sp] = t[0];
    out->kinetic_nr[sp] = t[1];
    for (int j = 0; j < 3; ++j) out->momentum[sp][j] = t[2 + j];
    for (int j = 0; j < 3; ++j) out->force[j] += t[5 + j];
    const int b0 = sp == PIC_ION ? 3 : 9;
    for (int j = 0; j < 6; ++j) out->force[b0 + j] = t[8 + j];
    out->energy_box[2 + sp] = t[20];
  }
  free(h);
  for (int sp = 0; sp < 2; ++sp) {
    int rc = count_live(ctx, sp, &out->count[sp]);
    if (rc) return rc;
  }
  return 0;
}

int upload_particles(pic_ctx *ctx, int sp, int64_t n, const double *const *h) {
  {
    int rc = settle(ctx, sp);
    if (rc) return rc;
  }
  const double dx = ctx->cfg.dx;
  const Geom &g = ctx->geom;
  const int64_t nn = n > 0 ? n : 1;
  int32_t *cell = (int32_t *)malloc(nn * sizeof(int32_t));
  float *f = (float *)malloc(nn * 6 * sizeof(float));
  if (!cell || !f) {
    free(cell);
    free(f);
    return set_error(ctx, PIC_E_NOMEM, "upload: host buffer");
  }
  const int S[3] = {g.SX, g.SY, g.SZ};
  for (int64_t i = 0; i < n; ++i) {
    int kc[3];
    float o[3];
    bool ok = true;
    for (int a = 0; a < 3; ++a) {
      const double u = h[a][i] / dx;
      if (!(u >= 0.0 && u < S[a])) {
        ok = false;
        break;
      }
      kc[a] = (int)u;
      o[a] = (float)(u - kc[a]);
      if (o[a] >= 1.0f) o[a] = 0.99999994f;
    }
    cell[i] = ok ? pack(kc[0], kc[1], kc[2]) : -1;
    for (int a = 0; a < 3; ++a) f[a * nn + i] = ok ? o[a] : 0.0f;
    for (int a = 0; a < 3; ++a) f[(3 + a) * nn + i] = (float)h[3 + a][i];
  }
  Particles32 P = p32(ctx, sp, ctx->cur[sp]);
  float *d[6] = {P.ox, P.oy, P.oz, P.vx, P.vy, P.vz};
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess && n > 0)
    e = hipMemcpy(P.cell, cell, n * sizeof(int32_t), hipMemcpyHostToDevice);
  for (int a = 0; a < 6 && e == hipSuccess && n > 0; ++a)
    e = hipMemcpy(d[a], f + a * nn, n * sizeof(float), hipMemcpyHostToDevice);
  free(cell);
  free(f);
  if (e != hipSuccess)
    return set_error(ctx, PIC_E_HIP, "upload: %s", hipGetErrorString(e));
  ctx->n[sp] = n;
  ctx->sorted_end[sp] = 0;
  return 0;
}

static int download_raw(pic_ctx *ctx, int sp, int32_t **cell, float **f,
                        int64_t *ns) {
  *ns = ctx->n[sp];
  const int64_t nn = *ns > 0 ? *ns : 1;
  *cell = (int32_t *)malloc(nn * sizeof(int32_t));
  *f = (float *)malloc(nn * 6 * sizeof(float));
  if (!*cell || !*f) {
    free(*cell);
    free(*f);
    return set_error(ctx, PIC_E_NOMEM, "download: host buffer");
  }
  Particles32 P = p32(ctx, sp, ctx->cur[sp]);
  const float *d[6] = {P.ox, P.oy, P.oz, P.vx, P.vy, P.vz};
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess && *ns > 0)
    e = hipMemcpy(*cell, P.cell, *ns * sizeof(int32_t), hipMemcpyDeviceToHost);
  for (int a = 0; a < 6 && e == hipSuccess && *ns > 0; ++a)
    e = hipMemcpy(*f + a * nn, d[a], *ns * sizeof(float), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    free(*cell);
    free(*f);
    return set_error(ctx, PIC_E_HIP, "download: %s", hipGetErrorString(e));
  }
  return 0;
}

int download_particles(pic_ctx *ctx, int sp, int64_t cap, int64_t *n,
                       double *const *h) {
  {
    int rc = settle(ctx, sp);
    if (rc) return rc;
  }
  int32_t *cell;
  float *f;
  int64_t ns;
  int rc = download_raw(ctx, sp, &cell, &f, &ns);
  if (rc) return rc;
  const int64_t nn = ns > 0 ? ns : 1;
  int64_t live = 0;
  for (int64_t i = 0; i < ns; ++i) live += cell[i] >= 0;
  *n = live;
  if (cap < live) {
    free(cell);
    free(f);
    return set_error(ctx, PIC_E_CAPACITY, "download: %lld live > cap %lld",
                     (long long)live, (long long)cap);
  }
  const double dx = ctx->cfg.dx;
  int64_t w = 0;
  for (int64_t i = 0; i < ns; ++i) {
    const int32_t c = cell[i];
    if (c < 0) continue;
    const int k = c >> 20, l = (c >> 10) & kPackMask, m = c & kPackMask;
    h[0][w] = (k + (double)f[0 * nn + i]) * dx;
    h[1][w] = (l + (double)f[1 * nn + i]) * dx;
    h[2][w] = (m + (double)f[2 * nn + i]) * dx;
    for (int a = 0; a < 3; ++a) h[3 + a][w] = f[(3 + a) * nn + i];
    w++;
  }
  free(cell);
  free(f);
  return 0;
}

int download_raw32(pic_ctx *ctx, int sp, int64_t cap, int64_t *n,
                   int32_t *cell_out, float *const *fo) {
  {
    int rc = settle(ctx, sp);
    if (rc) return rc;
  }
  int32_t *cell;
  float *f;
  int64_t ns;
  int rc = download_raw(ctx, sp, &cell, &f, &ns);
  if (rc) return rc;
  const int64_t nn = ns > 0 ? ns : 1;
  int64_t live = 0;
  for (int64_t i = 0; i < ns; ++i) live += cell[i] >= 0;
  *n = live;
  if (cap < live) {
    free(cell);
    free(f);
    return set_error(ctx, PIC_E_CAPACITY, "download: %lld live > cap %lld",
                     (long long)live, (long long)cap);
  }
  int64_t w = 0;
  for (int64_t i = 0; i < ns; ++i) {
    if (cell[i] < 0) continue;
    cell_out[w] = cell[i];
    for (int a = 0; a < 6; ++a) fo[a][w] = f[a * nn + i];
    w++;
  }
  free(cell);
  free(f);
  return 0;
}

}  // namespace f32
}  // namespace pic

extern "C" int pic_download_particles_raw32(pic_ctx *ctx, int species,
                                            int64_t cap, int64_t *n,
                                            int32_t *cell, float *ox, float *oy,
                                            float *oz, float *vx, float *vy,
                                            float *vz) {
  if (!ctx || !n || species < 0 || species > 1) return PIC_E_ARG;
  if (ctx->cfg.precision != PIC_FP32)
    return pic::set_error(ctx, PIC_E_STATE, "raw32 download needs the FP32 form");
  float *const f[6] = {ox, oy, oz, vx, vy, vz};
  return pic::f32::download_raw32(ctx, species, cap, n, cell, f);
}

// Particle routines of the 2D3V step: loading, the fused particle pass
// (bilinear gather, non-relativistic Boris push, J and n deposit, particle
// boundary), inflow, the per-tile sort, counting and host transfers, for
// both numeric forms.
//
// The particle pass runs one workgroup per sort tile (particles of a tile
// are contiguous after the sort) plus one per chunk of the unsorted tail.
// A workgroup sums current and density in an LDS window around its tile;
// deposits whose four nodes fall outside the window go to global memory
// directly.  The FP64 form sums doubles (ds_add_f64); the FP32 form sums
// 64-bit fixed-point integers (on gfx950 integer LDS atomics are an order
// of magnitude faster than fp32 ones) and flushes fp32 values.  Inside
// pic2d_step the FP32 tile sort copies no particles: the species' pass of
// the sorting step reads them through the permutation and writes the other
// buffer.
#include <hip/hip_runtime.h>
#include <new>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>

#include "philox.h"
#include "pic2d_internal.h"
#include "pic_tsort.h"
#include "pic_tiles.h"

namespace pic2d {

using pic::kMargin;
using pic::kTailChunk;
using pic::xcd_remap;

constexpr int kWin2 = 1024;  // LDS window capacity in nodes per quantity

#define PIC_LDS __attribute__((address_space(3)))
#define PIC_GLB __attribute__((address_space(1)))
typedef unsigned long long u64;

struct Run2 {
  const uint32_t *tile_start;  // ntiles + 1
  int ntiles;
  int64_t sorted_end;
  int64_t n;
};

struct Win2 {
  int k0, m0, WX, WZ;
  bool on;
  __device__ inline bool holds(int kb, int mb) const {
    return on && kb >= k0 && kb + 1 < k0 + WX && mb >= m0 && mb + 1 < m0 + WZ;
  }
};

// range and LDS window of workgroup bid (one thread); cell_of(i, k, m)
// returns false for a dead slot
template <typename CellOf>
__device__ inline void setup2(int bid, const Run2 &tr, const Tile2 &tg,
                              CellOf cell_of, int64_t &s, int64_t &e, Win2 &w) {
  int ak0, ak1, am0, am1;
  if (bid < tr.ntiles) {
    s = tr.tile_start[bid];
    e = tr.tile_start[bid + 1];
    const int tx = bid / tg.NTZ, tz = bid % tg.NTZ;
    ak0 = 2 + tx * tg.TX;
    ak1 = ak0 + tg.TX - 1;
    am0 = 2 + tz * tg.TZ;
    am1 = am0 + tg.TZ - 1;
  } else {
    const int c = bid - tr.ntiles;
    s = tr.sorted_end + (int64_t)c * kTailChunk;
    e = s + kTailChunk < tr.n ? s + kTailChunk : tr.n;
    ak0 = am0 = 1 << 30;
    ak1 = am1 = -1;
    for (int j = 0; j < 8; ++j) {
      const int64_t i = s + (e - s - 1) * j / 7;
      if (i < s || i >= e) continue;
      int k, m;
      if (!cell_of(i, k, m)) continue;
      ak0 = min(ak0, k);
      ak1 = max(ak1, k);
      am0 = min(am0, m);
      am1 = max(am1, m);
    }
  }
  w.on = ak1 >= ak0 && am1 >= am0;
  if (w.on) {
    w.k0 = ak0 - kMargin;
    w.m0 = am0 - kMargin;
    w.WX = ak1 - ak0 + 1 + 2 * kMargin;
    w.WZ = am1 - am0 + 1 + 2 * kMargin;
    if ((int64_t)w.WX * w.WZ > kWin2) w.on = false;
  }
}

// ===========================================================================
// FP64 form: the reference's arithmetic, expression by expression

struct F64Fields {
  const double *ex, *ey, *ez, *bx, *by, *bz;
};

// positions whose gather stencil lies inside the array
__device__ inline bool in_array64(const Geom2 &g, double x, double z, double dx) {
  const double xs = x / dx, zs = z / dx;
  return xs > -0.5 && xs < g.SX - 1 && zs > -0.5 && zs < g.SZ - 1;
}

// deposit gate of the particle loops: x in [dx,(Nx+3)dx), z in [2dx,(Nz+2)dx)
__device__ inline bool gate64(const Geom2 &g, double x, double z, double dx) {
  return x < (g.NX + 3.) * dx && x >= dx && z < (g.NZ + 2.) * dx && z >= 2. * dx;
}

// shape_func_*3 (:1948-2011)
__device__ inline void gather64(const Geom2 &g, const F64Fields &F, double x,
                                double z, double dx, double out[6]) {
  const int64_t dk = g.SZ;
  int k, m;
  int64_t i;
  double w0, w1, w2, w3;
  k = (int)(x / dx);
  m = (int)(z / dx - 0.5);
  i = g.idx(k, m);
  w0 = ((k + 1) * dx - x) * ((m + 1.5) * dx - z);
  w1 = (x - k * dx) * ((m + 1.5) * dx - z);
  w2 = ((k + 1) * dx - x) * (z - (m + 0.5) * dx);
  w3 = (x - k * dx) * (z - (m + 0.5) * dx);
  out[0] = (F.ex[i] * w0 + F.ex[i + dk] * w1 + F.ex[i + 1] * w2 +
            F.ex[i + dk + 1] * w3) /
           (dx * dx);
  out[5] = (F.bz[i] * w0 + F.bz[i + dk] * w1 + F.bz[i + 1] * w2 +
            F.bz[i + dk + 1] * w3) /
           (dx * dx);
  k = (int)(x / dx - 0.5);
  m = (int)(z / dx - 0.5);
  i = g.idx(k, m);
  out[1] = (F.ey[i] * ((k + 1.5) * dx - x) * ((m + 1.5) * dx - z) +
            F.ey[i + dk] * (x - (k + 0.5) * dx) * ((m + 1.5) * dx - z) +
            F.ey[i + 1] * ((k + 1.5) * dx - x) * (z - (m + 0.5) * dx) +
            F.ey[i + dk + 1] * (x - (k + 0.5) * dx) * (z - (m + 0.5) * dx)) /
           (dx * dx);
  k = (int)(x / dx - 0.5);
  m = (int)(z / dx);
  i = g.idx(k, m);
  w0 = ((k + 1.5) * dx - x) * ((m + 1) * dx - z);
  w1 = (x - (k + 0.5) * dx) * ((m + 1) * dx - z);
  w2 = ((k + 1.5) * dx - x) * (z - m * dx);
  w3 = (x - (k + 0.5) * dx) * (z - m * dx);
  out[2] = (F.ez[i] * w0 + F.ez[i + dk] * w1 + F.ez[i + 1] * w2 +
            F.ez[i + dk + 1] * w3) /
           (dx * dx);
  out[3] = (F.bx[i] * w0 + F.bx[i + dk] * w1 + F.bx[i + 1] * w2 +
            F.bx[i + dk + 1] * w3) /
           (dx * dx);
  k = (int)(x / dx);
  m = (int)(z / dx);
  i = g.idx(k, m);
  out[4] = (F.by[i] * ((k + 1) * dx - x) * ((m + 1) * dx - z) +
            F.by[i + dk] * (x - (k)*dx) * ((m + 1) * dx - z) +
            F.by[i + 1] * ((k + 1) * dx - x) * (z - (m)*dx) +
            F.by[i + dk + 1] * (x - (k)*dx) * (z - (m)*dx)) /
           (dx * dx);
}

// cal_track_c4 (:1520-1560)
__device__ inline void boris64(const Push2 &kp, const double f[6], double &x,
                               double &z, double &vx, double &vy, double &vz) {
  const double T = kp.T, dt = kp.dt;
  double ex = f[0], ey = f[1], ez = f[2];
  const double bx = f[3], by = f[4], bz = f[5];
  ex += kp.eimf_x;
  ey += kp.eimf_y;
  const double S = 2. * T / (1. + T * T * (bx * bx + by * by + bz * bz));
  const double ux = vx + ex * T;
  const double uy = vy + ey * T;
  const double uz = vz + ez * T;
  const double uox = ux + T * (uy * bz - uz * by);
  const double uoy = uy + T * (uz * bx - ux * bz);
  const double uoz = uz + T * (ux * by - uy * bx);
  vx = ux + S * (uoy * bz - uoz * by) + ex * T;
  vy = uy + S * (uoz * bx - uox * bz) + ey * T;
  vz = uz + S * (uox * by - uoy * bx) + ez * T;
  x += vx * dt * 0.5;
  z += vz * dt * 0.5;
}

__device__ inline void lds_add64(double *p, double v) {
  __hip_atomic_fetch_add((PIC_LDS double *)p, v, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void glb_add64(double *p, double v) {
  __hip_atomic_fetch_add((PIC_GLB double *)p, v, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}

// four node values of one deposit (shape_func_*0_* :2018-2144), to the LDS
// window or to the plane; lanes start at different nodes (the particles of
// a wave mostly share a cell)
__device__ inline void dep4_64(double *lds, double *glob, const Win2 &w,
                               const Geom2 &g, int k, int m, double v00,
                               double v10, double v01, double v11) {
  const double v[4] = {v00, v10, v01, v11};
  const int sw = threadIdx.x & 3;
  // the gate keeps nodes inside the array (for dx = 1 exactly; the check
  // only guards against x/dx rounding up to the gate's bound)
  if (k < 0 || m < 0 || k + 1 >= g.SX || m + 1 >= g.SZ) return;
  if (w.holds(k, m)) {
    const int i = (k - w.k0) * w.WZ + (m - w.m0);
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int id = n ^ sw;
      lds_add64(lds + i + ((id & 1) ? w.WZ : 0) + (id >> 1), v[id]);
    }
  } else {
    const int64_t i = g.idx(k, m);
#pragma unroll
    for (int n = 0; n < 4; ++n)
      glb_add64(glob + i + ((n & 1) ? g.SZ : 0) + (n >> 1), v[n]);
  }
}

__device__ inline void dep_jx64(double *lds, double *glob, const Win2 &w,
                                const Geom2 &g, double x, double z, double qn,
                                double vx, double dx) {
  const int k = (int)(x / dx), m = (int)(z / dx - 0.5);
  const double temp = qn * vx / (dx * dx);
  dep4_64(lds, glob, w, g, k, m,
          temp * ((k + 1) * dx - x) * ((m + 1.5) * dx - z),
          temp * (x - (k)*dx) * ((m + 1.5) * dx - z),
          temp * ((k + 1) * dx - x) * (z - (m + 0.5) * dx),
          temp * (x - (k)*dx) * (z - (m + 0.5) * dx));
}
__device__ inline void dep_jy64(double *lds, double *glob, const Win2 &w,
                                const Geom2 &g, double x, double z, double qn,
                                double vy, double dx) {
  const int k = (int)(x / dx - 0.5), m = (int)(z / dx - 0.5);
  const double temp = qn * vy / (dx * dx);
  dep4_64(lds, glob, w, g, k, m,
          temp * ((k + 1.5) * dx - x) * ((m + 1.5) * dx - z),
          temp * (x - (k + 0.5) * dx) * ((m + 1.5) * dx - z),
          temp * ((k + 1.5) * dx - x) * (z - (m + 0.5) * dx),
          temp * (x - (k + 0.5) * dx) * (z - (m + 0.5) * dx));
}
__device__ inline void dep_jz64(double *lds, double *glob, const Win2 &w,
                                const Geom2 &g, double x, double z, double qn,
                                double vz, double dx) {
  const int k = (int)(x / dx - 0.5), m = (int)(z / dx);
  const double temp = qn * vz / (dx * dx);
  dep4_64(lds, glob, w, g, k, m,
          temp * ((k + 1.5) * dx - x) * ((m + 1) * dx - z),
          temp * (x - (k + 0.5) * dx) * ((m + 1) * dx - z),
          temp * ((k + 1.5) * dx - x) * (z - (m)*dx),
          temp * (x - (k + 0.5) * dx) * (z - (m)*dx));
}
__device__ inline void dep_n64(double *lds, double *glob, const Win2 &w,
                               const Geom2 &g, double x, double z, double wt,
                               double dx) {
  const int k = (int)(x / dx - 0.5), m = (int)(z / dx - 0.5);
  const double temp = wt / (dx * dx);
  dep4_64(lds, glob, w, g, k, m,
          temp * ((k + 1.5) * dx - x) * ((m + 1.5) * dx - z),
          temp * (x - (k + 0.5) * dx) * ((m + 1.5) * dx - z),
          temp * ((k + 1.5) * dx - x) * (z - (m + 0.5) * dx),
          temp * (x - (k + 0.5) * dx) * (z - (m + 0.5) * dx));
}

struct Dep64 {
  double *jx, *jy, *jz, *n;
};

// PUSH_J: hot loop A (:665-720); DRIFT: the half drift of loop B; DEP_N: the
// density deposit of loop B (:752-788) or, without DRIFT, of the pre-loop
// pass (:586-612); BOUNDARY: the exchange and delete sweep (:812-901)
template <bool PUSH_J, bool DRIFT, bool DEP_N, bool BOUNDARY>
__global__ __launch_bounds__(256) void k_pass64(Geom2 g, Tile2 tg, F64Fields F,
                                                P64 P, Run2 tr, Push2 kp,
                                                Dep64 D) {
  __shared__ double acc[4 * kWin2];
  __shared__ Win2 sw;
  __shared__ int64_t s_range[2];
  const int tid = threadIdx.x;
  const int bid = xcd_remap(blockIdx.x, tr.ntiles, (int)gridDim.x);
  const double dx = kp.dx;
  if (tid == 0) {
    int64_t s, e;
    Win2 w;
    setup2(bid, tr, tg,
           [&](int64_t i, int &k, int &m) {
             const double x = P.x[i], z = P.z[i];
             if (!in_array64(g, x, z, dx)) return false;
             k = (int)(x / dx);
             m = (int)(z / dx);
             return true;
           },
           s, e, w);
    if (!(PUSH_J || DEP_N)) w.on = false;
    sw = w;
    s_range[0] = s;
    s_range[1] = e;
  }
  __syncthreads();
  const int64_t s = s_range[0], e = s_range[1];
  if (s >= e) return;
  const Win2 w = sw;
  const int WN = w.on ? w.WX * w.WZ : 0;
  for (int i = tid; i < 4 * WN; i += blockDim.x) acc[(i / WN) * kWin2 + i % WN] = 0.0;
  __syncthreads();
  for (int64_t i = s + tid; i < e; i += blockDim.x) {
    double x = P.x[i];
    if (isnan(x)) continue;
    double z = P.z[i], vx = P.vx[i], vy = P.vy[i], vz = P.vz[i];
    if (PUSH_J) {
      if (z < kp.blo[1] || z > kp.bhi[1] || x < kp.blo[0] || x > kp.bhi[0] ||
          !in_array64(g, x, z, dx)) {
        x += vx * kp.dt * 0.5;
        z += vz * kp.dt * 0.5;
      } else {
        double f[6];
        gather64(g, F, x, z, dx, f);
        boris64(kp, f, x, z, vx, vy, vz);
      }
      if (gate64(g, x, z, dx)) {
        dep_jy64(acc + 1 * kWin2, D.jy, w, g, x, z, kp.qw, vy, dx);
        dep_jx64(acc + 0 * kWin2, D.jx, w, g, x, z, kp.qw, vx, dx);
        dep_jz64(acc + 2 * kWin2, D.jz, w, g, x, z, kp.qw, vz, dx);
      }
    }
    if (DRIFT) {
      x += vx * kp.dt * 0.5;
      z += vz * kp.dt * 0.5;
    }
    if (DEP_N && gate64(g, x, z, dx)) dep_n64(acc + 3 * kWin2, D.n, w, g, x, z, kp.w, dx);
    if (BOUNDARY) {
      const double NXd = g.NX * dx;
      if (z < (g.NZ + 2.) * dx && z >= 2. * dx) {
        if (x < (g.NX + 3.) * dx && x >= (g.NX + 2.) * dx)
          x = x - NXd;
        else if (x > dx && x < 2. * dx)
          x = x + NXd;
      }
      if (x < 2. * dx || x >= (g.NX + 2.) * dx || z >= (g.NZ + 2.) * dx ||
          z < 2. * dx)
        x = __longlong_as_double(0x7ff8000000000000ll);
    }
    P.x[i] = x;
    P.z[i] = z;
    if (PUSH_J) {
      P.vx[i] = vx;
      P.vy[i] = vy;
      P.vz[i] = vz;
    }
  }
  if (!w.on) return;
  __syncthreads();
  for (int c = 0; c < 4; ++c) {
    if (c < 3 && !PUSH_J) continue;
    if (c == 3 && !DEP_N) continue;
    double *dst = c == 0 ? D.jx : c == 1 ? D.jy : c == 2 ? D.jz : D.n;
    for (int j = tid; j < WN; j += blockDim.x) {
      const double v = acc[c * kWin2 + j];
      if (v == 0.0) continue;
      const int kk = w.k0 + j / w.WZ, mm = w.m0 + j % w.WZ;
      if (kk < 0 || mm < 0 || kk >= g.SX || mm >= g.SZ) continue;
      glb_add64(dst + g.idx(kk, mm), v);
    }
  }
}

// ===========================================================================
// FP32 form: int32 cell + fp32 offsets.  This translation unit rounds every
// FP64 operation separately; the FP32 functions below may fuse.
#pragma clang fp contract(fast)

struct K32 {
  float T, hdt;          // q dt / (2m); dt / (2 dx)
  float qw, w;           // q w / dx^2, w / dx^2
  float eimf_x, eimf_y;
  float blo[2], bhi[2];  // ballistic bounds in cell units
  float fx;              // fixed-point units per unit value
  double fx_inv;
  uint32_t cmag;         // cell = k SZ + m: k = mulhi(cell, cmag) >> csh
  int csh;
};

// Exact k = n / SZ for 0 <= n < 2^31: with l = ceil(log2 SZ) and
// M = ceil(2^(31+l) / SZ) (below 2^32), n M / 2^(31+l) exceeds n / SZ by
// less than 1 / SZ, which never reaches the next integer.
inline void cell_magic(int SZ, uint32_t &mag, int &sh) {
  int l = 0;
  while ((1 << l) < SZ) ++l;
  if (l == 0) l = 1;
  const unsigned __int128 one = 1;
  mag = (uint32_t)(((one << (31 + l)) + (unsigned)SZ - 1) / (unsigned)SZ);
  sh = l - 1;
}

__device__ inline void split_cell(const Geom2 &g, const K32 &kp, int32_t cl, int &k,
                                  int &m) {
  k = (int)(__umulhi((uint32_t)cl, kp.cmag) >> kp.csh);
  m = cl - k * g.SZ;
}

struct F32Fields {
  const float *ex, *ey, *ez, *bx, *by, *bz;
};

struct Ax {
  int bi, bh;
  float i0, i1, h0, h1;
};

__device__ inline Ax axis2(int k, float o) {
  Ax a;
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

__device__ __forceinline__ const float *at(const float *base, uint32_t byte) {
  return (const float *)((const char *)base + byte);
}

// bilinear interpolation around the byte offset o of the stencil base (two
// adjacent m values per row; sk: byte stride of k)
__device__ __forceinline__ float lerp4o(const float *F, uint32_t o, uint32_t sk,
                                        float a0, float a1, float c0, float c1) {
  const float *p0 = at(F, o), *p1 = at(F, o + sk);
  return a0 * (c0 * p0[0] + c1 * p0[1]) + a1 * (c0 * p1[0] + c1 * p1[1]);
}

// shape_func_*3 (:1948-2011): the six staggered components from uniform plane
// bases and 32-bit byte offsets (the planes are below 2^31 bytes); o is the
// byte offset of the particle's cell, hk / hm the byte shifts of the
// half-node bases
__device__ inline void gather32(const F32Fields &F, uint32_t o, uint32_t sk,
                                uint32_t hk, uint32_t hm, const Ax &X, const Ax &Z,
                                float f[6]) {
  f[0] = lerp4o(F.ex, o - hm, sk, X.i0, X.i1, Z.h0, Z.h1);
  f[1] = lerp4o(F.ey, o - hk - hm, sk, X.h0, X.h1, Z.h0, Z.h1);
  f[2] = lerp4o(F.ez, o - hk, sk, X.h0, X.h1, Z.i0, Z.i1);
  f[3] = lerp4o(F.bx, o - hk, sk, X.h0, X.h1, Z.i0, Z.i1);
  f[4] = lerp4o(F.by, o, sk, X.i0, X.i1, Z.i0, Z.i1);
  f[5] = lerp4o(F.bz, o - hm, sk, X.i0, X.i1, Z.h0, Z.h1);
}

__device__ inline void boris32(const K32 &p, const float f[6], float &vx,
                               float &vy, float &vz) {
  const float T = p.T;
  const float ex = f[0] + p.eimf_x, ey = f[1] + p.eimf_y, ez = f[2];
  const float bx = f[3], by = f[4], bz = f[5];
  const float S = __fdividef(2.0f * T, 1.0f + T * T * (bx * bx + by * by + bz * bz));
  const float ux = vx + ex * T, uy = vy + ey * T, uz = vz + ez * T;
  const float ox = ux + T * (uy * bz - uz * by);
  const float oy = uy + T * (uz * bx - ux * bz);
  const float oz = uz + T * (ux * by - uy * bx);
  vx = ux + S * (oy * bz - oz * by) + ex * T;
  vy = uy + S * (oz * bx - ox * bz) + ey * T;
  vz = uz + S * (ox * by - oy * bx) + ez * T;
}

__device__ inline void drift32(int &k, float &o, float v, float hdt) {
  o += v * hdt;
  if (o >= 1.0f) {
    o -= 1.0f;
    k += 1;
    if (o >= 1.0f) {  // more than a cell per half step: keep it representable
      k += (int)o;
      o -= floorf(o);
    }
  } else if (o < 0.0f) {
    o += 1.0f;
    k -= 1;
    if (o < 0.0f) {
      const float fl = floorf(o);
      k += (int)fl;
      o -= fl;
    }
    if (o >= 1.0f) o = 0.99999994f;
  }
}

// Fixed-point contribution: the scale is folded into the deposit factor;
// the scaled value is rounded to the nearest unit and sign-extended.
__device__ inline u64 to_fx(float v) { return (u64)(long long)(int)rintf(v); }

// The lane's walk over the 4 nodes of a stencil.  Lanes of a wave mostly
// hold particles of the same cell (the data are cell sorted; at 100 per cell
// a wave's 256 particles span about 3 cells) and their LDS atomics to one
// node serialise.  Lane j visits node n ^ (j & 3) at step n, which spreads
// one instruction's lanes over the four nodes, and accumulates into copy
// (j >> 2) mod ncopy of the window (ncopy copies of WN nodes side by side
// in each plane, summed at the flush).
struct Walk2 {
  bool mx, mz;  // axes walked downwards
  int off[4];   // node offsets of the visits from the stencil base
};

__device__ inline Walk2 make_walk2(int WZ, int WN, int ncopy) {
  Walk2 k;
  const int sw = threadIdx.x & 3;
  const int copy = ((int)threadIdx.x >> 2) & (ncopy - 1);
  k.mx = sw & 1;
  k.mz = sw & 2;
  for (int n = 0; n < 4; ++n) {
    const int id = n ^ sw;
    k.off[n] = copy * WN + ((id & 1) ? WZ : 0) + (id >> 1);
  }
  return k;
}

// q a_i c_j at the 4 window nodes of a stencil based at window index i; q
// carries the fixed-point scale
__device__ __forceinline__ void lds_dep4(u64 *lds, int i, const Walk2 &wk, float q,
                                         float a0, float a1, float c0, float c1) {
  const float A0 = wk.mx ? a1 : a0, A1 = wk.mx ? a0 : a1;
  const float C0 = wk.mz ? c1 : c0, C1 = wk.mz ? c0 : c1;
  const float qa0 = q * A0, qa1 = q * A1;
  const float v[4] = {qa0 * C0, qa1 * C0, qa0 * C1, qa1 * C1};
#pragma unroll
  for (int n = 0; n < 4; ++n)
    __hip_atomic_fetch_add((PIC_LDS u64 *)(lds + i + wk.off[n]), to_fx(v[n]),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

struct Dep32 {
  float *jx, *jy, *jz, *n;
};

__device__ inline void glb_add32(float *p, float v) {
  __hip_atomic_fetch_add((PIC_GLB float *)p, v, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}

// one lane's stencil straight into a global plane of row stride SZ (taken by
// value: a reference to the kernel's geometry argument from this out-of-line
// function would keep a copy of it in private memory)
__device__ __noinline__ void glob_dep4(float *glob, int SZ, int k, int m, float q,
                                       float a0, float a1, float c0, float c1) {
  const int64_t i = (int64_t)k * SZ + m;
  const float qa0 = q * a0, qa1 = q * a1;
  glb_add32(glob + i, qa0 * c0);
  glb_add32(glob + i + SZ, qa1 * c0);
  glb_add32(glob + i + 1, qa0 * c1);
  glb_add32(glob + i + SZ + 1, qa1 * c1);
}

constexpr int kCoopMax = 8;

// Deposit of the particles outside the window (pend) into the global planes:
// the wave takes them one at a time, each active lane adding one node of one
// component (J: 12 nodes, n: 4); waves with more than kCoopMax such
// particles deposit lane by lane.  Reached by every active lane of the wave.
__device__ __forceinline__ void glob_coop2(const Geom2 &g, const Dep32 &D, bool pend,
                                           bool current, int k, int m, float ox,
                                           float oz, float q0, float q1, float q2) {
  u64 todo = __builtin_amdgcn_ballot_w64(pend);
  if (todo == 0) return;
  if (__builtin_popcountll(todo) > kCoopMax) {
    if (pend) {
      const Ax X = axis2(k, ox), Z = axis2(m, oz);
      if (current) {
        glob_dep4(D.jx, g.SZ, X.bi, Z.bh, q0, X.i0, X.i1, Z.h0, Z.h1);
        glob_dep4(D.jy, g.SZ, X.bh, Z.bh, q1, X.h0, X.h1, Z.h0, Z.h1);
        glob_dep4(D.jz, g.SZ, X.bh, Z.bi, q2, X.h0, X.h1, Z.i0, Z.i1);
      } else {
        glob_dep4(D.n, g.SZ, X.bh, Z.bh, q0, X.h0, X.h1, Z.h0, Z.h1);
      }
    }
    return;
  }
  const u64 ex = __builtin_amdgcn_read_exec();
  const int rank = (int)__builtin_amdgcn_mbcnt_hi(
      (uint32_t)(ex >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ex, 0u));
  const int nact = __builtin_popcountll(ex);
  const int nn = current ? 12 : 4;
  do {
    const int src = __builtin_ctzll(todo);
    todo &= todo - 1;
    const int sk = __builtin_amdgcn_readlane(k, src);
    const int sm = __builtin_amdgcn_readlane(m, src);
    const float sx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ox), src));
    const float sz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(oz), src));
    const float s0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(q0), src));
    float s1 = 0.0f, s2 = 0.0f;
    if (current) {
      s1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(q1), src));
      s2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(q2), src));
    }
    const Ax X = axis2(sk, sx), Z = axis2(sm, sz);
    for (int t = rank; t < nn; t += nact) {
      // node t: component t >> 2 (3: density), corner bits (x, z) of t & 3;
      // x takes the integer-node weights for Jx, z for Jz
      const int c = current ? t >> 2 : 3;
      const bool bx = t & 1, bz = t & 2;
      const float wx = c == 0 ? (bx ? X.i1 : X.i0) : (bx ? X.h1 : X.h0);
      const float wz = c == 2 ? (bz ? Z.i1 : Z.i0) : (bz ? Z.h1 : Z.h0);
      const int kx = (c == 0 ? X.bi : X.bh) + (bx ? 1 : 0);
      const int mz = (c == 2 ? Z.bi : Z.bh) + (bz ? 1 : 0);
      const float q = c == 1 ? s1 : c == 2 ? s2 : s0;
      float *dst = c == 0 ? D.jx : c == 1 ? D.jy : c == 2 ? D.jz : D.n;
      glb_add32(dst + g.idx(kx, mz), (q * wx) * wz);
    }
  } while (todo);
}

// the stencils of a particle at cell (k, m) (bases k-1..k, m-1..m) lie in the
// window
__device__ inline bool in_win2(const Win2 &w, int k, int m) {
  return w.on && (unsigned)(k - 1 - w.k0) <= (unsigned)(w.WX - 3) &&
         (unsigned)(m - 1 - w.m0) <= (unsigned)(w.WZ - 3);
}

template <bool PUSH_J, bool DRIFT, bool DEP_N, bool BOUNDARY>
__device__ __forceinline__ void step32(const Geom2 &g, const F32Fields &F,
                                       const K32 &kp, const Win2 &w,
                                       const Walk2 &wk, u64 *acc,
                                       const Dep32 &D, int32_t &cl, float &ox,
                                       float &oz, float &vx, float &vy,
                                       float &vz) {
  if (cl < 0) return;
  int k, m;
  split_cell(g, kp, cl, k, m);
  // deposits only from cells whose stencils lie inside the array
  const auto dep_gate = [&]() {
    return (unsigned)(k - 1) <= (unsigned)(g.NX + 1) &&
           (unsigned)(m - 2) <= (unsigned)(g.NZ - 1);
  };
  if (PUSH_J) {
    const float xf = (float)k + ox, zf = (float)m + oz;
    const bool safe = (unsigned)(k - 1) <= (unsigned)(g.SX - 3) &&
                      (unsigned)(m - 1) <= (unsigned)(g.SZ - 3);
    if (safe && !(zf < kp.blo[1] || zf > kp.bhi[1] || xf < kp.blo[0] ||
                  xf > kp.bhi[0])) {
      const uint32_t sk = 4u * (uint32_t)g.SZ;
      const uint32_t o = ((uint32_t)k * (uint32_t)g.SZ + (uint32_t)m) * 4u;
      const Ax X = axis2(k, ox), Z = axis2(m, oz);
      float f[6];
      gather32(F, o, sk, X.bh != k ? sk : 0u, Z.bh != m ? 4u : 0u, X, Z, f);
      boris32(kp, f, vx, vy, vz);
    }
    drift32(k, ox, vx, kp.hdt);
    drift32(m, oz, vz, kp.hdt);
    const bool dep = dep_gate();
    const bool lds_ok = in_win2(w, k, m);
    const float qs = kp.qw * kp.fx;
    if (dep && lds_ok) {
      const Ax X = axis2(k, ox), Z = axis2(m, oz);
      const int c0 = (k - w.k0) * w.WZ + (m - w.m0);
      const int hx = X.bh != k ? w.WZ : 0, hz = Z.bh != m ? 1 : 0;
      lds_dep4(acc + 0 * kWin2, c0 - hz, wk, qs * vx, X.i0, X.i1, Z.h0, Z.h1);
      lds_dep4(acc + 1 * kWin2, c0 - hx - hz, wk, qs * vy, X.h0, X.h1, Z.h0, Z.h1);
      lds_dep4(acc + 2 * kWin2, c0 - hx, wk, qs * vz, X.h0, X.h1, Z.i0, Z.i1);
    }
    glob_coop2(g, D, dep && !lds_ok, true, k, m, ox, oz, kp.qw * vx, kp.qw * vy,
               kp.qw * vz);
  }
  if (DRIFT) {
    drift32(k, ox, vx, kp.hdt);
    drift32(m, oz, vz, kp.hdt);
  }
  if (DEP_N) {
    const bool dep = dep_gate();
    const bool lds_ok = in_win2(w, k, m);
    if (dep && lds_ok) {
      const Ax X = axis2(k, ox), Z = axis2(m, oz);
      const int c0 = (k - w.k0) * w.WZ + (m - w.m0);
      const int hx = X.bh != k ? w.WZ : 0, hz = Z.bh != m ? 1 : 0;
      lds_dep4(acc + 3 * kWin2, c0 - hx - hz, wk, kp.w * kp.fx, X.h0, X.h1, Z.h0,
               Z.h1);
    }
    glob_coop2(g, D, dep && !lds_ok, false, k, m, ox, oz, kp.w, 0.0f, 0.0f);
  }
  bool dead = k < 0 || m < 0 || k >= g.SX || m >= g.SZ;
  if (BOUNDARY && !dead) {
    if (m >= 2 && m < g.NZ + 2) {
      if (k == g.NX + 2)
        k -= g.NX;
      else if (k == 1 && ox > 0.0f)
        k += g.NX;
    }
    dead = k < 2 || k >= g.NX + 2 || m < 2 || m >= g.NZ + 2;
  }
  cl = dead ? -1 : k * g.SZ + m;
}

// PERM: the pass also applies a sort's permutation: slot i takes its
// particle from slot perm[i] of Q and stores the result in P (the other
// buffer of the species), so the sort needs no separate copy
template <bool PUSH_J, bool DRIFT, bool DEP_N, bool BOUNDARY, bool PERM>
__global__ __launch_bounds__(256) void k_pass32(Geom2 g, Tile2 tg, F32Fields F,
                                                P32 P, Run2 tr, K32 kp, Dep32 D,
                                                P32 Q, const uint32_t *__restrict__ perm) {
  __shared__ u64 acc[4 * kWin2];
  __shared__ Win2 sw;
  __shared__ int64_t s_range[2];
  const int tid = threadIdx.x;
  const int bid = xcd_remap(blockIdx.x, tr.ntiles, (int)gridDim.x);
  if (tid == 0) {
    int64_t s, e;
    Win2 w;
    setup2(bid, tr, tg,
           [&](int64_t i, int &k, int &m) {
             const int32_t c = PERM ? Q.cell[perm[i]] : P.cell[i];
             if (c < 0) return false;
             split_cell(g, kp, c, k, m);
             return true;
           },
           s, e, w);
    if (!(PUSH_J || DEP_N)) w.on = false;
    sw = w;
    s_range[0] = s;
    s_range[1] = e;
  }
  __syncthreads();
  const int64_t s = s_range[0], e = s_range[1];
  if (s >= e) return;
  const Win2 w = sw;
  const int WN = w.on ? w.WX * w.WZ : 0;
  const int ncopy = 4 * WN <= kWin2 ? 4 : 2 * WN <= kWin2 ? 2 : 1;
  const int WC = ncopy * WN;  // nodes in use per plane
  for (int i = tid; i < 4 * WC; i += blockDim.x) acc[(i / WC) * kWin2 + i % WC] = 0;
  __syncthreads();
  const Walk2 wk = make_walk2(w.WZ, WN, ncopy);
  // four consecutive particles per thread with 16-byte accesses; slots of a
  // boundary quad outside [s, e) belong to a neighbouring range
  const int64_t q0 = s >> 2, q1 = (e + 3) >> 2;
  for (int64_t q = q0 + tid; q < q1; q += blockDim.x) {
    const int64_t i0 = q << 2;
    int cs[4];
    float xs[4], zs[4], us[4], vs[4], ws[4];
    if (PERM) {
      // the capacity is a multiple of 256, so the quad's perm entries are
      // allocated; slots outside the range read slot 0 and are not stored
      const uint4 p4 = *(const uint4 *)(perm + i0);
      const uint32_t pj[4] = {p4.x, p4.y, p4.z, p4.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t i = i0 + j;
        const uint32_t o = i >= s && i < e ? pj[j] : 0u;
        cs[j] = Q.cell[o];
        xs[j] = Q.ox[o];
        zs[j] = Q.oz[o];
        us[j] = Q.vx[o];
        vs[j] = Q.vy[o];
        ws[j] = Q.vz[o];
      }
    } else {
      const int4 c4 = *(const int4 *)(P.cell + i0);
      const float4 x4 = *(const float4 *)(P.ox + i0);
      const float4 z4 = *(const float4 *)(P.oz + i0);
      const float4 u4 = *(const float4 *)(P.vx + i0);
      const float4 v4 = *(const float4 *)(P.vy + i0);
      const float4 w4 = *(const float4 *)(P.vz + i0);
      cs[0] = c4.x, cs[1] = c4.y, cs[2] = c4.z, cs[3] = c4.w;
      xs[0] = x4.x, xs[1] = x4.y, xs[2] = x4.z, xs[3] = x4.w;
      zs[0] = z4.x, zs[1] = z4.y, zs[2] = z4.z, zs[3] = z4.w;
      us[0] = u4.x, us[1] = u4.y, us[2] = u4.z, us[3] = u4.w;
      vs[0] = v4.x, vs[1] = v4.y, vs[2] = v4.z, vs[3] = v4.w;
      ws[0] = w4.x, ws[1] = w4.y, ws[2] = w4.z, ws[3] = w4.w;
    }
    const bool full = i0 >= s && i0 + 3 < e;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t i = i0 + j;
      if (i < s || i >= e) continue;
      step32<PUSH_J, DRIFT, DEP_N, BOUNDARY>(g, F, kp, w, wk, acc, D, cs[j], xs[j],
                                             zs[j], us[j], vs[j], ws[j]);
    }
    if (full) {
      *(int4 *)(P.cell + i0) = make_int4(cs[0], cs[1], cs[2], cs[3]);
      *(float4 *)(P.ox + i0) = make_float4(xs[0], xs[1], xs[2], xs[3]);
      *(float4 *)(P.oz + i0) = make_float4(zs[0], zs[1], zs[2], zs[3]);
      if (PUSH_J) {
        *(float4 *)(P.vx + i0) = make_float4(us[0], us[1], us[2], us[3]);
        *(float4 *)(P.vy + i0) = make_float4(vs[0], vs[1], vs[2], vs[3]);
        *(float4 *)(P.vz + i0) = make_float4(ws[0], ws[1], ws[2], ws[3]);
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t i = i0 + j;
        if (i < s || i >= e) continue;
        P.cell[i] = cs[j];
        P.ox[i] = xs[j];
        P.oz[i] = zs[j];
        if (PUSH_J) {
          P.vx[i] = us[j];
          P.vy[i] = vs[j];
          P.vz[i] = ws[j];
        }
      }
    }
  }
  if (!w.on) return;
  __syncthreads();
  const double inv_s = kp.fx_inv;
  for (int c = 0; c < 4; ++c) {
    if (c < 3 && !PUSH_J) continue;
    if (c == 3 && !DEP_N) continue;
    float *dst = c == 0 ? D.jx : c == 1 ? D.jy : c == 2 ? D.jz : D.n;
    for (int j = tid; j < WN; j += blockDim.x) {
      u64 v = 0;
      for (int cp = 0; cp < ncopy; ++cp) v += acc[c * kWin2 + cp * WN + j];
      if (v == 0) continue;
      const int kk = w.k0 + j / w.WZ, mm = w.m0 + j % w.WZ;
      if (kk < 0 || mm < 0 || kk >= g.SX || mm >= g.SZ) continue;
      glb_add32(dst + g.idx(kk, mm), (float)((double)(long long)v * inv_s));
    }
  }
}

#pragma clang fp contract(off)

// ===========================================================================
// host side of the pass

Push2 push_const(const pic2d_ctx *ctx, int sp) {
  const pic2d_config &c = ctx->cfg;
  Push2 p;
  p.dx = c.dx;
  p.dt = c.dt;
  p.q = c.q;
  const double m = sp == 0 ? c.mass[0] : -c.mass[1];
  p.T = 0.5 * c.q * c.dt / m;
  p.qw = c.q * c.weight[sp];
  p.w = c.weight[sp];
  p.eimf_x = c.vdrift * c.imf[1];
  p.eimf_y = -c.vdrift * c.imf[0];
  for (int a = 0; a < 2; ++a) {
    p.blo[a] = c.ballistic_lo[a];
    p.bhi[a] = c.ballistic_hi[a];
  }
  return p;
}

static Run2 run_of(pic2d_ctx *ctx, int sp, unsigned *nblocks) {
  Run2 tr;
  tr.tile_start = ctx->tile_start[sp];
  tr.ntiles = ctx->tg.ntiles();
  tr.sorted_end = ctx->sorted_end[sp];
  tr.n = ctx->n[sp];
  const int64_t ntail = div_up(tr.n - tr.sorted_end, kTailChunk);
  *nblocks = (unsigned)(tr.ntiles + ntail);
  return tr;
}

static uint32_t *pass_perm(pic2d_ctx *ctx, int sp);

int push(pic2d_ctx *ctx, int sp, int push_j, int drift_n, int boundary,
         int density_only) {
  // a pending sort permutation is applied by the fused pass only
  const bool fused = push_j && drift_n && boundary && !density_only;
  if (!fused) {
    int rc = settle(ctx, sp);
    if (rc) return rc;
  }
  if (ctx->n[sp] == 0) return 0;
  unsigned nb;
  const Run2 tr = run_of(ctx, sp, &nb);
  const int jx = sp == 0 ? PIC_JIX : PIC_JEX, nn = sp == 0 ? PIC_NI : PIC_NE;
  const int sel = density_only ? 8 : (push_j ? 4 : 0) | (drift_n ? 2 : 0) | (boundary ? 1 : 0);
  hipStream_t st = ctx->rt.stream;
  if (ctx->f64) {
    F64Fields F = {plane<double>(ctx, PIC_EX), plane<double>(ctx, PIC_EY),
                   plane<double>(ctx, PIC_EZ), plane<double>(ctx, PIC_BX),
                   plane<double>(ctx, PIC_BY), plane<double>(ctx, PIC_BZ)};
    Dep64 D = {plane<double>(ctx, jx), plane<double>(ctx, jx + 1),
               plane<double>(ctx, jx + 2), plane<double>(ctx, nn)};
    const P64 P = p64(ctx, sp, ctx->cur[sp]);
    const Push2 kp = push_const(ctx, sp);
#define L64(a, b, c, d) \
  k_pass64<a, b, c, d><<<nb, 256, 0, st>>>(ctx->g, ctx->tg, F, P, tr, kp, D)
    switch (sel) {
      case 7: L64(true, true, true, true); break;
      case 4: L64(true, false, false, false); break;
      case 2: L64(false, true, true, false); break;
      case 1: L64(false, false, false, true); break;
      case 8: L64(false, false, true, false); break;
      default: return err(ctx, PIC_E_ARG, "push: unsupported routine combination");
    }
#undef L64
  } else {
    F32Fields F = {plane<float>(ctx, PIC_EX), plane<float>(ctx, PIC_EY),
                   plane<float>(ctx, PIC_EZ), plane<float>(ctx, PIC_BX),
                   plane<float>(ctx, PIC_BY), plane<float>(ctx, PIC_BZ)};
    Dep32 D = {plane<float>(ctx, jx), plane<float>(ctx, jx + 1),
               plane<float>(ctx, jx + 2), plane<float>(ctx, nn)};
    // with a pending permutation the pass reads the current buffer through
    // it and writes the other one
    const uint32_t *pm = ctx->pending[sp] ? pass_perm(ctx, sp) : nullptr;
    const P32 Q = p32(ctx, sp, ctx->cur[sp]);
    const P32 P = pm ? p32(ctx, sp, 1 - ctx->cur[sp]) : Q;
    const pic2d_config &c = ctx->cfg;
    const Push2 p = push_const(ctx, sp);
    K32 kp;
    kp.T = (float)p.T;
    kp.hdt = (float)(c.dt / (2.0 * c.dx));
    kp.qw = (float)(p.qw / (c.dx * c.dx));
    kp.w = (float)(p.w / (c.dx * c.dx));
    kp.eimf_x = (float)p.eimf_x;
    kp.eimf_y = (float)p.eimf_y;
    for (int a = 0; a < 2; ++a) {
      kp.blo[a] = (float)std::max(-1e30, std::min(1e30, p.blo[a] / c.dx));
      kp.bhi[a] = (float)std::max(-1e30, std::min(1e30, p.bhi[a] / c.dx));
    }
    kp.fx = (float)ctx->fx_scale;
    kp.fx_inv = 1.0 / ctx->fx_scale;
    cell_magic(ctx->g.SZ, kp.cmag, kp.csh);
#define L32(a, b, c, d) \
  k_pass32<a, b, c, d, false><<<nb, 256, 0, st>>>(ctx->g, ctx->tg, F, P, tr, kp, D, Q, pm)
    switch (sel) {
      case 7:
        if (pm)
          k_pass32<true, true, true, true, true>
              <<<nb, 256, 0, st>>>(ctx->g, ctx->tg, F, P, tr, kp, D, Q, pm);
        else
          L32(true, true, true, true);
        break;
      case 4: L32(true, false, false, false); break;
      case 2: L32(false, true, true, false); break;
      case 1: L32(false, false, false, true); break;
      case 8: L32(false, false, true, false); break;
      default: return err(ctx, PIC_E_ARG, "push: unsupported routine combination");
    }
#undef L32
    P2_LAUNCH_CHECK(ctx);
    if (pm) {
      ctx->cur[sp] = 1 - ctx->cur[sp];
      ctx->pending[sp] = false;
    }
    return 0;
  }
  P2_LAUNCH_CHECK(ctx);
  return 0;
}

int particle_boundary(pic2d_ctx *ctx, int sp) { return push(ctx, sp, 0, 0, 1, 0); }

// ===========================================================================
// loading and inflow: Philox streams keyed (seed, species, event, index)

// interior cell (k, m) of particle slot i when the slots follow the tile
// order, ppc particles per cell
__device__ inline void cell_of_slot(const Tile2 &tg, int64_t i, int ppc, int &k,
                                    int &m) {
  const int64_t key = i / ppc;
  const int cpt = tg.TX * tg.TZ;
  const int tile = (int)(key / cpt), in = (int)(key % cpt);
  k = (tile / tg.NTZ) * tg.TX + in / tg.TZ;
  m = (tile % tg.NTZ) * tg.TZ + in % tg.TZ;
}

__device__ inline float unit_f32(double u) {
  const float f = (float)u;
  return f < 1.0f ? f : 0.99999994f;
}

// init_particle (:389-525)
__global__ void k_load2(Geom2 g, Tile2 tg, int f64, P64 P, P32 Q, int64_t n,
                        int ppc, double dx, double vth, double V, uint64_t seed,
                        int sp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int k, m;
  cell_of_slot(tg, i, ppc, k, m);
  const pic::ParticleDraw d = pic::draw_particle(seed, sp, 0u, (uint64_t)i);
  const double vx = vth * d.n0 + 0.0, vz = vth * d.n2 + V;
  if (f64) {
    P.x[i] = (k + 2) * dx + d.ux * dx;
    P.z[i] = (m + 2) * dx + d.uz * dx;
    P.vx[i] = vx;
    P.vy[i] = 0.0;
    P.vz[i] = vz;
  } else {
    Q.cell[i] = (k + 2) * g.SZ + (m + 2);
    Q.ox[i] = unit_f32(d.ux);
    Q.oz[i] = unit_f32(d.uz);
    Q.vx[i] = (float)vx;
    Q.vy[i] = 0.0f;
    Q.vz[i] = (float)vz;
  }
}

static int set_tiles_loaded(pic2d_ctx *ctx, int sp, int64_t n) {
  const int nt = ctx->tg.ntiles();
  const int64_t per = (int64_t)ctx->tg.TX * ctx->tg.TZ * ctx->cfg.ppc;
  uint32_t *h = (uint32_t *)malloc(sizeof(uint32_t) * (nt + 1));
  if (!h) return err(ctx, PIC_E_NOMEM, "tile table");
  for (int t = 0; t <= nt; ++t) h[t] = (uint32_t)std::min<int64_t>(n, per * t);
  hipError_t e = hipMemcpy(ctx->tile_start[sp], h, sizeof(uint32_t) * (nt + 1),
                           hipMemcpyHostToDevice);
  free(h);
  P2_HIP(ctx, e);
  ctx->sorted_end[sp] = n;
  return 0;
}

static int clear_tiles(pic2d_ctx *ctx, int sp) {
  const int nt = ctx->tg.ntiles();
  P2_HIP(ctx, hipMemsetAsync(ctx->tile_start[sp], 0, sizeof(uint32_t) * (nt + 1),
                             ctx->rt.stream));
  ctx->sorted_end[sp] = 0;
  return 0;
}

int load_particles(pic2d_ctx *ctx) {
  {
    int rc = settle_sorts(ctx);
    if (rc) return rc;
  }
  const pic2d_config &c = ctx->cfg;
  const int64_t n = (int64_t)c.nx * c.nz * c.ppc;
  if (n > ctx->cap)
    return pic::set_error(&ctx->rt, PIC_E_CAPACITY,
                          "load: %lld particles > capacity %lld", (long long)n,
                          (long long)ctx->cap);
  for (int sp = 0; sp < 2; ++sp) {
    if (n > 0)
      k_load2<<<(unsigned)div_up(n, kBlock), kBlock, 0, ctx->rt.stream>>>(
          ctx->g, ctx->tg, ctx->f64 ? 1 : 0, p64(ctx, sp, ctx->cur[sp]),
          p32(ctx, sp, ctx->cur[sp]), n, c.ppc, c.dx, c.vth[sp], c.vdrift,
          c.seed, sp);
    P2_LAUNCH_CHECK(ctx);
    ctx->n[sp] = n;
    int rc = set_tiles_loaded(ctx, sp, n);
    if (rc) return rc;
  }
  return 0;
}

// add_*5 (:3311-3566), first half: delete z < 3dx or z > (Nz+1)dx
__global__ void k_inflow_kill2(Geom2 g, int f64, P64 P, P32 Q, int64_t n,
                               double dx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (f64) {
    const double z = P.z[i];
    if (isnan(P.x[i])) return;
    if (z < 3 * dx || z > (g.NZ + 1) * dx)
      P.x[i] = __longlong_as_double(0x7ff8000000000000ll);
  } else {
    const int32_t c = Q.cell[i];
    if (c < 0) return;
    const int m = c % g.SZ;
    if (m < 3 || m > g.NZ + 1 || (m == g.NZ + 1 && Q.oz[i] > 0.0f)) Q.cell[i] = -1;
  }
}

// second half: ppc particles per cell of the first and last interior rows,
// low face first, cells in k order; batch index b keys the stream
__global__ void k_refill2(Geom2 g, int f64, P64 P, P32 Q, int64_t base,
                          int64_t per_face, int ppc, double dx, double vth,
                          double V, uint64_t seed, int sp, uint32_t event) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= 2 * per_face) return;
  const int face = b >= per_face ? 1 : 0;
  const int k = (int)((b - face * per_face) / ppc);
  const pic::ParticleDraw d = pic::draw_particle(seed, sp, event, (uint64_t)b);
  const double vx = vth * d.n0 + 0.0, vz = vth * d.n2 + V;
  const int64_t i = base + b;
  if (f64) {
    const double zlo = face == 0 ? 2 * dx : (g.NZ + 1) * dx;
    P.x[i] = (k + 2) * dx + d.ux * dx;
    P.z[i] = zlo + d.uz * dx;
    P.vx[i] = vx;
    P.vy[i] = 0.0;
    P.vz[i] = vz;
  } else {
    const int mrow = face == 0 ? 2 : g.NZ + 1;
    Q.cell[i] = (k + 2) * g.SZ + mrow;
    Q.ox[i] = unit_f32(d.ux);
    Q.oz[i] = unit_f32(d.uz);
    Q.vx[i] = (float)vx;
    Q.vy[i] = 0.0f;
    Q.vz[i] = (float)vz;
  }
}

// live flags for compaction
__global__ void k_live_flags2(int f64, P64 P, P32 Q, int64_t n, uint8_t *flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flags[i] = f64 ? !isnan(P.x[i]) : Q.cell[i] >= 0;
}

__global__ void k_gather2(int f64, P64 Pd, P32 Qd, P64 Ps, P32 Qs,
                          const uint32_t *__restrict__ idx, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = idx[i];
  if (f64) {
    Pd.x[i] = Ps.x[j];
    Pd.z[i] = Ps.z[j];
    Pd.vx[i] = Ps.vx[j];
    Pd.vy[i] = Ps.vy[j];
    Pd.vz[i] = Ps.vz[j];
  } else {
    Qd.cell[i] = Qs.cell[j];
    Qd.ox[i] = Qs.ox[j];
    Qd.oz[i] = Qs.oz[j];
    Qd.vx[i] = Qs.vx[j];
    Qd.vy[i] = Qs.vy[j];
    Qd.vz[i] = Qs.vz[j];
  }
}

// the FP32 form's permutation gather, four slots per thread (the index as
// one uint4, each array as one 16-byte store; the arrays start at multiples
// of the capacity, itself a multiple of 256)
template <typename T, typename V>
__device__ inline void gather4_2(T *__restrict__ d, const T *__restrict__ s, const uint4 &j) {
  V v;
  v.x = s[j.x];
  v.y = s[j.y];
  v.z = s[j.z];
  v.w = s[j.w];
  *(V *)d = v;
}

__global__ void k_gather2_32(P32 d, P32 s, const uint32_t *__restrict__ idx, int64_t n) {
  const int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (i0 >= n) return;
  if (i0 + 4 <= n) {
    const uint4 j = *(const uint4 *)(idx + i0);
    gather4_2<int32_t, int4>(d.cell + i0, s.cell, j);
    gather4_2<float, float4>(d.ox + i0, s.ox, j);
    gather4_2<float, float4>(d.oz + i0, s.oz, j);
    gather4_2<float, float4>(d.vx + i0, s.vx, j);
    gather4_2<float, float4>(d.vy + i0, s.vy, j);
    gather4_2<float, float4>(d.vz + i0, s.vz, j);
    return;
  }
  for (int64_t i = i0; i < n; ++i) {
    const uint32_t j = idx[i];
    d.cell[i] = s.cell[j];
    d.ox[i] = s.ox[j];
    d.oz[i] = s.oz[j];
    d.vx[i] = s.vx[j];
    d.vy[i] = s.vy[j];
    d.vz[i] = s.vz[j];
  }
}

// dst[i] = src[idx[i]] for i < n in the context's form
static int gather_perm(pic2d_ctx *ctx, int sp, int src, int dst, int64_t n,
                       const uint32_t *idx) {
  if (n <= 0) return 0;
  if (ctx->f64)
    k_gather2<<<(unsigned)div_up(n, kBlock), kBlock, 0, ctx->rt.stream>>>(
        1, p64(ctx, sp, dst), p32(ctx, sp, dst), p64(ctx, sp, src), p32(ctx, sp, src),
        idx, n);
  else
    k_gather2_32<<<(unsigned)div_up(n, 4 * kBlock), kBlock, 0, ctx->rt.stream>>>(
        p32(ctx, sp, dst), p32(ctx, sp, src), idx, n);
  P2_LAUNCH_CHECK(ctx);
  return 0;
}

// where a deferred permutation waits for its pass
static uint32_t *pass_perm(pic2d_ctx *ctx, int sp) {
  return sp == 0 ? ctx->perm_ion : ctx->rt.idx;
}

int settle(pic2d_ctx *ctx, int sp) {
  if (!ctx->pending[sp]) return 0;
  const int src = ctx->cur[sp];
  int rc = gather_perm(ctx, sp, src, 1 - src, ctx->n[sp], pass_perm(ctx, sp));
  if (rc) return rc;  // a failed launch leaves the permutation waiting
  ctx->cur[sp] = 1 - src;
  ctx->pending[sp] = false;
  return 0;
}

int settle_sorts(pic2d_ctx *ctx) {
  int rc = settle(ctx, 0);
  const int re = settle(ctx, 1);
  return rc ? rc : re;
}

// stable compaction of species sp into its other buffer; the slots lose
// their tile order
static int compact(pic2d_ctx *ctx, int sp) {
  const int64_t n = ctx->n[sp];
  const int f64 = ctx->f64 ? 1 : 0;
  const int src = ctx->cur[sp], dst = 1 - src;
  if (n > 0) {
    k_live_flags2<<<(unsigned)div_up(n, kBlock), kBlock, 0, ctx->rt.stream>>>(
        f64, p64(ctx, sp, src), p32(ctx, sp, src), n, ctx->rt.flags);
    P2_LAUNCH_CHECK(ctx);
  }
  int64_t live = 0;
  int rc = pic::select_flagged(&ctx->rt, ctx->rt.flags, ctx->rt.idx, n, &live);
  if (rc) return rc;
  rc = gather_perm(ctx, sp, src, dst, live, ctx->rt.idx);
  if (rc) return rc;
  ctx->cur[sp] = dst;
  ctx->n[sp] = live;
  return clear_tiles(ctx, sp);
}

int inflow(pic2d_ctx *ctx, int sp, uint32_t event) {
  {
    int rc = settle(ctx, sp);
    if (rc) return rc;
  }
  const pic2d_config &c = ctx->cfg;
  const int f64 = ctx->f64 ? 1 : 0;
  if (ctx->n[sp] > 0) {
    k_inflow_kill2<<<(unsigned)div_up(ctx->n[sp], kBlock), kBlock, 0,
                     ctx->rt.stream>>>(ctx->g, f64, p64(ctx, sp, ctx->cur[sp]),
                                       p32(ctx, sp, ctx->cur[sp]), ctx->n[sp],
                                       c.dx);
    P2_LAUNCH_CHECK(ctx);
  }
  const int64_t per_face = (int64_t)c.nx * c.ppc;
  if (ctx->n[sp] + 2 * per_face > ctx->cap) {
    // out of tail room: drop the dead slots first
    int rc = compact(ctx, sp);
    if (rc) return rc;
    if (ctx->n[sp] + 2 * per_face > ctx->cap)
      return pic::set_error(&ctx->rt, PIC_E_CAPACITY,
                            "inflow: %lld + %lld particles > capacity %lld",
                            (long long)ctx->n[sp], (long long)(2 * per_face),
                            (long long)ctx->cap);
  }
  const int64_t base = ctx->n[sp];
  if (per_face > 0) {
    k_refill2<<<(unsigned)div_up(2 * per_face, kBlock), kBlock, 0,
                ctx->rt.stream>>>(ctx->g, f64, p64(ctx, sp, ctx->cur[sp]),
                                  p32(ctx, sp, ctx->cur[sp]), base, per_face,
                                  c.ppc, c.dx, c.vth[sp], c.vdrift, c.seed, sp,
                                  event);
    P2_LAUNCH_CHECK(ctx);
  }
  ctx->n[sp] = base + 2 * per_face;
  return 0;
}

// ===========================================================================
// sort: key per slot (interior cells only; others dropped), stable radix
// sort of (key, slot), tile ranges from the boundaries of the sorted keys,
// gather

__global__ void k_keys2(Geom2 g, Tile2 tg, int f64, P64 P, P32 Q, int64_t n,
                        double dx, uint32_t *keys) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t ncell = (uint32_t)g.NX * g.NZ;
  uint32_t key = ncell;
  int k = -1, m = -1;
  if (f64) {
    const double x = P.x[i], z = P.z[i];
    if (!isnan(x) && in_array64(g, x, z, dx)) {
      k = (int)(x / dx) - 2;
      m = (int)(z / dx) - 2;
    }
  } else {
    const int32_t c = Q.cell[i];
    if (c >= 0) {
      k = c / g.SZ - 2;
      m = c % g.SZ - 2;
    }
  }
  if (k >= 0 && k < g.NX && m >= 0 && m < g.NZ) key = tg.key(k, m);
  keys[i] = key;
}

// The tile sort's destination map of the 2D form (pic_tsort.h): cell index
// k * SZ + m, key = tile * TX * TZ + (k % TX) * TZ + m % TZ over interior
// cells, home slot from the tile coordinates modulo 8 (the 9 tiles around
// any tile take distinct slots).
struct KeyMap2 {
  Geom2 g;
  Tile2 tg;
  __host__ __device__ uint32_t cells() const { return (uint32_t)(tg.TX * tg.TZ); }
  __device__ bool dest(int32_t c, uint32_t &t, uint32_t &cin, int &h) const {
    if (c < 0) return false;
    const int k = c / g.SZ - 2, m = c % g.SZ - 2;
    if (k < 0 || k >= g.NX || m < 0 || m >= g.NZ) return false;
    const int tx = k / tg.TX, tz = m / tg.TZ;
    t = (uint32_t)tx * tg.NTZ + (uint32_t)tz;
    cin = (uint32_t)(k - tx * tg.TX) * tg.TZ + (uint32_t)(m - tz * tg.TZ);
    h = ((tx & 7) << 3) | (tz & 7);
    return true;
  }
};

int tsort_setup(pic2d_ctx *ctx) {
  const int nt = ctx->tg.ntiles();
  // (slots are 32-bit in the tile sort: larger stores take the radix sort)
  if (ctx->f64 || nt <= 1 || ctx->tg.TX * ctx->tg.TZ > pic::tsort::kTileCells ||
      ctx->cap > ((int64_t)1 << 30))
    return 0;
  ctx->ts = new (std::nothrow) pic::tsort::Bufs();
  if (!ctx->ts) return err(ctx, PIC_E_NOMEM, "tile sort lists");
  P2_HIP(ctx, hipMalloc(&ctx->perm_ion, (size_t)ctx->cap * sizeof(uint32_t)));
  return pic::tsort::reserve(&ctx->rt, *ctx->ts, nt, ctx->cap);
}

void tsort_free(pic2d_ctx *ctx) {
  if (ctx->perm_ion) (void)hipFree(ctx->perm_ion);
  ctx->perm_ion = nullptr;
  if (!ctx->ts) return;
  pic::tsort::release(*ctx->ts);
  delete ctx->ts;
  ctx->ts = nullptr;
}

// defer: a tile sort leaves its permutation for the species' next fused
// pass instead of copying the particles (pic2d_step, FP32 form)
static int sort_impl(pic2d_ctx *ctx, int sp, bool defer) {
  {
    int rc = settle(ctx, sp);
    if (rc) return rc;
  }
  const int64_t n = ctx->n[sp];
  const int nt = ctx->tg.ntiles();
  const int f64 = ctx->f64 ? 1 : 0;
  if (n == 0) return clear_tiles(ctx, sp);
  const int src = ctx->cur[sp], dst = 1 - src;
  hipStream_t st = ctx->rt.stream;
  if (ctx->ts && ctx->sorted_end[sp] > 0 && ctx->sorted_end[sp] <= n) {
    // the tiles of the last sort are current: the tile sort
    const KeyMap2 km{ctx->g, ctx->tg};
    bool done = false;
    int64_t valid = 0;
    uint32_t *perm = defer ? pass_perm(ctx, sp) : ctx->rt.idx;
    int rc = pic::tsort::run(&ctx->rt, km, nt, ctx->tile_start[sp], ctx->sorted_end[sp], n,
                             p32(ctx, sp, src).cell, *ctx->ts, perm, &done, &valid);
    if (rc) return rc;
    if (done && defer && valid > 0) {
      ctx->n[sp] = valid;
      ctx->sorted_end[sp] = valid;
      ctx->sort_tile_count++;
      ctx->pending[sp] = true;
      return 0;
    }
    if (done) {
      rc = gather_perm(ctx, sp, src, dst, valid, perm);
      if (rc) return rc;
      ctx->cur[sp] = dst;
      ctx->n[sp] = valid;
      ctx->sorted_end[sp] = valid;
      ctx->sort_tile_count++;
      return 0;
    }
  }
  ctx->sort_radix_count++;
  k_keys2<<<(unsigned)div_up(n, kBlock), kBlock, 0, st>>>(
      ctx->g, ctx->tg, f64, p64(ctx, sp, src), p32(ctx, sp, src), n,
      ctx->cfg.dx, ctx->rt.keys);
  P2_LAUNCH_CHECK(ctx);
  const uint32_t ncell = (uint32_t)ctx->g.NX * ctx->g.NZ;
  int bits = 1;
  while (bits < 32 && ((uint64_t)1 << bits) <= ncell) ++bits;
  int rc = pic::sort_keys(&ctx->rt, ctx->rt.keys, ctx->rt.keys_alt, ctx->rt.idx,
                          n, bits);
  if (rc) return rc;
  const uint32_t cpt = (uint32_t)(ctx->tg.TX * ctx->tg.TZ);
  rc = pic::tile_bounds(&ctx->rt, ctx->rt.keys_alt, n, cpt, nt, ctx->tile_start[sp]);
  if (rc) return rc;
  P2_HIP(ctx, hipMemcpyAsync(ctx->rt.host_counter, ctx->tile_start[sp] + nt,
                             sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  P2_HIP(ctx, hipStreamSynchronize(st));
  const int64_t valid = (int64_t)(*(uint32_t *)ctx->rt.host_counter);
  rc = gather_perm(ctx, sp, src, dst, valid, ctx->rt.idx);
  if (rc) return rc;
  ctx->cur[sp] = dst;
  ctx->n[sp] = valid;
  ctx->sorted_end[sp] = valid;
  return 0;
}

// ===========================================================================
// counting, transfers, particle diagnostics

__global__ __launch_bounds__(256) void k_count2(int f64, P64 P, P32 Q, int64_t n,
                                                unsigned long long *count) {
  unsigned int t = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    t += f64 ? !isnan(P.x[i]) : Q.cell[i] >= 0;
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
  __shared__ unsigned int part[256 / 64];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s = 0;
    for (int w = 0; w < 256 / 64; ++w) s += part[w];
    if (s) atomicAdd(count, s);
  }
}

int sort(pic2d_ctx *ctx, int sp) { return sort_impl(ctx, sp, false); }

int sort_for_pass(pic2d_ctx *ctx, int sp) { return sort_impl(ctx, sp, true); }

int count_live(pic2d_ctx *ctx, int sp, int64_t *out) {
  {
    int rc = settle(ctx, sp);
    if (rc) return rc;
  }
  const int64_t n = ctx->n[sp];
  *out = 0;
  if (n == 0) return 0;
  unsigned long long *dc = (unsigned long long *)ctx->rt.dev_counter;
  P2_HIP(ctx, hipMemsetAsync(dc, 0, sizeof(u64), ctx->rt.stream));
  k_count2<<<(unsigned)std::min<int64_t>(2048, div_up(n, kBlock)), kBlock, 0,
             ctx->rt.stream>>>(ctx->f64 ? 1 : 0, p64(ctx, sp, ctx->cur[sp]),
                               p32(ctx, sp, ctx->cur[sp]), n, dc);
  P2_LAUNCH_CHECK(ctx);
  P2_HIP(ctx, hipMemcpyAsync(ctx->rt.host_counter, dc, sizeof(u64),
                             hipMemcpyDeviceToHost, ctx->rt.stream));
  P2_HIP(ctx, hipStreamSynchronize(ctx->rt.stream));
  *out = (int64_t)ctx->rt.host_counter[0];
  return 0;
}

// FP32 device slots from reference coordinates
__global__ void k_from64(Geom2 g, P32 Q, const double *__restrict__ x,
                         const double *__restrict__ z, const double *__restrict__ vx,
                         const double *__restrict__ vy, const double *__restrict__ vz,
                         int64_t n, double dx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xs = x[i] / dx, zs = z[i] / dx;
  const double fk = floor(xs), fm = floor(zs);
  const bool ok = xs == xs && zs == zs && fk >= 0 && fm >= 0 && fk < g.SX && fm < g.SZ;
  const int k = ok ? (int)fk : 0, m = ok ? (int)fm : 0;
  Q.cell[i] = ok ? k * g.SZ + m : -1;
  Q.ox[i] = unit_f32(xs - fk);
  Q.oz[i] = unit_f32(zs - fm);
  Q.vx[i] = (float)vx[i];
  Q.vy[i] = (float)vy[i];
  Q.vz[i] = (float)vz[i];
}

__global__ void k_to64(Geom2 g, P32 Q, double *x, double *z, double *vx,
                       double *vy, double *vz, int64_t n, double dx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t c = Q.cell[i];
  const int k = c / g.SZ, m = c % g.SZ;
  x[i] = ((double)k + (double)Q.ox[i]) * dx;
  z[i] = ((double)m + (double)Q.oz[i]) * dx;
  vx[i] = Q.vx[i];
  vy[i] = Q.vy[i];
  vz[i] = Q.vz[i];
}

int upload_particles(pic2d_ctx *ctx, int sp, int64_t n, const double *const *h) {
  {
    int rc = settle(ctx, sp);
    if (rc) return rc;
  }
  if (n > ctx->cap)
    return pic::set_error(&ctx->rt, PIC_E_CAPACITY,
                          "upload: %lld particles > capacity %lld", (long long)n,
                          (long long)ctx->cap);
  const int cur = ctx->cur[sp];
  if (ctx->f64) {
    const P64 P = p64(ctx, sp, cur);
    double *d[5] = {P.x, P.z, P.vx, P.vy, P.vz};
    for (int j = 0; j < 5; ++j)
      if (n > 0) P2_HIP(ctx, hipMemcpy(d[j], h[j], n * sizeof(double), hipMemcpyHostToDevice));
  } else if (n > 0) {
    // stage the doubles in the other buffer (5 x 8 bytes fit in its 6 x 4
    // bytes per slot times the capacity only up to cap*24/40 slots, so
    // allocate a staging block instead)
    double *stage = nullptr;
    P2_HIP(ctx, hipMalloc(&stage, (size_t)5 * n * sizeof(double)));
    hipError_t e = hipSuccess;
    for (int j = 0; j < 5 && e == hipSuccess; ++j)
      e = hipMemcpy(stage + j * n, h[j], n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      k_from64<<<(unsigned)div_up(n, kBlock), kBlock, 0, ctx->rt.stream>>>(
          ctx->g, p32(ctx, sp, cur), stage, stage + n, stage + 2 * n,
          stage + 3 * n, stage + 4 * n, n, ctx->cfg.dx);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->rt.stream);
    (void)hipFree(stage);
    P2_HIP(ctx, e);
  }
  ctx->n[sp] = n;
  return clear_tiles(ctx, sp);
}

int download_particles(pic2d_ctx *ctx, int sp, int64_t cap, int64_t *nout,
                       double *const *h) {
  {
    int rc = settle(ctx, sp);
    if (rc) return rc;
  }
  const int64_t n = ctx->n[sp];
  const int f64 = ctx->f64 ? 1 : 0;
  const int cur = ctx->cur[sp];
  hipStream_t st = ctx->rt.stream;
  if (n > 0) {
    k_live_flags2<<<(unsigned)div_up(n, kBlock), kBlock, 0, st>>>(
        f64, p64(ctx, sp, cur), p32(ctx, sp, cur), n, ctx->rt.flags);
    P2_LAUNCH_CHECK(ctx);
  }
  int64_t live = 0;
  int rc = pic::select_flagged(&ctx->rt, ctx->rt.flags, ctx->rt.idx, n, &live);
  if (rc) return rc;
  if (live > cap)
    return pic::set_error(&ctx->rt, PIC_E_CAPACITY, "download: %lld live > cap %lld",
                          (long long)live, (long long)cap);
  *nout = live;
  if (live == 0) return 0;
  // gather the live slots (stable) into the idle buffer, then copy out
  const int tmp = 1 - cur;
  k_gather2<<<(unsigned)div_up(live, kBlock), kBlock, 0, st>>>(
      f64, p64(ctx, sp, tmp), p32(ctx, sp, tmp), p64(ctx, sp, cur),
      p32(ctx, sp, cur), ctx->rt.idx, live);
  P2_LAUNCH_CHECK(ctx);
  if (f64) {
    const P64 P = p64(ctx, sp, tmp);
    const double *d[5] = {P.x, P.z, P.vx, P.vy, P.vz};
    for (int j = 0; j < 5; ++j)
      P2_HIP(ctx, hipMemcpyAsync(h[j], d[j], live * sizeof(double),
                                 hipMemcpyDeviceToHost, st));
    P2_HIP(ctx, hipStreamSynchronize(st));
    return 0;
  }
  double *stage = nullptr;
  P2_HIP(ctx, hipMalloc(&stage, (size_t)5 * live * sizeof(double)));
  k_to64<<<(unsigned)div_up(live, kBlock), kBlock, 0, st>>>(
      ctx->g, p32(ctx, sp, tmp), stage, stage + live, stage + 2 * live,
      stage + 3 * live, stage + 4 * live, live, ctx->cfg.dx);
  hipError_t e = hipGetLastError();
  for (int j = 0; j < 5 && e == hipSuccess; ++j)
    e = hipMemcpyAsync(h[j], stage + j * live, live * sizeof(double),
                       hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(stage);
  P2_HIP(ctx, e);
  return 0;
}

// per-block FP64 sums of v^2, vx, vy, vz over live particles
__global__ __launch_bounds__(256) void k_pdiag2(int f64, P64 P, P32 Q, int64_t n,
                                                double *__restrict__ part) {
  __shared__ double red[4][256];
  double v[4] = {0, 0, 0, 0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double a, b, c;
    if (f64) {
      if (isnan(P.x[i])) continue;
      a = P.vx[i];
      b = P.vy[i];
      c = P.vz[i];
    } else {
      if (Q.cell[i] < 0) continue;
      a = Q.vx[i];
      b = Q.vy[i];
      c = Q.vz[i];
    }
    v[0] += a * a + b * b + c * c;
    v[1] += a;
    v[2] += b;
    v[3] += c;
  }
  for (int c = 0; c < 4; ++c) red[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 4) part[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

int diagnose_particles(pic2d_ctx *ctx, pic2d_diag *d) {
  {
    int rc = settle_sorts(ctx);
    if (rc) return rc;
  }
  const pic2d_config &c = ctx->cfg;
  const int nbk = 512;
  double *part = nullptr;
  P2_HIP(ctx, hipMalloc(&part, (size_t)nbk * 4 * sizeof(double)));
  double h[nbk * 4];
  for (int sp = 0; sp < 2; ++sp) {
    const int64_t n = ctx->n[sp];
    double s[4] = {0, 0, 0, 0};
    if (n > 0) {
      k_pdiag2<<<nbk, 256, 0, ctx->rt.stream>>>(
          ctx->f64 ? 1 : 0, p64(ctx, sp, ctx->cur[sp]), p32(ctx, sp, ctx->cur[sp]),
          n, part);
      hipError_t e = hipGetLastError();
      if (e == hipSuccess)
        e = hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, ctx->rt.stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->rt.stream);
      if (e != hipSuccess) {
        (void)hipFree(part);
        P2_HIP(ctx, e);
      }
      for (int b = 0; b < nbk; ++b)
        for (int j = 0; j < 4; ++j) s[j] += h[b * 4 + j];
    }
    const double mw = c.mass[sp] * c.weight[sp] * c.dx * c.dx;
    d->kinetic[sp] = 0.5 * mw * s[0];
    d->momentum[sp][0] = mw * s[1];
    d->momentum[sp][1] = mw * s[2];
    d->momentum[sp][2] = mw * s[3];
    int64_t live = 0;
    int rc = count_live(ctx, sp, &live);
    if (rc) {
      (void)hipFree(part);
      return rc;
    }
    d->count[sp] = live;
  }
  (void)hipFree(part);
  return 0;
}

}  // namespace pic2d

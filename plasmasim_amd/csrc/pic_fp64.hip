// Plain FP64 form of the step on gfx950.
//
// Arithmetic mirrors ashida-lab/PLASMAsim PIC/3d_pic458_domain.c expression
// by expression; this translation unit is compiled with -ffp-contract=off so
// that no multiply-add is fused and every operation rounds as the
// reference's C does.  Layout: one plane per field component (m contiguous),
// particles in SoA per species; a dead particle carries z = NaN until the
// next compaction (inflow or sort).
#include <hip/hip_runtime.h>
#include <new>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "philox.h"
#include "pic_internal.h"
#include "pic_tsort.h"
#include "pic_tiles.h"

namespace pic {
namespace f64 {

// ---------------------------------------------------------------------------
// field kernels

__global__ void k_zero(double *__restrict__ p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0.0;
}

__device__ inline double sigma_at(const SigmaProfile &s, int k2, int l2,
                                  int m2) {
  const double a4 = s.d4[0][k2] + s.d4[1][l2] + s.d4[2][m2];
  if (a4 == 0.0) return 0.0;  // outside every layer: all depths are zero
  const double a12 = s.d12[0][k2] + s.d12[1][l2] + s.d12[2][m2];
  return s.amp * (0.1 * sqrt(a4) + 0.01 * sqrt(a12));
}

// renew_b4_2 (:2580-2611)
__global__ __launch_bounds__(256) void k_renew_b(Geom g, Fields<double> F,
                                                 SigmaProfile sg, double dt,
                                                 double dx, double eps0) {
  const int m = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int l = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 2 + blockIdx.z;
  if (m >= g.NZ + 2 || l >= g.NY + 2) return;
  const double temp = 1. / eps0 * dt * 0.5 * 0.5;
  const double dth = 0.5 * dt;
  const double temp_24 = 1. / 24.;
  const double inv_dx = 1. / dx;
  const int64_t c = g.idx(k, l, m);
  const int64_t dl = g.SZ, dk = (int64_t)g.SY * g.SZ;
  const double *ex = F.f[PIC_EX], *ey = F.f[PIC_EY], *ez = F.f[PIC_EZ];
  double f, h, s;

  s = sigma_at(sg, 2 * k + 1, 2 * l, 2 * m);
  f = s == 0.0 ? 1.0 : (1. - s * temp) / (1. + s * temp);
  h = s == 0.0 ? dth : dth / (1. + s * temp);
  {
    const double b0 = F.f[PIC_B0X][c];
    F.f[PIC_BX][c] =
        b0 + f * (F.f[PIC_BX][c] - b0) -
        h * temp_24 *
            (27. * (ez[c] - ez[c - dl] - ey[c] + ey[c - 1]) -
             (ez[c + dl] - ez[c - 2 * dl] - ey[c + 1] + ey[c - 2])) *
            inv_dx;
  }
  s = sigma_at(sg, 2 * k, 2 * l + 1, 2 * m);
  f = s == 0.0 ? 1.0 : (1. - s * temp) / (1. + s * temp);
  h = s == 0.0 ? dth : dth / (1. + s * temp);
  {
    const double b0 = F.f[PIC_B0Y][c];
    F.f[PIC_BY][c] =
        b0 + f * (F.f[PIC_BY][c] - b0) -
        h * temp_24 *
            (27. * (ex[c] - ex[c - 1] - ez[c] + ez[c - dk]) -
             (ex[c + 1] - ex[c - 2] - ez[c + dk] + ez[c - 2 * dk])) *
            inv_dx;
  }
  s = sigma_at(sg, 2 * k, 2 * l, 2 * m + 1);
  f = s == 0.0 ? 1.0 : (1. - s * temp) / (1. + s * temp);
  h = s == 0.0 ? dth : dth / (1. + s * temp);
  {
    const double b0 = F.f[PIC_B0Z][c];
    F.f[PIC_BZ][c] =
        b0 + f * (F.f[PIC_BZ][c] - b0) -
        h * temp_24 *
            (27. * (ey[c] - ey[c - dk] - ex[c] + ex[c - dl]) -
             (ey[c + dk] - ey[c - 2 * dk] - ex[c + dl] + ex[c - 2 * dl])) *
            inv_dx;
  }
}

// renew_e4_2 (:2613-2652); the curl acts on B - B0 recomputed per term, as
// the reference writes it.
__global__ __launch_bounds__(256) void k_renew_e(Geom g, Fields<double> F,
                                                 SigmaProfile sg, double dt,
                                                 double dx, double c_light,
                                                 double mu0) {
  const int m = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int l = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 2 + blockIdx.z;
  if (m >= g.NZ + 2 || l >= g.NY + 2) return;
  const double temp = c_light * c_light * dt;
  const double temp_24 = 1. / 24.;
  const double inv_dx = 1. / dx;
  const int64_t c = g.idx(k, l, m);
  const int64_t dl = g.SZ, dk = (int64_t)g.SY * g.SZ;
  const double *bx = F.f[PIC_BX], *by = F.f[PIC_BY], *bz = F.f[PIC_BZ];
  const double *b0x = F.f[PIC_B0X], *b0y = F.f[PIC_B0Y], *b0z = F.f[PIC_B0Z];
  double s, f, h;
#define D1(B, B0, a, b) (B[a] - B0[a] - B[b] + B0[b])

  s = sigma_at(sg, 2 * k, 2 * l + 1, 2 * m + 1);
  f = s == 0.0 ? 1.0 : (1. - 0.5 * mu0 * s * temp) / (1. + 0.5 * mu0 * s * temp);
  h = s == 0.0 ? temp : temp / (1. + 0.5 * mu0 * s * temp);
  F.f[PIC_EX][c] =
      f * F.f[PIC_EX][c] +
      h * (temp_24 *
               (27. * (D1(bz, b0z, c + dl, c) - D1(by, b0y, c + 1, c)) -
                (D1(bz, b0z, c + 2 * dl, c - dl) - D1(by, b0y, c + 2, c - 1))) *
               inv_dx -
           mu0 * (F.f[PIC_JIX][c] - F.f[PIC_JEX][c]));

  s = sigma_at(sg, 2 * k + 1, 2 * l, 2 * m + 1);
  f = s == 0.0 ? 1.0 : (1. - 0.5 * mu0 * s * temp) / (1. + 0.5 * mu0 * s * temp);
  h = s == 0.0 ? temp : temp / (1. + 0.5 * mu0 * s * temp);
  F.f[PIC_EY][c] =
      f * F.f[PIC_EY][c] +
      h * (temp_24 *
               (27. * (D1(bx, b0x, c + 1, c) - D1(bz, b0z, c + dk, c)) -
                (D1(bx, b0x, c + 2, c - 1) - D1(bz, b0z, c + 2 * dk, c - dk))) *
               inv_dx -
           mu0 * (F.f[PIC_JIY][c] - F.f[PIC_JEY][c]));

  s = sigma_at(sg, 2 * k + 1, 2 * l + 1, 2 * m);
  f = s == 0.0 ? 1.0 : (1. - 0.5 * mu0 * s * temp) / (1. + 0.5 * mu0 * s * temp);
  h = s == 0.0 ? temp : temp / (1. + 0.5 * mu0 * s * temp);
  F.f[PIC_EZ][c] =
      f * F.f[PIC_EZ][c] +
      h * (temp_24 *
               (27. * (D1(by, b0y, c + dk, c) - D1(bx, b0x, c + dl, c)) -
                (D1(by, b0y, c + 2 * dk, c - dk) - D1(bx, b0x, c + 2 * dl, c - dl))) *
               inv_dx -
           mu0 * (F.f[PIC_JIZ][c] - F.f[PIC_JEZ][c]));
#undef D1
}

// sr_b / sr_e on one rank (:1001-1387): copy along x over all (l,m), then
// along y over all (k,m) in a second launch so that corners see the x pass.
__global__ void k_halo_x(Geom g, double *a, double *b, double *c) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)g.SY * g.SZ;
  if (i >= n) return;
  const int64_t dk = n;
  double *p[3] = {a, b, c};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    double *f = p[q] + i;
    f[0] = f[g.NX * dk];
    f[1 * dk] = f[(g.NX + 1) * dk];
    f[(g.NX + 2) * dk] = f[2 * dk];
    f[(g.NX + 3) * dk] = f[3 * dk];
  }
}

__global__ void k_halo_y(Geom g, double *a, double *b, double *c) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)g.SX * g.SZ;
  if (i >= n) return;
  const int k = (int)(i / g.SZ), m = (int)(i % g.SZ);
  const int64_t dl = g.SZ;
  double *p[3] = {a, b, c};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    double *f = p[q] + g.idx(k, 0, m);
    f[0] = f[g.NY * dl];
    f[1 * dl] = f[(g.NY + 1) * dl];
    f[(g.NY + 2) * dl] = f[2 * dl];
    f[(g.NY + 3) * dl] = f[3 * dl];
  }
}

// sr_current2 (:1805-2039) on one rank: 2,3 += N+2,N+3 then N,N+1 += 0,1
// along x over all (l,m); then along y over all (k,m).  8 planes.
__global__ void k_fold_x(Geom g, Fields<double> F) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)g.SY * g.SZ;
  if (i >= n) return;
  const int64_t dk = n;
#pragma unroll
  for (int q = PIC_JIX; q <= PIC_NE; ++q) {
    double *f = F.f[q] + i;
    f[2 * dk] += f[(g.NX + 2) * dk];
    f[3 * dk] += f[(g.NX + 3) * dk];
    f[g.NX * dk] += f[0];
    f[(g.NX + 1) * dk] += f[1 * dk];
  }
}

__global__ void k_fold_y(Geom g, Fields<double> F) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)g.SX * g.SZ;
  if (i >= n) return;
  const int k = (int)(i / g.SZ), m = (int)(i % g.SZ);
  const int64_t dl = g.SZ;
#pragma unroll
  for (int q = PIC_JIX; q <= PIC_NE; ++q) {
    double *f = F.f[q] + g.idx(k, 0, m);
    f[2 * dl] += f[(g.NY + 2) * dl];
    f[3 * dl] += f[(g.NY + 3) * dl];
    f[g.NY * dl] += f[0];
    f[(g.NY + 1) * dl] += f[1 * dl];
  }
}

// ---------------------------------------------------------------------------
// particle kernels

// Gather of one staggered component (shape_func_v2 :3223-3273): the eight
// products are formed and summed in the reference's order.
__device__ inline double gather1(const double *__restrict__ F, const Geom &g,
                                 int a, int b, int c, double wa0, double wa1,
                                 double wb0, double wb1, double wc0,
                                 double wc1) {
  const int64_t i = g.idx(a, b, c);
  const int64_t dl = g.SZ, dk = (int64_t)g.SY * g.SZ;
  const double p1 = F[i] * (wa0 * wb0 * wc0) + F[i + 1] * (wa0 * wb0 * wc1);
  const double p2 =
      F[i + dl] * (wa0 * wb1 * wc0) + F[i + dl + 1] * (wa0 * wb1 * wc1);
  const double p3 =
      F[i + dk] * (wa1 * wb0 * wc0) + F[i + dk + 1] * (wa1 * wb0 * wc1);
  const double p4 = F[i + dk + dl] * (wa1 * wb1 * wc0) +
                    F[i + dk + dl + 1] * (wa1 * wb1 * wc1);
  return p1 + p2 + p3 + p4;
}

struct AxisW {
  int bi, bh;        // (int)(x/dx), (int)(x/dx - 0.5)
  double iu, il;     // integer-located weights: upper, lower
  double hu, hl;     // half-located weights
};

__device__ inline AxisW axis_weights(double x, double dx) {
  AxisW w;
  w.bi = (int)(x / dx);
  w.iu = ((w.bi + 1) * dx - x);
  w.il = (x - w.bi * dx);
  w.bh = (int)(x / dx - 0.5);
  w.hu = (((w.bh + 1) + 0.5) * dx - x);
  w.hl = (x - (w.bh + 0.5) * dx);
  return w;
}

// shape_func_v2 (:3017-3326): out = ex, ey, ez, bx, by, bz
__device__ inline void gather(const Fields<double> &F, const Geom &g, double x,
                              double y, double z, double dx, double out[6]) {
  const double inv_dx3 = 1. / (dx * dx * dx);
  const AxisW X = axis_weights(x, dx), Y = axis_weights(y, dx),
              Z = axis_weights(z, dx);
  out[0] = gather1(F.f[PIC_EX], g, X.bi, Y.bh, Z.bh, X.iu, X.il, Y.hu, Y.hl,
                   Z.hu, Z.hl) * inv_dx3;
  out[1] = gather1(F.f[PIC_EY], g, X.bh, Y.bi, Z.bh, X.hu, X.hl, Y.iu, Y.il,
                   Z.hu, Z.hl) * inv_dx3;
  out[2] = gather1(F.f[PIC_EZ], g, X.bh, Y.bh, Z.bi, X.hu, X.hl, Y.hu, Y.hl,
                   Z.iu, Z.il) * inv_dx3;
  out[3] = gather1(F.f[PIC_BX], g, X.bh, Y.bi, Z.bi, X.hu, X.hl, Y.iu, Y.il,
                   Z.iu, Z.il) * inv_dx3;
  out[4] = gather1(F.f[PIC_BY], g, X.bi, Y.bh, Z.bi, X.iu, X.il, Y.hu, Y.hl,
                   Z.iu, Z.il) * inv_dx3;
  out[5] = gather1(F.f[PIC_BZ], g, X.bi, Y.bi, Z.bh, X.iu, X.il, Y.iu, Y.il,
                   Z.hu, Z.hl) * inv_dx3;
}

// cal_track4 (:2804-2870)
__device__ inline void boris(const StepConst &k, const double f[6], double &x,
                             double &y, double &z, double &vx, double &vy,
                             double &vz) {
  const double C = k.c, dt = k.dt;
  const double T = 0.5 * k.q * dt / k.mass;
  double ex = f[0], ey = f[1], ez = f[2];
  const double bx = f[3], by = f[4], bz = f[5];
  ex += k.vdrift_z * k.imf_y;
  ey += -k.vdrift_z * k.imf_x;
  double gamma1 = C / sqrt(C * C - vx * vx - vy * vy - vz * vz);
  if (isnan(gamma1)) gamma1 = 0.;
  double ux = gamma1 * vx, uy = gamma1 * vy, uz = gamma1 * vz;
  double gamma2 = C / sqrt(C * C + ux * ux + uy * uy + uz * uz);
  const double bux = gamma2 * (bx), buy = gamma2 * (by), buz = gamma2 * (bz);
  const double S = 2. * T / (1. + T * T * (bux * bux + buy * buy + buz * buz));
  ux += (ex) * T;
  uy += (ey) * T;
  uz += (ez) * T;
  const double uox = ux + T * (uy * buz - uz * buy);
  const double uoy = uy + T * (uz * bux - ux * buz);
  const double uoz = uz + T * (ux * buy - uy * bux);
  double upx = ux + S * (uoy * buz - uoz * buy);
  double upy = uy + S * (uoz * bux - uox * buz);
  double upz = uz + S * (uox * buy - uoy * bux);
  upx += (ex) * T;
  upy += (ey) * T;
  upz += (ez) * T;
  gamma2 = C / sqrt(C * C + upx * upx + upy * upy + upz * upz);
  vx = gamma2 * upx;
  vy = gamma2 * upy;
  vz = gamma2 * upz;
  x += vx * dt * 0.5;
  y += vy * dt * 0.5;
  z += vz * dt * 0.5;
}

__device__ inline void add(double *p, double v) { unsafeAtomicAdd(p, v); }

// shape_func_{ion,electron}0_j_2 (:3397-3553)
__device__ inline void deposit_j(const Geom &g, double *jx, double *jy,
                                 double *jz, double x, double y, double z,
                                 double vx, double vy, double vz, double wgt,
                                 double q, double dx) {
  const double dx3 = dx * dx * dx;
  const int64_t dl = g.SZ, dk = (int64_t)g.SY * g.SZ;
  int k = (int)(x / dx);
  int l = (int)(y / dx - 0.5);
  int m = (int)(z / dx - 0.5);
  const double qn = q * wgt / dx3;
  double temp = qn * vx;
  double kkkp = ((k + 1) * dx - x);
  double lllp = (((l + 1) + 0.5) * dx - y);
  double mmmp = (((m + 1) + 0.5) * dx - z);
  double pk = (x - k * dx);
  double pl = (y - (l + 0.5) * dx);
  double pm = (z - (m + 0.5) * dx);
  double tk = temp * kkkp, tp = temp * pk;
  int64_t i = g.idx(k, l, m);
  add(jx + i, (tk) * (lllp * mmmp));
  add(jx + i + 1, (tk) * (lllp * pm));
  add(jx + i + dl, (tk) * (pl * mmmp));
  add(jx + i + dl + 1, (tk) * (pl * pm));
  add(jx + i + dk, (tp) * (lllp * mmmp));
  add(jx + i + dk + 1, (tp) * (lllp * pm));
  add(jx + i + dk + dl, (tp) * (pl * mmmp));
  add(jx + i + dk + dl + 1, (tp) * (pl * pm));

  k = (int)(x / dx - 0.5);
  l = (int)(y / dx);
  temp = qn * vy;
  kkkp = (((k + 1) + 0.5) * dx - x);
  lllp = ((l + 1) * dx - y);
  pk = (x - (k + 0.5) * dx);
  pl = (y - l * dx);
  tk = temp * kkkp;
  tp = temp * pk;
  i = g.idx(k, l, m);
  add(jy + i, (tk) * (lllp * mmmp));
  add(jy + i + 1, (tk) * (lllp * pm));
  add(jy + i + dl, (tk) * (pl * mmmp));
  add(jy + i + dl + 1, (tk) * (pl * pm));
  add(jy + i + dk, (tp) * (lllp * mmmp));
  add(jy + i + dk + 1, (tp) * (lllp * pm));
  add(jy + i + dk + dl, (tp) * (pl * mmmp));
  add(jy + i + dk + dl + 1, (tp) * (pl * pm));

  l = (int)(y / dx - 0.5);
  m = (int)(z / dx);
  temp = qn * vz;
  lllp = (((l + 1) + 0.5) * dx - y);
  mmmp = ((m + 1) * dx - z);
  pl = (y - (l + 0.5) * dx);
  pm = (z - m * dx);
  tk = temp * kkkp;
  tp = temp * pk;
  i = g.idx(k, l, m);
  add(jz + i, (tk) * (lllp * mmmp));
  add(jz + i + 1, (tk) * (lllp * pm));
  add(jz + i + dl, (tk) * (pl * mmmp));
  add(jz + i + dl + 1, (tk) * (pl * pm));
  add(jz + i + dk, (tp) * (lllp * mmmp));
  add(jz + i + dk + 1, (tp) * (lllp * pm));
  add(jz + i + dk + dl, (tp) * (pl * mmmp));
  add(jz + i + dk + dl + 1, (tp) * (pl * pm));
}

// shape_func_{ion,electron}0_n_2 (:3333-3395)
__device__ inline void deposit_n(const Geom &g, double *n, double x, double y,
                                 double z, double wgt, double dx) {
  const double dx3 = dx * dx * dx;
  const int64_t dl = g.SZ, dk = (int64_t)g.SY * g.SZ;
  const int k = (int)(x / dx - 0.5);
  const int l = (int)(y / dx - 0.5);
  const int m = (int)(z / dx - 0.5);
  const double temp = wgt / dx3;
  const double kkkp = (((k + 1) + 0.5) * dx - x);
  const double lllp = (((l + 1) + 0.5) * dx - y);
  const double mmmp = (((m + 1) + 0.5) * dx - z);
  const double pk = (x - (k + 0.5) * dx);
  const double pl = (y - (l + 0.5) * dx);
  const double pm = (z - (m + 0.5) * dx);
  const double tk = temp * kkkp, tp = temp * pk;
  const int64_t i = g.idx(k, l, m);
  add(n + i, (tk) * (lllp * mmmp));
  add(n + i + 1, (tk) * (lllp * pm));
  add(n + i + dl, (tk) * (pl * mmmp));
  add(n + i + dl + 1, (tk) * (pl * pm));
  add(n + i + dk, (tp) * (lllp * mmmp));
  add(n + i + dk + 1, (tp) * (lllp * pm));
  add(n + i + dk + dl, (tp) * (pl * mmmp));
  add(n + i + dk + dl + 1, (tp) * (pl * pm));
}

// Array-safety rule (the oracle states the same): nodes base..base+1 stay
// inside the array when -0.5 < x/dx < S-1 on every axis.
__device__ inline bool in_array(const Geom &g, double x, double y, double z,
                                double dx) {
  const double xd = x / dx, yd = y / dx, zd = z / dx;
  return xd > -0.5 && xd < g.SX - 1 && yd > -0.5 && yd < g.SY - 1 &&
         zd > -0.5 && zd < g.SZ - 1;
}

// deposit gate (:747-749)
__device__ inline bool in_gate(const Geom &g, double x, double y, double z,
                               double dx) {
  return x < (g.NX + 4) * dx && x >= 0. && y < (g.NY + 4) * dx && y >= 0. &&
         z < (g.NZ + 4) * dx && z >= 0.;
}

// One particle pass.  PUSH_J: hot loop A (:731-812); DRIFT_N: hot loop B
// (:818-867); BOUNDARY: sr_particle (:2041-2531) on one rank.  The fused
// form (all three) is exact because none of the field updates between those
// loops in fdtd() reads or writes particle state.
template <bool PUSH_J, bool DRIFT_N, bool BOUNDARY>
__global__ __launch_bounds__(256) void k_particles(
    Geom g, Fields<double> F, Particles64 P, int64_t n, StepConst kc,
    double *jx, double *jy, double *jz, double *nd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double z = P.z[i];
  if (isnan(z)) return;  // dead until the next compaction
  double x = P.x[i], y = P.y[i];
  double vx = P.vx[i], vy = P.vy[i], vz = P.vz[i];
  const double dx = kc.dx;
  const double dth = kc.dt * 0.5;
  if (PUSH_J) {
    if (x < kc.blo[0] || x > kc.bhi[0] || y < kc.blo[1] || y > kc.bhi[1] ||
        z < kc.blo[2] || z > kc.bhi[2] || !in_array(g, x, y, z, dx)) {
      x += vx * dth;
      y += vy * dth;
      z += vz * dth;
    } else {
      double f[6];
      gather(F, g, x, y, z, dx, f);
      boris(kc, f, x, y, z, vx, vy, vz);
    }
    if (in_gate(g, x, y, z, dx) && in_array(g, x, y, z, dx))
      deposit_j(g, jx, jy, jz, x, y, z, vx, vy, vz, kc.weight, kc.q, dx);
  }
  if (DRIFT_N) {
    x += vx * dth;
    y += vy * dth;
    z += vz * dth;
    if (in_gate(g, x, y, z, dx) && in_array(g, x, y, z, dx))
      deposit_n(g, nd, x, y, z, kc.weight, dx);
  }
  if (BOUNDARY) {
    const double dx2 = 2. * dx;
    const double gx2 = (g.NX + 2) * dx, gy2 = (g.NY + 2) * dx,
                 gz2 = (g.NZ + 2) * dx;
    bool dead = !(z < gz2 && z >= dx2);
    if (!dead) {
      if (x >= gx2) {
        if (kc.periodic[0]) x = x - g.NX * dx; else dead = true;
      } else if (x < dx2) {
        if (kc.periodic[0]) x = x + g.NX * dx; else dead = true;
      }
      if (y >= gy2) {
        if (kc.periodic[1]) y = y - g.NY * dx; else dead = true;
      } else if (y < dx2) {
        if (kc.periodic[1]) y = y + g.NY * dx; else dead = true;
      }
    }
    if (dead) {
      P.z[i] = __longlong_as_double(0x7ff8000000000000LL);
      return;
    }
  }
  P.x[i] = x;
  P.y[i] = y;
  P.z[i] = z;
  P.vx[i] = vx;
  P.vy[i] = vy;
  P.vz[i] = vz;
}

// ---------------------------------------------------------------------------
// tiled particle pass: one workgroup per sort tile, J and n accumulated in an
// LDS window with FP64 atomics, then added to the planes (see pic_tiles.h)

constexpr int kWin64 = 1400;  // window capacity in nodes (4 x double each)

#define PIC_LDS __attribute__((address_space(3)))
#define PIC_GLB __attribute__((address_space(1)))

__device__ inline void lds_add64(double *p, double v) {
  __hip_atomic_fetch_add((PIC_LDS double *)p, v, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void glob_add64(double *p, double v) {
  __hip_atomic_fetch_add((PIC_GLB double *)p, v, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}

// The 8 contributions (temp*wa)*(wb*wc) of one component, as the reference
// forms them (:3412-3425); lane j visits the nodes in the order n ^ (j & 7)
// so that lanes holding particles of one cell hit different LDS nodes.
__device__ inline void dep8_64(double *lds, double *glob, bool in_lds,
                               const Win &w, const Geom &g, int a, int b,
                               int c, double wa0, double wa1, double wb0,
                               double wb1, double wc0, double wc1,
                               double temp) {
  const double ta0 = temp * wa0, ta1 = temp * wa1;
  if (in_lds) {
    const int WZ = w.WZ, WYZ = w.WY * w.WZ;
    const int i = ((a - w.k0) * w.WY + (b - w.l0)) * WZ + (c - w.m0);
    const int sw = threadIdx.x & 7;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      const int id = n ^ sw;
      const double ta = (id & 4) ? ta1 : ta0;
      const double wb = (id & 2) ? wb1 : wb0;
      const double wc = (id & 1) ? wc1 : wc0;
      const int off = ((id & 4) ? WYZ : 0) + ((id & 2) ? WZ : 0) + (id & 1);
      lds_add64(lds + i + off, (ta) * (wb * wc));
    }
  } else {
    const int64_t dl = g.SZ, dk = (int64_t)g.SY * g.SZ;
    const int64_t i = g.idx(a, b, c);
    glob_add64(glob + i, (ta0) * (wb0 * wc0));
    glob_add64(glob + i + 1, (ta0) * (wb0 * wc1));
    glob_add64(glob + i + dl, (ta0) * (wb1 * wc0));
    glob_add64(glob + i + dl + 1, (ta0) * (wb1 * wc1));
    glob_add64(glob + i + dk, (ta1) * (wb0 * wc0));
    glob_add64(glob + i + dk + 1, (ta1) * (wb0 * wc1));
    glob_add64(glob + i + dk + dl, (ta1) * (wb1 * wc0));
    glob_add64(glob + i + dk + dl + 1, (ta1) * (wb1 * wc1));
  }
}

// thrust_fave partials of one workgroup (six per-thread sums added over the
// workgroup in a fixed order), stored at slot bid
__device__ inline void thrust_store64(const double (&th)[6], double *out, int bid) {
  __shared__ double sh[6][kBlock / 64];
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
    for (int j = 1; j < kBlock / 64; ++j) a += sh[threadIdx.x][j];
    out[(int64_t)bid * 6 + threadIdx.x] = a;
  }
}

// THRUST: also add thrust_fave (:5124-5285) of the previous step for the
// particles of the central box: their positions and velocities at the start
// of the pass are that step's end state, E is unchanged since, and Fp holds
// the total B of its end (E planes shared with F).
template <bool PUSH_J, bool DRIFT_N, bool BOUNDARY, bool THRUST>
__global__ __launch_bounds__(256) void k_particles_tiled(
    Geom g, TileGeom tg, Fields<double> F, Particles64 P, TileRun tr,
    StepConst kc, double *jx, double *jy, double *jz, double *nd,
    Fields<double> Fp, double *thrust_part) {
  __shared__ double acc[4 * kWin64];
  __shared__ Win sw;
  __shared__ int64_t s_range[2];
  const int tid = threadIdx.x;
  const int bid = xcd_remap(blockIdx.x, tr.ntiles, (int)gridDim.x);
  const double dx = kc.dx;
  if (tid == 0) {
    int64_t s, e;
    Win w;
    tile_setup(bid, tr, tg, kWin64,
               [&](int64_t i, int &k, int &l, int &m) {
                 const double x = P.x[i], y = P.y[i], z = P.z[i];
                 if (!in_array(g, x, y, z, dx)) return false;
                 k = (int)(x / dx);
                 l = (int)(y / dx);
                 m = (int)(z / dx);
                 return true;
               },
               s, e, w);
    if (!(PUSH_J || DRIFT_N)) w.on = false;
    sw = w;
    s_range[0] = s;
    s_range[1] = e;
  }
  __syncthreads();
  const int64_t s = s_range[0], e = s_range[1];
  double th[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (s >= e) {
    if (THRUST && tid < 6) thrust_part[(int64_t)bid * 6 + tid] = 0.0;
    return;
  }
  const Win w = sw;
  const int WN = w.on ? w.WX * w.WY * w.WZ : 0;
  for (int i = tid; i < 4 * WN; i += blockDim.x) acc[(i / WN) * kWin64 + i % WN] = 0.0;
  __syncthreads();
  const double dth = kc.dt * 0.5;
  for (int64_t i = s + tid; i < e; i += blockDim.x) {
    double z = P.z[i];
    if (isnan(z)) continue;
    double x = P.x[i], y = P.y[i];
    double vx = P.vx[i], vy = P.vy[i], vz = P.vz[i];
    const bool safe0 = in_array(g, x, y, z, dx);
    const bool lds_ok =
        safe0 && in_win(w, (int)(x / dx), (int)(y / dx), (int)(z / dx));
    if (THRUST && safe0 && x < g.NX * dx * 0.8 && x > g.NX * dx * 0.2 &&
        y < g.NY * dx * 0.8 && y > g.NY * dx * 0.2 && z < g.NZ * dx * 0.8 &&
        z > g.NZ * dx * 0.2) {
      double f[6];
      gather(Fp, g, x, y, z, dx, f);
      th[0] += f[0];
      th[1] += f[1];
      th[2] += f[2];
      th[3] += vy * f[5] - vz * f[4];
      th[4] += vz * f[3] - vx * f[5];
      th[5] += vx * f[4] - vy * f[3];
    }
    if (PUSH_J) {
      if (x < kc.blo[0] || x > kc.bhi[0] || y < kc.blo[1] || y > kc.bhi[1] ||
          z < kc.blo[2] || z > kc.bhi[2] || !safe0) {
        x += vx * dth;
        y += vy * dth;
        z += vz * dth;
      } else {
        double f[6];
        gather(F, g, x, y, z, dx, f);
        boris(kc, f, x, y, z, vx, vy, vz);
      }
      if (in_gate(g, x, y, z, dx) && in_array(g, x, y, z, dx)) {
        // shape_func_*0_j_2 (:3397-3553) with the reference's weights
        const double qn = kc.q * kc.weight / (dx * dx * dx);
        const AxisW X = axis_weights(x, dx), Y = axis_weights(y, dx),
                    Z = axis_weights(z, dx);
        dep8_64(acc + 0 * kWin64, jx, lds_ok, w, g, X.bi, Y.bh, Z.bh, X.iu,
                X.il, Y.hu, Y.hl, Z.hu, Z.hl, qn * vx);
        dep8_64(acc + 1 * kWin64, jy, lds_ok, w, g, X.bh, Y.bi, Z.bh, X.hu,
                X.hl, Y.iu, Y.il, Z.hu, Z.hl, qn * vy);
        dep8_64(acc + 2 * kWin64, jz, lds_ok, w, g, X.bh, Y.bh, Z.bi, X.hu,
                X.hl, Y.hu, Y.hl, Z.iu, Z.il, qn * vz);
      }
    }
    if (DRIFT_N) {
      x += vx * dth;
      y += vy * dth;
      z += vz * dth;
      if (in_gate(g, x, y, z, dx) && in_array(g, x, y, z, dx)) {
        // shape_func_*0_n_2 (:3333-3395)
        const AxisW X = axis_weights(x, dx), Y = axis_weights(y, dx),
                    Z = axis_weights(z, dx);
        dep8_64(acc + 3 * kWin64, nd, lds_ok, w, g, X.bh, Y.bh, Z.bh, X.hu,
                X.hl, Y.hu, Y.hl, Z.hu, Z.hl, kc.weight / (dx * dx * dx));
      }
    }
    if (BOUNDARY) {
      const double dx2 = 2. * dx;
      const double gx2 = (g.NX + 2) * dx, gy2 = (g.NY + 2) * dx,
                   gz2 = (g.NZ + 2) * dx;
      bool dead = !(z < gz2 && z >= dx2);
      if (!dead) {
        if (x >= gx2) {
          if (kc.periodic[0]) x = x - g.NX * dx; else dead = true;
        } else if (x < dx2) {
          if (kc.periodic[0]) x = x + g.NX * dx; else dead = true;
        }
        if (y >= gy2) {
          if (kc.periodic[1]) y = y - g.NY * dx; else dead = true;
        } else if (y < dx2) {
          if (kc.periodic[1]) y = y + g.NY * dx; else dead = true;
        }
      }
      if (dead) {
        P.z[i] = __longlong_as_double(0x7ff8000000000000LL);
        continue;
      }
    }
    P.x[i] = x;
    P.y[i] = y;
    P.z[i] = z;
    P.vx[i] = vx;
    P.vy[i] = vy;
    P.vz[i] = vz;
  }
  if (THRUST) thrust_store64(th, thrust_part, bid);
  if (!w.on) return;
  __syncthreads();
  const int WYZ = w.WY * w.WZ;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    double *dst = c == 0 ? jx : c == 1 ? jy : c == 2 ? jz : nd;
    for (int j = tid; j < WN; j += blockDim.x) {
      const double v = acc[c * kWin64 + j];
      if (v == 0.0) continue;
      const int kk = w.k0 + j / WYZ, ll = w.l0 + (j / w.WZ) % w.WY,
                mm = w.m0 + j % w.WZ;
      if (kk < 0 || ll < 0 || mm < 0 || kk >= g.SX || ll >= g.SY || mm >= g.SZ)
        continue;
      glob_add64(dst + g.idx(kk, ll, mm), v);
    }
  }
}

// init_particle_flat (:412-625): particle p sits in interior cell p/ppc,
// cells in (k,l,m) order with m fastest; Philox event 0.
__global__ void k_load(Geom g, Particles64 P, int64_t n, int ppc, double dx,
                       double vth, double mu0, double mu1, double mu2,
                       uint64_t seed, int sp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t cell = i / ppc;
  const int m = (int)(cell % g.NZ);
  const int l = (int)((cell / g.NZ) % g.NY);
  const int k = (int)(cell / ((int64_t)g.NZ * g.NY));
  const ParticleDraw d = draw_particle(seed, sp, 0u, (uint64_t)i);
  P.x[i] = (k + 2) * dx + d.ux * dx;
  P.y[i] = (l + 2) * dx + d.uy * dx;
  P.z[i] = (m + 2) * dx + d.uz * dx;
  P.vx[i] = vth * d.n0 + mu0;
  P.vy[i] = vth * d.n1 + mu1;
  P.vz[i] = vth * d.n2 + mu2;
}

// add_*_flat refill (:6042-6104): batch index b = ((face*NX + k)*NY + l)*2*ppc
// + m*ppc + j, written after the survivors at offset base.
__global__ void k_refill(Geom g, Particles64 P, int64_t base, int64_t count,
                         int ppc, double dx, double vth, double mu0, double mu1,
                         double mu2, uint64_t seed, int sp, uint32_t event) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= count) return;
  const int64_t per_face = (int64_t)g.NX * g.NY * 2 * ppc;
  const int face = (int)(b / per_face);
  const int64_t r = b - face * per_face;
  const int m = (int)((r / ppc) % 2);
  const int l = (int)((r / (2 * ppc)) % g.NY);
  const int k = (int)(r / ((int64_t)2 * ppc * g.NY));
  const ParticleDraw d = draw_particle(seed, sp, event, (uint64_t)b);
  const double zlo = face == 0 ? (m + 2) * dx : (m + g.NZ) * dx;
  const int64_t o = base + b;
  P.x[o] = (k + 2) * dx + d.ux * dx;
  P.y[o] = (l + 2) * dx + d.uy * dx;
  P.z[o] = zlo + d.uz * dx;
  P.vx[o] = vth * d.n0 + mu0;
  P.vy[o] = vth * d.n1 + mu1;
  P.vz[o] = vth * d.n2 + mu2;
}

// inflow deletion flags (:6019-6021): keep live particles outside the slabs
__global__ void k_inflow_flags(const double *__restrict__ z, int64_t n,
                               double lo, double hi, uint8_t *flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = z[i];
  flags[i] = (!isnan(v) && !(v < lo || v > hi)) ? 1 : 0;
}

__global__ void k_live_flags(const double *__restrict__ z, int64_t n,
                             uint8_t *flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flags[i] = isnan(z[i]) ? 0 : 1;
}

__global__ void k_gather_particles(Particles64 dst, Particles64 src,
                                   const uint32_t *__restrict__ idx,
                                   int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = idx[i];
  dst.x[i] = src.x[j];
  dst.y[i] = src.y[j];
  dst.z[i] = src.z[j];
  dst.vx[i] = src.vx[j];
  dst.vy[i] = src.vy[j];
  dst.vz[i] = src.vz[j];
}

// sorting key (:7488-7497): ((int)(x/dx)-2, (int)(y/dx)-2, (int)(z/dx)-2),
// tile-major (the reference's lexicographic order with m fastest when one
// tile spans the interior); particles outside the interior (and dead ones)
// get the key ncell and are dropped after the sort.
__global__ __launch_bounds__(256) void k_sort_keys(Geom g, TileGeom tg,
                                                   Particles64 P, int64_t n,
                                                   double dx, uint32_t *keys) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t ncell = (uint32_t)g.NX * g.NY * g.NZ;
  uint32_t key = ncell;
  const double x = P.x[i], y = P.y[i], z = P.z[i];
  const double xd = x / dx, yd = y / dx, zd = z / dx;
  // range test before the conversions so that (int) never sees a value
  // outside int range or NaN
  if (xd > -1.0 && xd < g.NX + 3 && yd > -1.0 && yd < g.NY + 3 &&
      zd > -1.0 && zd < g.NZ + 3) {
    const int k = (int)(xd) - 2, l = (int)(yd) - 2, m = (int)(zd) - 2;
    if (k >= 0 && k < g.NX && l >= 0 && l < g.NY && m >= 0 && m < g.NZ)
      key = tg.key(k, l, m);
  }
  keys[i] = key;
}

__global__ __launch_bounds__(256) void k_count_live(
    const double *__restrict__ z, int64_t n, int64_t *count) {
  unsigned int t = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    t += !isnan(z[i]);
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
// diagnostics: per-block partial sums, summed on the host in block order

constexpr int kDiagField = 7;  // E2, B2, B1^2, charge, gauss^2, gauss max, -
constexpr int kDiagBox = 2;
constexpr int kDiagPart = 1 + 1 + 3 + 15 + 1;  // kin, kin_nr, mom, force, box

template <int NQ>
__device__ inline void block_reduce_store(double (&v)[NQ], double *out,
                                          const bool *is_max) {
  __shared__ double sh[NQ][kBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double a = v[q];
    for (int o = 32; o > 0; o >>= 1) {
      const double b = __shfl_xor(a, o);
      a = is_max[q] ? fmax(a, b) : a + b;
    }
    if (lane == 0) sh[q][w] = a;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double a = sh[q][0];
      for (int j = 1; j < kBlock / 64; ++j)
        a = is_max[q] ? fmax(a, sh[q][j]) : a + sh[q][j];
      out[(int64_t)blockIdx.x * NQ + q] = a;
    }
  }
}

__global__ __launch_bounds__(256) void k_diag_fields(Geom g, Fields<double> F,
                                                     double dx, double eps0,
                                                     double mu0, double q,
                                                     double *out) {
  constexpr int NQ = kDiagField + kDiagBox;
  double v[NQ] = {0};
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
    const double bx = F.f[PIC_BX][c], by = F.f[PIC_BY][c], bz = F.f[PIC_BZ][c];
    const double b1x = bx - F.f[PIC_B0X][c], b1y = by - F.f[PIC_B0Y][c],
                 b1z = bz - F.f[PIC_B0Z][c];
    const double e2 = 0.5 * (ex * ex + ey * ey + ez * ez) * eps0;
    const double b2 = 0.5 * (bx * bx + by * by + bz * bz) / mu0;
    v[0] += e2 * dx3;
    v[1] += b2 * dx3;
    v[2] += 0.5 * (b1x * b1x + b1y * b1y + b1z * b1z) / mu0 * dx3;
    const double rho = F.f[PIC_NI][c] - F.f[PIC_NE][c];
    v[3] += q * rho * dx3;
    const double *Ex = F.f[PIC_EX], *Ey = F.f[PIC_EY], *Ez = F.f[PIC_EZ];
    const double r =
        -q * rho / eps0 +
        (-Ex[c + 2 * dk] + 27 * Ex[c + dk] - 27 * Ex[c] + Ex[c - dk] -
         Ey[c + 2 * dl] + 27 * Ey[c + dl] - 27 * Ey[c] + Ey[c - dl] -
         Ez[c + 2] + 27 * Ez[c + 1] - 27 * Ez[c] + Ez[c - 1]) /
            (24. * dx);
    v[4] += r * r;
    v[5] = fmax(v[5], fabs(r));
    if (k - 2 < g.NX * 0.8 && k - 2 > g.NX * 0.2 && l - 2 < g.NY * 0.8 &&
        l - 2 > g.NY * 0.2 && m - 2 < g.NZ * 0.8 && m - 2 > g.NZ * 0.2) {
      v[7] += e2;
      v[8] += b2;
    }
  }
  const bool is_max[NQ] = {false, false, false, false, false, true,
                           false, false, false};
  block_reduce_store<NQ>(v, out, is_max);
}

__global__ __launch_bounds__(256) void k_diag_particles(
    Geom g, Fields<double> F, Particles64 P, int64_t n, double dx, double c,
    double mass, double w, double qs, double *out) {
  constexpr int NQ = kDiagPart;
  double v[NQ] = {0};
  const double dx3 = dx * dx * dx;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double z = P.z[i];
    if (isnan(z)) continue;
    const double x = P.x[i], y = P.y[i];
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
        in_array(g, x, y, z, dx)) {
      double f[6];
      gather(F, g, x, y, z, dx, f);
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
      v[14 + 6] += 0.5 * mass * w * dx3 * v2;
    }
  }
  bool is_max[NQ];
  for (int q = 0; q < NQ; ++q) is_max[q] = false;
  block_reduce_store<NQ>(v, out, is_max);
}

// ---------------------------------------------------------------------------
// host side

static dim3 interior_grid(const Geom &g, dim3 &block) {
  block = dim3(64, 4, 1);
  return dim3((unsigned)div_up(g.NZ, 64), (unsigned)div_up(g.NY, 4),
              (unsigned)g.NX);
}

int renew_grid(pic_ctx *ctx) {
  Fields<double> F = fields<double>(ctx);
  // J (6 planes) and n (2 planes) are contiguous: PIC_JIX..PIC_NE
  const int64_t n = 8 * ctx->geom.plane;
  k_zero<<<2048, kBlock, 0, ctx->stream>>>(F.f[PIC_JIX], n);
  PIC_LAUNCH_CHECK(ctx);
  return 0;
}

int renew_b_half(pic_ctx *ctx) {
  const pic_config &c = ctx->cfg;
  dim3 b, gr = interior_grid(ctx->geom, b);
  k_renew_b<<<gr, b, 0, ctx->stream>>>(ctx->geom, fields<double>(ctx),
                                       sigma(ctx), c.dt, c.dx, c.eps0);
  PIC_LAUNCH_CHECK(ctx);
  return 0;
}

int renew_e(pic_ctx *ctx) {
  const pic_config &c = ctx->cfg;
  dim3 b, gr = interior_grid(ctx->geom, b);
  k_renew_e<<<gr, b, 0, ctx->stream>>>(ctx->geom, fields<double>(ctx),
                                       sigma(ctx), c.dt, c.dx, c.c, c.mu0);
  PIC_LAUNCH_CHECK(ctx);
  return 0;
}

int halo(pic_ctx *ctx, int first) {
  const Geom &g = ctx->geom;
  Fields<double> F = fields<double>(ctx);
  double *a = F.f[first], *b = F.f[first + 1], *c = F.f[first + 2];
  k_halo_x<<<(unsigned)div_up((int64_t)g.SY * g.SZ, kBlock), kBlock, 0,
             ctx->stream>>>(g, a, b, c);
  k_halo_y<<<(unsigned)div_up((int64_t)g.SX * g.SZ, kBlock), kBlock, 0,
             ctx->stream>>>(g, a, b, c);
  PIC_LAUNCH_CHECK(ctx);
  return 0;
}

int fold_current(pic_ctx *ctx) {
  const Geom &g = ctx->geom;
  Fields<double> F = fields<double>(ctx);
  k_fold_x<<<(unsigned)div_up((int64_t)g.SY * g.SZ, kBlock), kBlock, 0,
             ctx->stream>>>(g, F);
  k_fold_y<<<(unsigned)div_up((int64_t)g.SX * g.SZ, kBlock), kBlock, 0,
             ctx->stream>>>(g, F);
  PIC_LAUNCH_CHECK(ctx);
  return 0;
}

int load_particles(pic_ctx *ctx) {
  const pic_config &c = ctx->cfg;
  const int64_t n = (int64_t)c.nx * c.ny * c.nz * c.ppc;
  if (n > ctx->cap)
    return set_error(ctx, PIC_E_CAPACITY, "load: %lld particles > capacity %lld",
                     (long long)n, (long long)ctx->cap);
  for (int sp = 0; sp < 2; ++sp) {
    Particles64 P = p64(ctx, sp, ctx->cur[sp]);
    if (n > 0)
      k_load<<<(unsigned)div_up(n, kBlock), kBlock, 0, ctx->stream>>>(
          ctx->geom, P, n, c.ppc, c.dx, c.vth[sp], c.vdrift[0], c.vdrift[1],
          c.vdrift[2], c.seed, sp);
    ctx->n[sp] = n;
    ctx->sorted_end[sp] = 0;
  }
  PIC_LAUNCH_CHECK(ctx);
  // the loaded order is the reference's; bring it to tile order (the same
  // order when one tile spans the interior)
  for (int sp = 0; sp < 2; ++sp) {
    int rc = sort(ctx, sp);
    if (rc) return rc;
  }
  return 0;
}

int push(pic_ctx *ctx, int sp, int do_push_j, int do_drift_n,
         int do_boundary) {
  const int64_t n = ctx->n[sp];
  if (n == 0) return 0;
  Fields<double> F = fields<double>(ctx);
  Particles64 P = p64(ctx, sp, ctx->cur[sp]);
  StepConst kc = step_const(ctx, sp);
  double *jx = F.f[sp == PIC_ION ? PIC_JIX : PIC_JEX];
  double *jy = F.f[sp == PIC_ION ? PIC_JIY : PIC_JEY];
  double *jz = F.f[sp == PIC_ION ? PIC_JIZ : PIC_JEZ];
  double *nd = F.f[sp == PIC_ION ? PIC_NI : PIC_NE];
  const TileGeom tg = tile_geom(ctx);
  TileRun tr;
  tr.tile_start = ctx->tile_start[sp];
  tr.sorted_end = ctx->sorted_end[sp];
  // a single tile (the reference's order) is processed as tail chunks
  tr.ntiles = (tr.sorted_end > 0 && tg.ntiles() > 1) ? tg.ntiles() : 0;
  if (tr.ntiles == 0) tr.sorted_end = 0;
  tr.n = n;
  tr.ntail = (int)div_up(n - tr.sorted_end, kTailChunk);
  const unsigned nb = (unsigned)(tr.ntiles + tr.ntail);
  const int sel = (do_push_j ? 4 : 0) | (do_drift_n ? 2 : 0) | (do_boundary ? 1 : 0);
  // the fused pass of pic_step adds the previous step's thrust_fave
  const bool thrust = ctx->thrust_now && sel == 7;
  Fields<double> Fp = F;
  double *tp = nullptr;
  if (thrust) {
    if ((int64_t)nb > ctx->thrust_part_slots)
      return set_error(ctx, PIC_E_CAPACITY, "thrust: %u workgroups, %lld partial slots",
                       nb, (long long)ctx->thrust_part_slots);
    const double *bp = (const double *)ctx->thrust_bprev;
    for (int q = 0; q < 3; ++q) Fp.f[PIC_BX + q] = (double *)bp + q * ctx->geom.plane;
    tp = ctx->thrust_part;
  }
#define LAUNCH(A, B, C)                                                    \
  k_particles_tiled<A, B, C, false><<<nb, kBlock, 0, ctx->stream>>>(        \
      ctx->geom, tg, F, P, tr, kc, jx, jy, jz, nd, Fp, tp)
  switch (thrust ? 8 : sel) {
    case 8:
      k_particles_tiled<true, true, true, true><<<nb, kBlock, 0, ctx->stream>>>(
          ctx->geom, tg, F, P, tr, kc, jx, jy, jz, nd, Fp, tp);
      break;
    case 7: LAUNCH(true, true, true); break;
    case 6: LAUNCH(true, true, false); break;
    case 4: LAUNCH(true, false, false); break;
    case 2: LAUNCH(false, true, false); break;
    case 1: LAUNCH(false, false, true); break;
    case 3: LAUNCH(false, true, true); break;
    case 5: LAUNCH(true, false, true); break;
    default: return 0;
  }
#undef LAUNCH
  PIC_LAUNCH_CHECK(ctx);
  if (thrust) {
    const pic_config &c = ctx->cfg;
    const double qs = sp == PIC_ION ? c.q : -c.q;
    return thrust_add(ctx, sp, (int)nb, qs * c.weight[sp] * c.dx * c.dx * c.dx);
  }
  return 0;
}

int particle_boundary(pic_ctx *ctx, int sp) { return push(ctx, sp, 0, 0, 1); }

// Stable compaction of the current buffer into the other one through the
// flags already written to ctx->flags; returns the survivor count.
static int compact(pic_ctx *ctx, int sp, int64_t *kept) {
  const int64_t n = ctx->n[sp];
  int rc = select_flagged(ctx, ctx->flags, ctx->idx, n, kept);
  if (rc) return rc;
  Particles64 src = p64(ctx, sp, ctx->cur[sp]);
  Particles64 dst = p64(ctx, sp, 1 - ctx->cur[sp]);
  if (*kept > 0)
    k_gather_particles<<<(unsigned)div_up(*kept, kBlock), kBlock, 0,
                         ctx->stream>>>(dst, src, ctx->idx, *kept);
  PIC_LAUNCH_CHECK(ctx);
  ctx->cur[sp] = 1 - ctx->cur[sp];
  ctx->n[sp] = *kept;
  return 0;
}

// add_*_flat: slab particles become dead slots (z = NaN) in place and the
// refill is appended, so tile ranges stay valid until the next sort; the
// store is compacted first only when the refill would not fit
__global__ void k_inflow_kill(double *z, int64_t n, double lo, double hi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = z[i];
  if (!isnan(v) && (v < lo || v > hi)) z[i] = __longlong_as_double(0x7ff8000000000000LL);
}

int inflow(pic_ctx *ctx, int sp, uint32_t event) {
  const pic_config &c = ctx->cfg;
  const Geom &g = ctx->geom;
  if (ctx->n[sp] > 0) {
    Particles64 P = p64(ctx, sp, ctx->cur[sp]);
    k_inflow_kill<<<(unsigned)div_up(ctx->n[sp], kBlock), kBlock, 0,
                    ctx->stream>>>(P.z, ctx->n[sp], 4 * c.dx, g.NZ * c.dx);
    PIC_LAUNCH_CHECK(ctx);
  }
  const int nface = c.refill_far ? 2 : 1;
  const int64_t add = (int64_t)g.NX * g.NY * 2 * c.ppc * nface;
  if (ctx->n[sp] + add > ctx->cap) {
    Particles64 P = p64(ctx, sp, ctx->cur[sp]);
    k_live_flags<<<(unsigned)div_up(ctx->n[sp], kBlock), kBlock, 0,
                   ctx->stream>>>(P.z, ctx->n[sp], ctx->flags);
    PIC_LAUNCH_CHECK(ctx);
    int64_t kept = 0;
    int rc = compact(ctx, sp, &kept);
    if (rc) return rc;
    ctx->sorted_end[sp] = 0;
  }
  if (ctx->n[sp] + add > ctx->cap)
    return set_error(ctx, PIC_E_CAPACITY,
                     "inflow: %lld + %lld particles exceed capacity %lld",
                     (long long)ctx->n[sp], (long long)add, (long long)ctx->cap);
  if (add > 0) {
    Particles64 P = p64(ctx, sp, ctx->cur[sp]);
    k_refill<<<(unsigned)div_up(add, kBlock), kBlock, 0, ctx->stream>>>(
        g, P, ctx->n[sp], add, c.ppc, c.dx, c.vth[sp], c.vdrift[0],
        c.vdrift[1], c.vdrift[2], c.seed, sp, event);
    PIC_LAUNCH_CHECK(ctx);
  }
  ctx->n[sp] += add;
  return 0;
}

// The tile sort's destination map of the FP64 form (pic_tsort.h): the
// particles carry no cell index, so the sort runs on the key array of
// k_sort_keys (tile * cells per tile + cell in the tile, or the drop key);
// the home slot is the tile coordinates modulo 4, as in the FP32 form.
struct KeyMapKey {
  TileGeom tg;
  uint32_t drop;  // tiles x cells per tile
  __host__ __device__ uint32_t cells() const { return (uint32_t)(tg.TX * tg.TY * tg.TZ); }
  __device__ bool dest(int32_t c, uint32_t &t, uint32_t &cin, int &h) const {
    const uint32_t key = (uint32_t)c;
    if (key >= drop) return false;
    t = tg.tile_of(key);
    cin = tg.cell_of(key);
    const uint32_t tz = t % (uint32_t)tg.NTZ, r = t / (uint32_t)tg.NTZ;
    const uint32_t ty = r % (uint32_t)tg.NTY, tx = r / (uint32_t)tg.NTY;
    h = (int)(((tx & 3u) << 4) | ((ty & 3u) << 2) | (tz & 3u));
    return true;
  }
};

int tsort_setup(pic_ctx *ctx) {
  const TileGeom tg = tile_geom(ctx);
  // (slots are 32-bit in the tile sort: larger stores take the radix sort)
  if (tg.ntiles() <= 1 || tg.TX * tg.TY * tg.TZ > tsort::kTileCells ||
      ctx->cap > ((int64_t)1 << 30))
    return 0;
  ctx->ts64 = new (std::nothrow) tsort::Bufs();
  if (!ctx->ts64) return set_error(ctx, PIC_E_NOMEM, "tile sort lists");
  return tsort::reserve(ctx, *ctx->ts64, tg.ntiles(), ctx->cap);
}

void tsort_free(pic_ctx *ctx) {
  if (!ctx->ts64) return;
  tsort::release(*ctx->ts64);
  delete ctx->ts64;
  ctx->ts64 = nullptr;
}

int sort(pic_ctx *ctx, int sp) {
  const Geom &g = ctx->geom;
  const int64_t n = ctx->n[sp];
  if (n == 0) return 0;
  Particles64 src = p64(ctx, sp, ctx->cur[sp]);
  const TileGeom tg = tile_geom(ctx);
  const int nt = tg.ntiles();
  k_sort_keys<<<(unsigned)div_up(n, kBlock), kBlock, 0, ctx->stream>>>(
      g, tg, src, n, ctx->cfg.dx, ctx->keys);
  PIC_LAUNCH_CHECK(ctx);
  const uint32_t ncell = (uint32_t)g.NX * g.NY * g.NZ;
  if (ctx->ts64 && ctx->sorted_end[sp] > 0 && ctx->sorted_end[sp] <= n) {
    // the tiles of the last sort are current: the tile sort of the keys
    const KeyMapKey km{tg, (uint32_t)nt * (uint32_t)(tg.TX * tg.TY * tg.TZ)};
    bool done = false;
    int64_t valid = 0;
    int rc = tsort::run(ctx, km, nt, ctx->tile_start[sp], ctx->sorted_end[sp], n,
                        (const int32_t *)ctx->keys, *ctx->ts64, ctx->idx, &done, &valid);
    if (rc) return rc;
    if (done) {
      Particles64 dst = p64(ctx, sp, 1 - ctx->cur[sp]);
      if (valid > 0)
        k_gather_particles<<<(unsigned)div_up(valid, kBlock), kBlock, 0, ctx->stream>>>(
            dst, src, ctx->idx, valid);
      PIC_LAUNCH_CHECK(ctx);
      ctx->cur[sp] = 1 - ctx->cur[sp];
      ctx->n[sp] = valid;
      ctx->sorted_end[sp] = valid;
      ctx->sort_tile_count++;
      return 0;
    }
  }
  ctx->sort_radix_count++;
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
  Particles64 dst = p64(ctx, sp, 1 - ctx->cur[sp]);
  if (valid > 0)
    k_gather_particles<<<(unsigned)div_up(valid, kBlock), kBlock, 0,
                         ctx->stream>>>(dst, src, ctx->idx, valid);
  PIC_LAUNCH_CHECK(ctx);
  ctx->cur[sp] = 1 - ctx->cur[sp];
  ctx->n[sp] = valid;
  ctx->sorted_end[sp] = valid;
  return 0;
}

int count_live(pic_ctx *ctx, int sp, int64_t *out) {
  const int64_t n = ctx->n[sp];
  *out = 0;
  if (n == 0) return 0;
  Particles64 P = p64(ctx, sp, ctx->cur[sp]);
  PIC_HIP(ctx, hipMemsetAsync(ctx->dev_counter, 0, sizeof(int64_t), ctx->stream));
  k_count_live<<<(unsigned)(div_up(n, kBlock) < 2048 ? div_up(n, kBlock) : 2048),
                 kBlock, 0, ctx->stream>>>(P.z, n, ctx->dev_counter);
  PIC_LAUNCH_CHECK(ctx);
  PIC_HIP(ctx, hipMemcpyAsync(ctx->host_counter, ctx->dev_counter,
                              sizeof(int64_t), hipMemcpyDeviceToHost,
                              ctx->stream));
  PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *out = ctx->host_counter[0];
  return 0;
}

static_assert(kDiagField + kDiagBox == kDiagNQF && kDiagPart == kDiagNQP,
              "diagnostics layout");

int diagnose_partials(pic_ctx *ctx, double *pf, unsigned long long *counts) {
  const pic_config &c = ctx->cfg;
  const Geom &g = ctx->geom;
  Fields<double> F = fields<double>(ctx);
  double *pp[2] = {pf + (int64_t)kDiagNBF * kDiagNQF,
                   pf + (int64_t)kDiagNBF * kDiagNQF + (int64_t)kDiagNBP * kDiagNQP};
  k_diag_fields<<<kDiagNBF, kBlock, 0, ctx->stream>>>(g, F, c.dx, c.eps0, c.mu0,
                                                      c.q, pf);
  PIC_LAUNCH_CHECK(ctx);
  for (int sp = 0; sp < 2; ++sp) {
    Particles64 P = p64(ctx, sp, ctx->cur[sp]);
    k_diag_particles<<<kDiagNBP, kBlock, 0, ctx->stream>>>(
        g, F, P, ctx->n[sp], c.dx, c.c, c.mass[sp], c.weight[sp],
        sp == PIC_ION ? c.q : -c.q, pp[sp]);
    PIC_LAUNCH_CHECK(ctx);
    if (counts && ctx->n[sp] > 0) {
      k_count_live<<<(unsigned)(div_up(ctx->n[sp], kBlock) < 2048
                                    ? div_up(ctx->n[sp], kBlock) : 2048),
                     kBlock, 0, ctx->stream>>>(P.z, ctx->n[sp],
                                               (int64_t *)(counts + sp));
      PIC_LAUNCH_CHECK(ctx);
    }
  }
  return 0;
}

int diagnose(pic_ctx *ctx, pic_diag *out) {
  const int nbf = kDiagNBF, nbp = kDiagNBP;
  constexpr int NQF = kDiagNQF, NQP = kDiagNQP;
  const int64_t need = kDiagPartials;
  if (ctx->diag_partial_cap < need) {
    if (ctx->diag_partial) (void)hipFree(ctx->diag_partial);
    ctx->diag_partial = nullptr;
    PIC_HIP(ctx, hipMalloc(&ctx->diag_partial, need * sizeof(double)));
    ctx->diag_partial_cap = need;
  }
  double *pf = ctx->diag_partial;
  double *pp[2] = {pf + (int64_t)nbf * NQF, pf + (int64_t)nbf * NQF + (int64_t)nbp * NQP};
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
  double s[NQF] = {0};
  for (int b = 0; b < nbf; ++b)
    for (int q = 0; q < NQF; ++q) {
      const double v = h[(int64_t)b * NQF + q];
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
    double t[NQP] = {0};
    const double *hp = h + (pp[sp] - pf);
    for (int b = 0; b < nbp; ++b)
      for (int q = 0; q < NQP; ++q) t[q] += hp[(int64_t)b * NQP + q];
    out->kinetic[sp] = t[0];
    out->kinetic_nr[sp] = t[1];
    out->momentum[sp][0] = t[2];
    out->momentum[sp][1] = t[3];
    out->momentum[sp][2] = t[4];
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

}  // namespace f64
}  // namespace pic

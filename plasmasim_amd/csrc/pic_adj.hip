// Hand-written discrete adjoint of the 3D step on gfx950: gradients of the
// step-averaged Maxwell-stress flux through a sphere around the dipole with
// respect to p = (Mx, My, Mz, V), the dipole moment vector and the inflow
// drift speed.
//
// Forward (pic_adj_forward): the context's own step (either numeric form)
// with the cell sort at its cadence and the inflow appended to the store
// (each injected batch contiguous; the particle boundary marks dead slots
// in place); a checkpoint of the whole state every ckpt_every steps.
//
// Backward (pic_adj_backward): per segment, from the last, the checkpoint is
// restored and the segment recomputed while its tape is recorded (particles
// and fields in FP64 at the start of each step, and each sort's
// permutation); the steps are then reversed
// in FP64: transposed (2,4) FDTD, guard copies and fold, the loss and the
// particle pass (transposed gather, relativistic Boris step and current
// deposit, recomputed from the tape), and across a sort the particle
// adjoints move back to the slots the particles held before it.  All
// adjoint planes, particle adjoints and the parameter gradients accumulate
// in FP64.
//
// The forward state is the B1 = B - B0 formulation: B0 = sum_a M_a G_a + IMF
// enters the gather, the loss and the guard copies of the total B; V enters
// the initial and injected v_z = V + sigma N(0,1) and the motional field.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "pic_internal.h"

namespace pic {
namespace {

constexpr double kPi = 3.14159265358979323846;

#define PIC_GLB __attribute__((address_space(1)))

#define PIC_LDS __attribute__((address_space(3)))

__device__ inline void gadd(double *p, double v) {
  __hip_atomic_fetch_add((PIC_GLB double *)p, v, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}

// staggering (x, y, z; 1 = half) of Ex Ey Ez Bx By Bz and of Jx Jy Jz
__constant__ int kStg[6][3] = {{0, 1, 1}, {1, 0, 1}, {1, 1, 0},
                               {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

struct Stn {
  int b[3];
  double w[3][2];
};

// q = p / dx, computed once per position by callers that build several
// stencils at it
__device__ inline void stencil_q(double dx, const double p[3], const double q[3], int c,
                                 Stn &s) {
  for (int a = 0; a < 3; ++a) {
    const double h = kStg[c][a] ? 0.5 : 0.0;
    const int k = (int)(q[a] - h);
    s.b[a] = k;
    s.w[a][0] = (k + 1 + h) * dx - p[a];
    s.w[a][1] = p[a] - (k + h) * dx;
  }
}

__device__ inline void stencil(double dx, const double p[3], int c, Stn &s) {
  const double q[3] = {p[0] / dx, p[1] / dx, p[2] / dx};
  stencil_q(dx, p, q, c, s);
}

// value and position gradient of sum_n F[n] W_n (W unnormalised)
__device__ inline double st_gather(const Geom &g, const double *F, const Stn &s,
                                   double gr[3]) {
  double v = 0;
  gr[0] = gr[1] = gr[2] = 0;
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const int i = n >> 2, j = (n >> 1) & 1, k = n & 1;
    const double f = F[g.idx(s.b[0] + i, s.b[1] + j, s.b[2] + k)];
    const double wx = s.w[0][i], wy = s.w[1][j], wz = s.w[2][k];
    v += f * wx * wy * wz;
    gr[0] += f * (i ? 1.0 : -1.0) * wy * wz;
    gr[1] += f * wx * (j ? 1.0 : -1.0) * wz;
    gr[2] += f * wx * wy * (k ? 1.0 : -1.0);
  }
  return v;
}

__device__ inline double st_value(const Geom &g, const double *F, const Stn &s) {
  double v = 0;
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const int i = n >> 2, j = (n >> 1) & 1, k = n & 1;
    v += F[g.idx(s.b[0] + i, s.b[1] + j, s.b[2] + k)] * s.w[0][i] * s.w[1][j] *
         s.w[2][k];
  }
  return v;
}

__device__ inline void st_scatter(const Geom &g, double *F, const Stn &s, double val) {
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const int i = n >> 2, j = (n >> 1) & 1, k = n & 1;
    gadd(F + g.idx(s.b[0] + i, s.b[1] + j, s.b[2] + k),
         val * s.w[0][i] * s.w[1][j] * s.w[2][k]);
  }
}

__device__ inline bool in_array_q(const Geom &g, const double q[3]) {
  return q[0] > -0.5 && q[0] < g.SX - 1 && q[1] > -0.5 && q[1] < g.SY - 1 &&
         q[2] > -0.5 && q[2] < g.SZ - 1;
}

__device__ inline double sigma_at(const SigmaProfile &s, int k2, int l2, int m2) {
  const double a4 = s.d4[0][k2] + s.d4[1][l2] + s.d4[2][m2];
  if (a4 == 0.0) return 0.0;
  const double a12 = s.d12[0][k2] + s.d12[1][l2] + s.d12[2][m2];
  return s.amp * (0.1 * sqrt(a4) + 0.01 * sqrt(a12));
}

// planes of the adjoint: E, B1, B0 (3 each), J ion and electron (3 each),
// and three scratch planes
struct Lam {
  double *E[3], *B[3], *B0[3], *J[2][3], *S[3];
};

// the FP64 field set the reverse pass reads: E and the total B
struct F6 {
  double *f[6];
};

// ---------------------------------------------------------------------------
// transposed field operators

__device__ inline double at(const Geom &g, const double *p, int k, int l, int m) {
  if (k < 0 || l < 0 || m < 0 || k >= g.SX || l >= g.SY || m >= g.SZ) return 0.0;
  return p[g.idx(k, l, m)];
}

// (D_a^T q)[j], D_a f[c] = 27 (f[c] - f[c-a]) - (f[c+a] - f[c-2a])
__device__ inline double dT_back(const Geom &g, const double *q, int k, int l,
                                 int m, int a) {
  const int dk = a == 0, dl = a == 1, dm = a == 2;
  return 27. * at(g, q, k, l, m) - 27. * at(g, q, k + dk, l + dl, m + dm) -
         at(g, q, k - dk, l - dl, m - dm) + at(g, q, k + 2 * dk, l + 2 * dl, m + 2 * dm);
}

// (G_a^T q)[j], G_a f[c] = 27 (f[c+a] - f[c]) - (f[c+2a] - f[c-a])
__device__ inline double dT_fwd(const Geom &g, const double *q, int k, int l,
                                int m, int a) {
  const int dk = a == 0, dl = a == 1, dm = a == 2;
  return 27. * at(g, q, k - dk, l - dl, m - dm) - 27. * at(g, q, k, l, m) -
         at(g, q, k - 2 * dk, l - 2 * dl, m - 2 * dm) + at(g, q, k + dk, l + dl, m + dm);
}

__device__ inline bool interior(const Geom &g, int k, int l, int m) {
  return k >= 2 && k < g.NX + 2 && l >= 2 && l < g.NY + 2 && m >= 2 && m < g.NZ + 2;
}

// renew_b4_2 transposed, phase 1: S = h sc lB, lB *= f (interior)
__global__ void k_rb_T1(Geom g, Lam L, SigmaProfile sg, double temp, double dth,
                        double sc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nc = (int64_t)g.SX * g.SY * g.SZ;
  if (t >= nc) return;
  const int m = (int)(t % g.SZ), l = (int)((t / g.SZ) % g.SY), k = (int)(t / ((int64_t)g.SZ * g.SY));
  if (!interior(g, k, l, m)) {
    L.S[0][t] = L.S[1][t] = L.S[2][t] = 0.0;
    return;
  }
  const double s[3] = {sigma_at(sg, 2 * k + 1, 2 * l, 2 * m),
                       sigma_at(sg, 2 * k, 2 * l + 1, 2 * m),
                       sigma_at(sg, 2 * k, 2 * l, 2 * m + 1)};
  for (int a = 0; a < 3; ++a) {
    const double f = (1. - s[a] * temp) / (1. + s[a] * temp);
    const double h = dth / (1. + s[a] * temp);
    L.S[a][t] = h * sc * L.B[a][t];
    L.B[a][t] *= f;
  }
}

// phase 2: lE += -curl^T S over the whole array
__global__ void k_rb_T2(Geom g, Lam L) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nc = (int64_t)g.SX * g.SY * g.SZ;
  if (t >= nc) return;
  const int m = (int)(t % g.SZ), l = (int)((t / g.SZ) % g.SY), k = (int)(t / ((int64_t)g.SZ * g.SY));
  L.E[2][t] += -dT_back(g, L.S[0], k, l, m, 1) + dT_back(g, L.S[1], k, l, m, 0);
  L.E[1][t] += dT_back(g, L.S[0], k, l, m, 2) - dT_back(g, L.S[2], k, l, m, 0);
  L.E[0][t] += -dT_back(g, L.S[1], k, l, m, 2) + dT_back(g, L.S[2], k, l, m, 1);
}

// renew_e4_2 transposed, phase 1: S = h sc lE, lJ, lE *= f (interior)
__global__ void k_re_T1(Geom g, Lam L, SigmaProfile sg, double temp, double mu0,
                        double sc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nc = (int64_t)g.SX * g.SY * g.SZ;
  if (t >= nc) return;
  const int m = (int)(t % g.SZ), l = (int)((t / g.SZ) % g.SY), k = (int)(t / ((int64_t)g.SZ * g.SY));
  if (!interior(g, k, l, m)) {
    L.S[0][t] = L.S[1][t] = L.S[2][t] = 0.0;
    return;
  }
  const double s[3] = {sigma_at(sg, 2 * k, 2 * l + 1, 2 * m + 1),
                       sigma_at(sg, 2 * k + 1, 2 * l, 2 * m + 1),
                       sigma_at(sg, 2 * k + 1, 2 * l + 1, 2 * m)};
  for (int a = 0; a < 3; ++a) {
    const double f = (1. - 0.5 * mu0 * s[a] * temp) / (1. + 0.5 * mu0 * s[a] * temp);
    const double h = temp / (1. + 0.5 * mu0 * s[a] * temp);
    const double le = L.E[a][t];
    L.S[a][t] = h * sc * le;
    L.J[0][a][t] += -mu0 * h * le;
    L.J[1][a][t] += mu0 * h * le;
    L.E[a][t] = f * le;
  }
}

// phase 2: lB += curl^T S
__global__ void k_re_T2(Geom g, Lam L) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nc = (int64_t)g.SX * g.SY * g.SZ;
  if (t >= nc) return;
  const int m = (int)(t % g.SZ), l = (int)((t / g.SZ) % g.SY), k = (int)(t / ((int64_t)g.SZ * g.SY));
  L.B[2][t] += dT_fwd(g, L.S[0], k, l, m, 1) - dT_fwd(g, L.S[1], k, l, m, 0);
  L.B[1][t] += -dT_fwd(g, L.S[0], k, l, m, 2) + dT_fwd(g, L.S[2], k, l, m, 0);
  L.B[0][t] += dT_fwd(g, L.S[1], k, l, m, 2) - dT_fwd(g, L.S[2], k, l, m, 1);
}

// one guard copy dst <- src transposed (B: total-field copy, so the B0
// adjoint takes +lam at src and -lam at dst)
__device__ inline void copy_T(double *lam, double *lam0, int64_t dst, int64_t src) {
  const double v = lam[dst];
  lam[src] += v;
  lam[dst] = 0.0;
  if (lam0) {
    lam0[src] += v;
    lam0[dst] -= v;
  }
}

// sr_b / sr_e transposed: y faces (launch first), then x faces
__global__ void k_halo_T_y(Geom g, double *l0, double *l1, double *l2, double *m0,
                           double *m1, double *m2) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)g.SX * g.SZ) return;
  const int k = (int)(t / g.SZ), m = (int)(t % g.SZ), NY = g.NY;
  double *L[3] = {l0, l1, l2}, *M[3] = {m0, m1, m2};
  for (int c = 0; c < 3; ++c) {
    copy_T(L[c], M[c], g.idx(k, NY + 3, m), g.idx(k, 3, m));
    copy_T(L[c], M[c], g.idx(k, NY + 2, m), g.idx(k, 2, m));
    copy_T(L[c], M[c], g.idx(k, 1, m), g.idx(k, NY + 1, m));
    copy_T(L[c], M[c], g.idx(k, 0, m), g.idx(k, NY, m));
  }
}

__global__ void k_halo_T_x(Geom g, double *l0, double *l1, double *l2, double *m0,
                           double *m1, double *m2) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)g.SY * g.SZ) return;
  const int l = (int)(t / g.SZ), m = (int)(t % g.SZ), NX = g.NX;
  double *L[3] = {l0, l1, l2}, *M[3] = {m0, m1, m2};
  for (int c = 0; c < 3; ++c) {
    copy_T(L[c], M[c], g.idx(NX + 3, l, m), g.idx(3, l, m));
    copy_T(L[c], M[c], g.idx(NX + 2, l, m), g.idx(2, l, m));
    copy_T(L[c], M[c], g.idx(1, l, m), g.idx(NX + 1, l, m));
    copy_T(L[c], M[c], g.idx(0, l, m), g.idx(NX, l, m));
  }
}

// sr_current2 transposed: y part (launch first), then x part
__global__ void k_fold_T_y(Geom g, double *p0, double *p1, double *p2) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)g.SX * g.SZ) return;
  const int k = (int)(t / g.SZ), m = (int)(t % g.SZ), NY = g.NY;
  double *P[3] = {p0, p1, p2};
  for (int c = 0; c < 3; ++c) {
    double *L = P[c];
    L[g.idx(k, 0, m)] += L[g.idx(k, NY, m)];
    L[g.idx(k, 1, m)] += L[g.idx(k, NY + 1, m)];
    L[g.idx(k, NY + 2, m)] += L[g.idx(k, 2, m)];
    L[g.idx(k, NY + 3, m)] += L[g.idx(k, 3, m)];
  }
}

__global__ void k_fold_T_x(Geom g, double *p0, double *p1, double *p2) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)g.SY * g.SZ) return;
  const int l = (int)(t / g.SZ), m = (int)(t % g.SZ), NX = g.NX;
  double *P[3] = {p0, p1, p2};
  for (int c = 0; c < 3; ++c) {
    double *L = P[c];
    L[g.idx(0, l, m)] += L[g.idx(NX, l, m)];
    L[g.idx(1, l, m)] += L[g.idx(NX + 1, l, m)];
    L[g.idx(NX + 2, l, m)] += L[g.idx(2, l, m)];
    L[g.idx(NX + 3, l, m)] += L[g.idx(3, l, m)];
  }
}

// ---------------------------------------------------------------------------
// forward pieces recomputed in FP64 for the reverse pass

// renew_b4_2 on the total B of F with B0 (interior)
__global__ void k_rb_fwd(Geom g, F6 F, const double *b0x, const double *b0y,
                         const double *b0z, SigmaProfile sg, double temp,
                         double dth, double sc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nc = (int64_t)g.SX * g.SY * g.SZ;
  if (t >= nc) return;
  const int m = (int)(t % g.SZ), l = (int)((t / g.SZ) % g.SY), k = (int)(t / ((int64_t)g.SZ * g.SY));
  if (!interior(g, k, l, m)) return;
  const int64_t c = t, dl = g.SZ, dk = (int64_t)g.SY * g.SZ;
  const double *ex = F.f[0], *ey = F.f[1], *ez = F.f[2];
  const double *b0[3] = {b0x, b0y, b0z};
  const double s[3] = {sigma_at(sg, 2 * k + 1, 2 * l, 2 * m),
                       sigma_at(sg, 2 * k, 2 * l + 1, 2 * m),
                       sigma_at(sg, 2 * k, 2 * l, 2 * m + 1)};
  const double curl[3] = {
      27. * (ez[c] - ez[c - dl] - ey[c] + ey[c - 1]) -
          (ez[c + dl] - ez[c - 2 * dl] - ey[c + 1] + ey[c - 2]),
      27. * (ex[c] - ex[c - 1] - ez[c] + ez[c - dk]) -
          (ex[c + 1] - ex[c - 2] - ez[c + dk] + ez[c - 2 * dk]),
      27. * (ey[c] - ey[c - dk] - ex[c] + ex[c - dl]) -
          (ey[c + dk] - ey[c - 2 * dk] - ex[c + dl] + ex[c - 2 * dl])};
  for (int a = 0; a < 3; ++a) {
    const double f = (1. - s[a] * temp) / (1. + s[a] * temp);
    const double h = dth / (1. + s[a] * temp);
    F.f[3 + a][c] = b0[a][c] + f * (F.f[3 + a][c] - b0[a][c]) - h * sc * curl[a];
  }
}

// sr_b: total-B guard copies, x faces then y faces (two launches)
__global__ void k_halo_x_fwd(Geom g, double *a, double *b, double *c) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)g.SY * g.SZ) return;
  const int l = (int)(t / g.SZ), m = (int)(t % g.SZ), NX = g.NX;
  double *P[3] = {a, b, c};
  for (int j = 0; j < 3; ++j) {
    double *F = P[j];
    F[g.idx(0, l, m)] = F[g.idx(NX, l, m)];
    F[g.idx(1, l, m)] = F[g.idx(NX + 1, l, m)];
    F[g.idx(NX + 2, l, m)] = F[g.idx(2, l, m)];
    F[g.idx(NX + 3, l, m)] = F[g.idx(3, l, m)];
  }
}

__global__ void k_halo_y_fwd(Geom g, double *a, double *b, double *c) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)g.SX * g.SZ) return;
  const int k = (int)(t / g.SZ), m = (int)(t % g.SZ), NY = g.NY;
  double *P[3] = {a, b, c};
  for (int j = 0; j < 3; ++j) {
    double *F = P[j];
    F[g.idx(k, 0, m)] = F[g.idx(k, NY, m)];
    F[g.idx(k, 1, m)] = F[g.idx(k, NY + 1, m)];
    F[g.idx(k, NY + 2, m)] = F[g.idx(k, 2, m)];
    F[g.idx(k, NY + 3, m)] = F[g.idx(k, 3, m)];
  }
}

// cal_track4 (relativistic Boris, B / gamma(u^{n-1/2}), gamma1 = 0 when
// |v| >= c) without the position update
struct BorisK {
  double c, T, eimf_x, eimf_y, imf_x, imf_y;
};

__device__ inline void cross(double r[3], const double p[3], const double q[3]) {
  r[0] = p[1] * q[2] - p[2] * q[1];
  r[1] = p[2] * q[0] - p[0] * q[2];
  r[2] = p[0] * q[1] - p[1] * q[0];
}

__device__ inline void boris_fwd(const BorisK &K, const double f[6],
                                 const double v[3], double vo[3]) {
  const double C = K.c, T = K.T;
  const double E[3] = {f[0] + K.eimf_x, f[1] + K.eimf_y, f[2]};
  const double s1 = C * C - (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  const double g1 = s1 > 0 ? C / sqrt(s1) : 0.0;
  double u[3], b[3], t[3], uo[3], up[3];
  for (int a = 0; a < 3; ++a) u[a] = g1 * v[a];
  const double g2 = C / sqrt(C * C + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (int a = 0; a < 3; ++a) b[a] = g2 * f[3 + a];
  const double S = 2. * T / (1. + T * T * (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]));
  for (int a = 0; a < 3; ++a) u[a] += E[a] * T;
  cross(t, u, b);
  for (int a = 0; a < 3; ++a) uo[a] = u[a] + T * t[a];
  cross(t, uo, b);
  for (int a = 0; a < 3; ++a) up[a] = u[a] + S * t[a] + E[a] * T;
  const double g3 = C / sqrt(C * C + up[0] * up[0] + up[1] * up[1] + up[2] * up[2]);
  for (int a = 0; a < 3; ++a) vo[a] = g3 * up[a];
}

// transposed: lvp (adjoint of v') -> lv, lf (E then B), dV of the
// motional field
__device__ inline void boris_T(const BorisK &K, const double f[6], const double v[3],
                               const double lvp[3], double lv[3], double lf[6],
                               double &gV) {
  const double C = K.c, T = K.T;
  const double E[3] = {f[0] + K.eimf_x, f[1] + K.eimf_y, f[2]};
  const double *B = f + 3;
  const double s1 = C * C - (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  const double g1 = s1 > 0 ? C / sqrt(s1) : 0.0;
  double u[3], b[3], um[3], uo[3], uq[3], t[3];
  for (int a = 0; a < 3; ++a) u[a] = g1 * v[a];
  const double s2 = C * C + u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const double g2 = C / sqrt(s2);
  for (int a = 0; a < 3; ++a) b[a] = g2 * B[a];
  const double S = 2. * T / (1. + T * T * (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]));
  for (int a = 0; a < 3; ++a) um[a] = u[a] + E[a] * T;
  cross(t, um, b);
  for (int a = 0; a < 3; ++a) uo[a] = um[a] + T * t[a];
  cross(t, uo, b);
  for (int a = 0; a < 3; ++a) uq[a] = um[a] + S * t[a] + E[a] * T;
  const double s3 = C * C + uq[0] * uq[0] + uq[1] * uq[1] + uq[2] * uq[2];
  const double g3 = C / sqrt(s3);
  double luq[3], lE[3], lum[3], luo[3], lb[3], lu[3], uoxb[3];
  const double d3 = lvp[0] * uq[0] + lvp[1] * uq[1] + lvp[2] * uq[2];
  const double k3 = C / (s3 * sqrt(s3));
  for (int a = 0; a < 3; ++a) luq[a] = g3 * lvp[a] - k3 * d3 * uq[a];
  for (int a = 0; a < 3; ++a) lE[a] = T * luq[a];
  cross(uoxb, uo, b);
  const double lS = luq[0] * uoxb[0] + luq[1] * uoxb[1] + luq[2] * uoxb[2];
  cross(t, b, luq);
  for (int a = 0; a < 3; ++a) {
    lum[a] = luq[a];
    luo[a] = S * t[a];
  }
  cross(t, luq, uo);
  for (int a = 0; a < 3; ++a) lb[a] = S * t[a];
  cross(t, b, luo);
  for (int a = 0; a < 3; ++a) lum[a] += luo[a] + T * t[a];
  cross(t, luo, um);
  for (int a = 0; a < 3; ++a) lb[a] += T * t[a];
  for (int a = 0; a < 3; ++a) lb[a] += lS * (-S * S * T * b[a]);
  for (int a = 0; a < 3; ++a) {
    lu[a] = lum[a];
    lE[a] += T * lum[a];
  }
  const double lg2 = lb[0] * B[0] + lb[1] * B[1] + lb[2] * B[2];
  for (int a = 0; a < 3; ++a) lf[3 + a] = g2 * lb[a];
  const double k2 = C / (s2 * sqrt(s2));
  for (int a = 0; a < 3; ++a) lu[a] -= lg2 * k2 * u[a];
  const double lg1 = lu[0] * v[0] + lu[1] * v[1] + lu[2] * v[2];
  for (int a = 0; a < 3; ++a) lv[a] = g1 * lu[a];
  if (s1 > 0) {
    const double k1 = C / (s1 * sqrt(s1));
    for (int a = 0; a < 3; ++a) lv[a] += lg1 * k1 * v[a];
  }
  for (int a = 0; a < 3; ++a) lf[a] = lE[a];
  gV += lE[0] * K.imf_y - lE[1] * K.imf_x;
}

// FP64 particle SoA of the tape
struct PD {
  double *x, *y, *z, *vx, *vy, *vz;
};

struct PassK {
  double dx, dth, qn, inv3;
  double blo[3], bhi[3];
  BorisK bk;
};

// the particle pass transposed for one species: lp (6 doubles per slot)
// holds the adjoints after the step and returns those before it
// B1 and B0 adjoints += the particles' gather-transpose scatter held in S
__global__ void k_add_gather_T(int64_t n, Lam L) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double d = L.S[a][t];
    L.B[a][t] += d;
    L.B0[a][t] += d;
  }
}

// LDS window of the transposed gather: six FP64 planes of kAdjWin nodes
// (36 KB) around the workgroup's particles
constexpr int kAdjWin = 768;

__global__ __launch_bounds__(256) void k_pass_T(Geom g, PD P, int64_t n, F6 F,
                                                Lam L, int sp, PassK K,
                                                double *lp, double *gV_out) {
  __shared__ double win[6 * kAdjWin];
  __shared__ int bb[6];  // node box of the workgroup's gathers: lo k, l, m, hi k, l, m
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int tid = threadIdx.x, lane = tid & 63;
  const double dx = K.dx, dth = K.dth;
  bool live = i < n && !isnan(P.z[i]);
  double x0[3] = {0, 0, 0}, v0[3] = {0, 0, 0}, q0[3] = {0, 0, 0};
  bool pushed = false;
  if (live) {
    x0[0] = P.x[i];
    x0[1] = P.y[i];
    x0[2] = P.z[i];
    v0[0] = P.vx[i];
    v0[1] = P.vy[i];
    v0[2] = P.vz[i];
    const bool ball = x0[0] < K.blo[0] || x0[0] > K.bhi[0] || x0[1] < K.blo[1] ||
                      x0[1] > K.bhi[1] || x0[2] < K.blo[2] || x0[2] > K.bhi[2];
    for (int a = 0; a < 3; ++a) q0[a] = x0[a] / dx;
    pushed = !ball && in_array_q(g, q0);
  }
  // the nodes the gathers at x0 touch: from the half-staggered base (int)(x -
  // 1/2) to the integer base (int)(x) + 1 on each axis
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = pushed ? (int)(q0[a] - 0.5) : (1 << 29);
    hi[a] = pushed ? (int)q0[a] + 1 : -(1 << 29);
  }
  for (int o = 32; o > 0; o >>= 1)
    for (int a = 0; a < 3; ++a) {
      lo[a] = min(lo[a], __shfl_xor(lo[a], o));
      hi[a] = max(hi[a], __shfl_xor(hi[a], o));
    }
  if (tid < 3) {
    bb[tid] = 1 << 29;
    bb[3 + tid] = -(1 << 29);
  }
  __syncthreads();
  if (lane == 0)
    for (int a = 0; a < 3; ++a) {
      atomicMin(&bb[a], lo[a]);
      atomicMax(&bb[3 + a], hi[a]);
    }
  __syncthreads();
  const int k0 = bb[0], l0 = bb[1], m0 = bb[2];
  const int WX = bb[3] - k0 + 1, WY = bb[4] - l0 + 1, WZ = bb[5] - m0 + 1;
  const bool on = bb[3] >= k0 && (int64_t)WX * WY * WZ <= kAdjWin;
  const int WN = on ? WX * WY * WZ : 0;
  for (int j = tid; j < 6 * WN; j += blockDim.x) win[(j / WN) * kAdjWin + j % WN] = 0.0;
  __syncthreads();
  double gV = 0;
  if (live) {
    double f[6] = {0, 0, 0, 0, 0, 0}, vp[3], x1[3];
    if (pushed) {
      for (int c = 0; c < 6; ++c) {
        Stn s;
        stencil_q(dx, x0, q0, c, s);
        f[c] = st_value(g, F.f[c], s) * K.inv3;
      }
      boris_fwd(K.bk, f, v0, vp);
    } else {
      for (int a = 0; a < 3; ++a) vp[a] = v0[a];
    }
    double q1[3];
    for (int a = 0; a < 3; ++a) {
      x1[a] = x0[a] + vp[a] * dth;
      q1[a] = x1[a] / dx;
    }
    const bool dep = x1[0] < (g.NX + 4) * dx && x1[0] >= 0 && x1[1] < (g.NY + 4) * dx &&
                     x1[1] >= 0 && x1[2] < (g.NZ + 4) * dx && x1[2] >= 0 &&
                     in_array_q(g, q1);
    double *lpi = lp + 6 * i;
    double lx[3] = {lpi[0], lpi[1], lpi[2]};
    double lvp[3] = {lpi[3], lpi[4], lpi[5]};
    for (int a = 0; a < 3; ++a) lvp[a] += lx[a] * dth;
    if (dep) {
      for (int c = 0; c < 3; ++c) {
        Stn s;
        double gr[3];
        stencil_q(dx, x1, q1, c, s);
        const double val = st_gather(g, L.J[sp][c], s, gr);
        lvp[c] += K.qn * val;
        for (int a = 0; a < 3; ++a) lx[a] += K.qn * vp[c] * gr[a];
      }
    }
    for (int a = 0; a < 3; ++a) lvp[a] += lx[a] * dth;
    double lv[3];
    if (pushed) {
      double lf[6];
      boris_T(K.bk, f, v0, lvp, lv, lf, gV);
      for (int c = 0; c < 6; ++c) {
        Stn s;
        double gr[3];
        stencil_q(dx, x0, q0, c, s);
        st_gather(g, F.f[c], s, gr);
        const double w = lf[c] * K.inv3;
        for (int a = 0; a < 3; ++a) lx[a] += w * gr[a];
        // the total B = B0 + B1 is gathered: both adjoints receive the same
        // scatter, collected once in the scratch planes S (added to each
        // after the pass, k_add_gather_T); it goes to the workgroup's LDS
        // window first, each lane walking the 8 nodes in its own order so
        // that lanes of one cell add to different nodes at a time
        if (on) {
          const int bi = s.b[0] - k0, bj = s.b[1] - l0, bk = s.b[2] - m0;
          for (int t = 0; t < 8; ++t) {
            const int nn = t ^ (lane & 7);
            const int a = nn >> 2, b2 = (nn >> 1) & 1, c2 = nn & 1;
            const int j = ((bi + a) * WY + (bj + b2)) * WZ + (bk + c2);
            // selects, not s.w[.][lane-dependent index]: that form puts the
            // stencil in private memory
            const double wa = a ? s.w[0][1] : s.w[0][0];
            const double wb = b2 ? s.w[1][1] : s.w[1][0];
            const double wc = c2 ? s.w[2][1] : s.w[2][0];
            __hip_atomic_fetch_add((PIC_LDS double *)&win[c * kAdjWin + j], w * wa * wb * wc,
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
        } else {
          st_scatter(g, c < 3 ? L.E[c] : L.S[c - 3], s, w);
        }
      }
    } else {
      for (int a = 0; a < 3; ++a) lv[a] = lvp[a];
    }
    for (int a = 0; a < 3; ++a) {
      lpi[a] = lx[a];
      lpi[3 + a] = lv[a];
    }
  }
  if (on) {
    __syncthreads();
    const int WYZ = WY * WZ;
    for (int c = 0; c < 6; ++c) {
      double *dst = c < 3 ? L.E[c] : L.S[c - 3];
      for (int j = tid; j < WN; j += blockDim.x) {
        const double v = win[c * kAdjWin + j];
        if (v == 0.0) continue;
        gadd(dst + g.idx(k0 + j / WYZ, l0 + (j / WZ) % WY, m0 + j % WZ), v);
      }
    }
  }
  // per-block sum of the motional-field V derivative
  __shared__ double red[256];
  red[threadIdx.x] = gV;
  __syncthreads();
  for (int s2 = blockDim.x / 2; s2 > 0; s2 >>= 1) {
    if ((int)threadIdx.x < s2) red[threadIdx.x] += red[threadIdx.x + s2];
    __syncthreads();
  }
  if (threadIdx.x == 0 && red[0] != 0.0) gadd(gV_out, red[0]);
}

// ---------------------------------------------------------------------------
// loss: Maxwell-stress z flux through the sphere (Fibonacci lattice)

struct Sphere {
  double cx, cy, cz, R, w, eps0, mu0, dx, inv3;
  int nq;
};

__device__ inline void sphere_point(const Sphere &S, int q, double P[3], double nrm[3]) {
  const double z = 1.0 - (2.0 * q + 1.0) / S.nq;
  const double r = sqrt(1.0 - z * z);
  const double phi = q * kPi * (3.0 - sqrt(5.0));
  nrm[0] = r * cos(phi);
  nrm[1] = r * sin(phi);
  nrm[2] = z;
  P[0] = S.cx + S.R * nrm[0];
  P[1] = S.cy + S.R * nrm[1];
  P[2] = S.cz + S.R * nrm[2];
}

__global__ __launch_bounds__(256) void k_force(Geom g, F6 F, Sphere S, double *out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0;
  if (q < S.nq) {
    double P[3], n[3], f[6];
    sphere_point(S, q, P, n);
    for (int c = 0; c < 6; ++c) {
      Stn s;
      stencil(S.dx, P, c, s);
      f[c] = st_value(g, F.f[c], s) * S.inv3;
    }
    const double *E = f, *B = f + 3;
    const double En = E[0] * n[0] + E[1] * n[1] + E[2] * n[2];
    const double Bn = B[0] * n[0] + B[1] * n[1] + B[2] * n[2];
    const double E2 = E[0] * E[0] + E[1] * E[1] + E[2] * E[2];
    const double B2 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
    v = S.w * (S.eps0 * (E[2] * En - 0.5 * n[2] * E2) + (B[2] * Bn - 0.5 * n[2] * B2) / S.mu0);
  }
  __shared__ double red[256];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_force_T(Geom g, F6 F, Sphere S, double scale,
                                                 Lam L) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= S.nq) return;
  double P[3], n[3], f[6];
  sphere_point(S, q, P, n);
  Stn st[6];
  for (int c = 0; c < 6; ++c) {
    stencil(S.dx, P, c, st[c]);
    f[c] = st_value(g, F.f[c], st[c]) * S.inv3;
  }
  const double *E = f, *B = f + 3;
  const double En = E[0] * n[0] + E[1] * n[1] + E[2] * n[2];
  const double Bn = B[0] * n[0] + B[1] * n[1] + B[2] * n[2];
  const double w = S.w * scale;
  for (int a = 0; a < 3; ++a) {
    const double le = w * S.eps0 * ((a == 2 ? En : 0.0) + E[2] * n[a] - n[2] * E[a]);
    const double lb = w / S.mu0 * ((a == 2 ? Bn : 0.0) + B[2] * n[a] - n[2] * B[a]);
    st_scatter(g, L.E[a], st[a], le * S.inv3);
    st_scatter(g, L.B[a], st[3 + a], lb * S.inv3);
    st_scatter(g, L.B0[a], st[3 + a], lb * S.inv3);
  }
}

// ---------------------------------------------------------------------------
// conversions into the FP64 tape

// FP64 form: E and B1 = B - B0; FP32 form: E and B1 as stored
template <typename T>
__global__ void k_fields_to_d(int64_t n, const T *ex, const T *ey, const T *ez,
                              const T *bx, const T *by, const T *bz,
                              const T *b0x, const T *b0y, const T *b0z, int b_total,
                              double *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = ex[i];
  out[n + i] = ey[i];
  out[2 * n + i] = ez[i];
  if (b_total) {
    out[3 * n + i] = (double)bx[i] - (double)b0x[i];
    out[4 * n + i] = (double)by[i] - (double)b0y[i];
    out[5 * n + i] = (double)bz[i] - (double)b0z[i];
  } else {
    out[3 * n + i] = bx[i];
    out[4 * n + i] = by[i];
    out[5 * n + i] = bz[i];
  }
}

// F6 (E, total B) from a tape entry (E, B1) and B0
__global__ void k_tape_to_f6(int64_t n, const double *tp, const double *b0, F6 F) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < 3; ++c) {
    F.f[c][i] = tp[c * n + i];
    F.f[3 + c][i] = b0[c * n + i] + tp[(3 + c) * n + i];
  }
}

// FP32 particles into FP64 reference coordinates; dead -> z = NaN
__global__ void k_p32_to_d(Particles32 Q, int64_t n, double dx, PD P) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t c = Q.cell[i];
  if (c < 0) {
    P.x[i] = P.y[i] = 0.0;
    P.z[i] = __longlong_as_double(0x7ff8000000000000ll);
    P.vx[i] = P.vy[i] = P.vz[i] = 0.0;
    return;
  }
  const int k = c >> 20, l = (c >> 10) & 1023, m = c & 1023;
  P.x[i] = ((double)k + (double)Q.ox[i]) * dx;
  P.y[i] = ((double)l + (double)Q.oy[i]) * dx;
  P.z[i] = ((double)m + (double)Q.oz[i]) * dx;
  P.vx[i] = Q.vx[i];
  P.vy[i] = Q.vy[i];
  P.vz[i] = Q.vz[i];
}

// per-block sums: out[b] = sum over [lo, hi) of lp[6 i + 5] (v_z adjoint),
// the slots zeroed afterwards when zero_after is set; alive[] filter
__global__ __launch_bounds__(256) void k_sum_vz(double *lp, int64_t lo, int64_t hi,
                                                const double *z, int zero_after,
                                                double *out) {
  double v = 0;
  for (int64_t i = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < hi;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (z && isnan(z[i])) continue;
    v += lp[6 * i + 5];
    if (zero_after)
      for (int a = 0; a < 6; ++a) lp[6 * i + a] = 0.0;
  }
  __shared__ double red[256];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// the particle adjoints across a cell sort, backwards: the particle at slot
// i after the sort sat at slot perm[i] before it; the slots of particles
// the sort dropped (dead ones) keep the zero they were given
__global__ void k_unsort_lp(int64_t na, const uint32_t *__restrict__ perm,
                            const double *__restrict__ src, double *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= na) return;
  const int64_t j = perm[i];
  for (int a = 0; a < 6; ++a) dst[6 * j + a] = src[6 * i + a];
}

// out[b * 3 + a] = per-block sums of lB0 . G_a
__global__ __launch_bounds__(256) void k_dot_g(int64_t n, Lam L, const double *G,
                                               double *out) {
  double v[3] = {0, 0, 0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) v[a] += L.B0[c][i] * G[(3 * a + c) * n + i];
  __shared__ double red[3][256];
  for (int a = 0; a < 3; ++a) red[a][threadIdx.x] = v[a];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int a = 0; a < 3; ++a) red[a][threadIdx.x] += red[a][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 3) out[blockIdx.x * 3 + threadIdx.x] = red[threadIdx.x][0];
}

// ---------------------------------------------------------------------------
// host side

struct Ckpt {
  void *fields = nullptr;
  void *parts[2] = {nullptr, nullptr};
  uint32_t *tiles[2] = {nullptr, nullptr};
  int64_t n[2], sorted_end[2];
  int cur[2];
  int32_t drift_halves[2];
  int64_t step;
  int32_t c_add, c_sort;
  uint32_t inflow_event;
};

// One recorded step.  A step that began with the cell sort keeps, per
// species, the count before it (nb) and after it (na); its permutation
// (slot after = perm index, slot before = perm value) goes to the caller's
// buffers.
struct TapeStep {
  int inflow, sorted;
  int64_t n[2], lo[2], hi[2], nb[2], na[2];
};

struct AdjState {
  double p[4];
  int nsteps = 0, ckpt = 0, nq = 0;
  double R = 0, loss = 0;
  bool have_forward = false;
  double *G = nullptr;   // 9 planes: unit-moment B0 (device)
  double *b0 = nullptr;  // 3 planes: B0 of the forward (device)
  double *lam = nullptr; // 18 planes of Lam (device)
  double *f6 = nullptr;  // 6 planes: the F6 scratch (device)
  double *tmp = nullptr; // 6 planes: field conversion (device)
  double *red = nullptr; // reduction partials (device, 4096)
  std::vector<Ckpt> ck;
};

AdjState *state(pic_ctx *ctx) { return (AdjState *)ctx->adj; }

int64_t ncell(const pic_ctx *ctx) {
  return (int64_t)ctx->geom.SX * ctx->geom.SY * ctx->geom.SZ;
}

Lam lam_view(pic_ctx *ctx, AdjState *A) {
  const int64_t n = ncell(ctx);
  double *b = A->lam;
  Lam L;
  for (int a = 0; a < 3; ++a) {
    L.E[a] = b + (0 + a) * n;
    L.B[a] = b + (3 + a) * n;
    L.B0[a] = b + (6 + a) * n;
    L.J[0][a] = b + (9 + a) * n;
    L.J[1][a] = b + (12 + a) * n;
    L.S[a] = b + (15 + a) * n;
  }
  return L;
}

F6 f6_view(pic_ctx *ctx, double *base) {
  const int64_t n = ncell(ctx);
  F6 F;
  for (int c = 0; c < 6; ++c) F.f[c] = base + c * n;
  return F;
}

Sphere sphere(pic_ctx *ctx, AdjState *A) {
  const pic_config &c = ctx->cfg;
  const Geom &g = ctx->geom;
  Sphere S;
  S.dx = c.dx;
  S.cx = ((g.NX + 4) / 2) * c.dx;
  S.cy = ((g.NY + 4) / 2) * c.dx;
  S.cz = ((g.NZ + 4) / 2) * c.dx;
  S.R = A->R * c.dx;
  S.nq = A->nq;
  S.w = 4.0 * kPi * S.R * S.R / A->nq;
  S.eps0 = c.eps0;
  S.mu0 = c.mu0;
  S.inv3 = 1.0 / (c.dx * c.dx * c.dx);
  return S;
}

unsigned nb(int64_t n) { return (unsigned)div_up(n, kBlock); }

// unit-moment B0 planes on the host (the forward's dipole formula with a
// general moment vector, the softened core and the rmax cutoff)
void unit_fields_host(const pic_config &c, const Geom &g, double *G) {
  const int64_t nc = (int64_t)g.SX * g.SY * g.SZ;
  const double dx = c.dx;
  const int cx = (g.NX + 4) / 2, cy = (g.NY + 4) / 2, cz = (g.NZ + 4) / 2;
  auto field = [&](double x, double y, double z, int a, int comp) {
    double r = sqrt(x * x + y * y + z * z);
    if (c.dipole_soft > 0.0 && r < c.dipole_soft * dx) {
      const double s = c.dipole_soft * dx / r;
      x *= s;
      y *= s;
      z *= s;
      r = sqrt(x * x + y * y + z * z);
    }
    const double rv[3] = {x, y, z};
    const double mr = rv[a];
    return c.mu0 * (3. * mr * rv[comp] - r * r * (a == comp ? 1.0 : 0.0)) /
           (4. * kPi * pow(r, 5));
  };
  for (int k = 0; k < g.SX; ++k)
    for (int l = 0; l < g.SY; ++l)
      for (int m = 0; m < g.SZ; ++m) {
        const int64_t i = g.idx(k, l, m);
        const double x = (k - cx) * dx, y = (l - cy) * dx, z = (m - cz) * dx;
        const bool in = sqrt(x * x + y * y + z * z) < c.dipole_rmax * dx;
        for (int a = 0; a < 3; ++a) {
          G[(3 * a + 0) * nc + i] = in ? field((k - cx + 0.5) * dx, (l - cy) * dx, (m - cz) * dx, a, 0) : 0.0;
          G[(3 * a + 1) * nc + i] = in ? field((k - cx) * dx, (l - cy + 0.5) * dx, (m - cz) * dx, a, 1) : 0.0;
          G[(3 * a + 2) * nc + i] = in ? field((k - cx) * dx, (l - cy) * dx, (m - cz + 0.5) * dx, a, 2) : 0.0;
        }
      }
}

int ensure_state(pic_ctx *ctx) {
  if (ctx->adj) return 0;
  AdjState *A = new (std::nothrow) AdjState();
  if (!A) return set_error(ctx, PIC_E_NOMEM, "adjoint state");
  ctx->adj = A;
  const int64_t n = ncell(ctx);
  PIC_HIP(ctx, hipMalloc(&A->G, 9 * n * sizeof(double)));
  PIC_HIP(ctx, hipMalloc(&A->b0, 3 * n * sizeof(double)));
  PIC_HIP(ctx, hipMalloc(&A->lam, 18 * n * sizeof(double)));
  PIC_HIP(ctx, hipMalloc(&A->f6, 6 * n * sizeof(double)));
  PIC_HIP(ctx, hipMalloc(&A->tmp, 6 * n * sizeof(double)));
  PIC_HIP(ctx, hipMalloc(&A->red, 4096 * sizeof(double)));
  double *h = (double *)malloc(9 * n * sizeof(double));
  if (!h) return set_error(ctx, PIC_E_NOMEM, "adjoint unit fields");
  unit_fields_host(ctx->cfg, ctx->geom, h);
  hipError_t e = hipMemcpy(A->G, h, 9 * n * sizeof(double), hipMemcpyHostToDevice);
  free(h);
  PIC_HIP(ctx, e);
  return 0;
}

void free_ckpts(AdjState *A) {
  for (Ckpt &c : A->ck) {
    if (c.fields) (void)hipFree(c.fields);
    for (int s = 0; s < 2; ++s) {
      if (c.parts[s]) (void)hipFree(c.parts[s]);
      if (c.tiles[s]) (void)hipFree(c.tiles[s]);
    }
  }
  A->ck.clear();
}

size_t slot_bytes(const pic_ctx *ctx) {
  return ctx->cfg.precision == PIC_FP64 ? 6 * sizeof(double)
                                        : sizeof(int32_t) + 6 * sizeof(float);
}

// the 7 (FP32) or 6 (FP64) sub-arrays of a particle buffer: base and
// element size
void sub_arrays(const pic_ctx *ctx, void *buf, int64_t cap, char **ptr, size_t *esz,
                int *count) {
  char *b = (char *)buf;
  if (ctx->cfg.precision == PIC_FP64) {
    *count = 6;
    for (int j = 0; j < 6; ++j) {
      ptr[j] = b + (size_t)j * cap * sizeof(double);
      esz[j] = sizeof(double);
    }
  } else {
    *count = 7;
    ptr[0] = b;
    esz[0] = sizeof(int32_t);
    for (int j = 0; j < 6; ++j) {
      ptr[1 + j] = b + (size_t)4 * cap + (size_t)j * cap * sizeof(float);
      esz[1 + j] = sizeof(float);
    }
  }
}

int save_ckpt(pic_ctx *ctx, Ckpt &c) {
  PIC_HIP(ctx, hipMalloc(&c.fields, ctx->field_bytes));
  PIC_HIP(ctx, hipMemcpyAsync(c.fields, ctx->field_mem, ctx->field_bytes,
                              hipMemcpyDeviceToDevice, ctx->stream));
  const int nt = tile_geom(ctx).ntiles();
  for (int s = 0; s < 2; ++s) {
    c.n[s] = ctx->n[s];
    c.sorted_end[s] = ctx->sorted_end[s];
    c.cur[s] = ctx->cur[s];
    c.drift_halves[s] = ctx->drift_halves[s];
    const int64_t n = ctx->n[s];
    PIC_HIP(ctx, hipMalloc(&c.parts[s], (size_t)(n + 1) * slot_bytes(ctx)));
    char *src[7], *dst[7];
    size_t esz[7];
    int cnt;
    sub_arrays(ctx, ctx->pbuf[s][ctx->cur[s]], ctx->cap, src, esz, &cnt);
    sub_arrays(ctx, c.parts[s], n + 1, dst, esz, &cnt);
    for (int j = 0; j < cnt; ++j)
      if (n > 0)
        PIC_HIP(ctx, hipMemcpyAsync(dst[j], src[j], n * esz[j], hipMemcpyDeviceToDevice,
                                    ctx->stream));
    PIC_HIP(ctx, hipMalloc(&c.tiles[s], (size_t)(nt + 1) * sizeof(uint32_t)));
    PIC_HIP(ctx, hipMemcpyAsync(c.tiles[s], ctx->tile_start[s], (size_t)(nt + 1) * sizeof(uint32_t),
                                hipMemcpyDeviceToDevice, ctx->stream));
  }
  c.step = ctx->step;
  c.c_add = ctx->c_add;
  c.c_sort = ctx->c_sort;
  c.inflow_event = ctx->inflow_event;
  PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int restore_ckpt(pic_ctx *ctx, const Ckpt &c) {
  PIC_HIP(ctx, hipMemcpyAsync(ctx->field_mem, c.fields, ctx->field_bytes,
                              hipMemcpyDeviceToDevice, ctx->stream));
  const int nt = tile_geom(ctx).ntiles();
  for (int s = 0; s < 2; ++s) {
    ctx->n[s] = c.n[s];
    ctx->sorted_end[s] = c.sorted_end[s];
    ctx->cur[s] = c.cur[s];
    ctx->drift_halves[s] = c.drift_halves[s];
    const int64_t n = c.n[s];
    char *src[7], *dst[7];
    size_t esz[7];
    int cnt;
    sub_arrays(ctx, c.parts[s], n + 1, src, esz, &cnt);
    sub_arrays(ctx, ctx->pbuf[s][c.cur[s]], ctx->cap, dst, esz, &cnt);
    for (int j = 0; j < cnt; ++j)
      if (n > 0)
        PIC_HIP(ctx, hipMemcpyAsync(dst[j], src[j], n * esz[j], hipMemcpyDeviceToDevice,
                                    ctx->stream));
    PIC_HIP(ctx, hipMemcpyAsync(ctx->tile_start[s], c.tiles[s], (size_t)(nt + 1) * sizeof(uint32_t),
                                hipMemcpyDeviceToDevice, ctx->stream));
  }
  ctx->step = c.step;
  ctx->c_add = c.c_add;
  ctx->c_sort = c.c_sort;
  ctx->inflow_event = c.inflow_event;
  if (ctx->cfg.precision == PIC_FP32) {
    int rc = f32::refresh_total_b(ctx);
    if (rc) return rc;
  }
  PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

// E and B1 of the live state into `out` (6 planes, FP64)
int fields_to_d(pic_ctx *ctx, double *out) {
  const int64_t n = ncell(ctx);
  const Geom &g = ctx->geom;
  if (ctx->cfg.precision == PIC_FP64) {
    const double *b = (const double *)ctx->field_mem;
    k_fields_to_d<double><<<nb(n), kBlock, 0, ctx->stream>>>(
        n, b + PIC_EX * g.plane, b + PIC_EY * g.plane, b + PIC_EZ * g.plane,
        b + PIC_BX * g.plane, b + PIC_BY * g.plane, b + PIC_BZ * g.plane,
        b + PIC_B0X * g.plane, b + PIC_B0Y * g.plane, b + PIC_B0Z * g.plane, 1, out);
  } else {
    const float *b = (const float *)ctx->field_mem;
    k_fields_to_d<float><<<nb(n), kBlock, 0, ctx->stream>>>(
        n, b + PIC_EX * g.plane, b + PIC_EY * g.plane, b + PIC_EZ * g.plane,
        b + PIC_BX * g.plane, b + PIC_BY * g.plane, b + PIC_BZ * g.plane,
        b + PIC_B0X * g.plane, b + PIC_B0Y * g.plane, b + PIC_B0Z * g.plane, 0, out);
  }
  PIC_LAUNCH_CHECK(ctx);
  return 0;
}

// the loss term of the live state
int force_now(pic_ctx *ctx, AdjState *A, double *F) {
  int rc = fields_to_d(ctx, A->tmp);
  if (rc) return rc;
  const int64_t n = ncell(ctx);
  k_tape_to_f6<<<nb(n), kBlock, 0, ctx->stream>>>(n, A->tmp, A->b0, f6_view(ctx, A->f6));
  const unsigned blocks = nb(A->nq);
  k_force<<<blocks, kBlock, 0, ctx->stream>>>(ctx->geom, f6_view(ctx, A->f6),
                                              sphere(ctx, A), A->red);
  PIC_LAUNCH_CHECK(ctx);
  std::vector<double> h(blocks);
  PIC_HIP(ctx, hipMemcpyAsync(h.data(), A->red, blocks * sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
  PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double s = 0;
  for (double v : h) s += v;
  *F = s;
  return 0;
}

// one forward step: the cell sort at the step loop's cadence (:7475-7590,
// its permutation into perm[s] when given), the inflow appended to the
// store (its ranges recorded), and, when tape_fields is given, the tape
// entry (after the inflow, before the first half step)
int fwd_step(pic_ctx *ctx, TapeStep &ts, double *tape_fields, PD tape_p[2],
             uint32_t *const perm[2]) {
  const pic_config &c = ctx->cfg;
  int rc;
  if ((rc = pic_renew_grid(ctx))) return rc;
  ts.sorted = 0;
  if (ctx->c_sort > c.sort_every - 1) {
    ctx->c_sort = 0;
    for (int s = 0; s < 2; ++s) {
      ts.nb[s] = ctx->n[s];
      if ((rc = pic_sort(ctx, s))) return rc;
      ts.na[s] = ctx->n[s];
      if (perm && perm[s] && ts.na[s] > 0)
        PIC_HIP(ctx, hipMemcpyAsync(perm[s], ctx->idx, (size_t)ts.na[s] * sizeof(uint32_t),
                                    hipMemcpyDeviceToDevice, ctx->stream));
    }
    ts.sorted = 1;
  }
  ts.inflow = 0;
  if (ctx->c_add > c.inflow_every - 1) {
    const int nface = c.refill_far ? 2 : 1;
    const int64_t add = (int64_t)c.nx * c.ny * 2 * c.ppc * nface;
    for (int s = 0; s < 2; ++s)
      if (ctx->n[s] + add > ctx->cap)
        return set_error(ctx, PIC_E_CAPACITY,
                         "adjoint: inflow would compact the particle store; "
                         "create the context with capacity >= %lld",
                         (long long)(ctx->n[s] + add));
    for (int s = 0; s < 2; ++s) ts.lo[s] = ctx->n[s];
    if ((rc = pic_inflow(ctx))) return rc;
    for (int s = 0; s < 2; ++s) ts.hi[s] = ctx->n[s];
    ctx->c_add = 0;
    ts.inflow = 1;
  }
  for (int s = 0; s < 2; ++s) ts.n[s] = ctx->n[s];
  if (tape_fields) {
    if ((rc = fields_to_d(ctx, tape_fields))) return rc;
    for (int s = 0; s < 2; ++s) {
      const int64_t n = ctx->n[s];
      if (n == 0) continue;
      if (c.precision == PIC_FP64) {
        const Particles64 P = p64(ctx, s, ctx->cur[s]);
        const double *src[6] = {P.x, P.y, P.z, P.vx, P.vy, P.vz};
        double *dst[6] = {tape_p[s].x, tape_p[s].y, tape_p[s].z, tape_p[s].vx,
                          tape_p[s].vy, tape_p[s].vz};
        for (int j = 0; j < 6; ++j)
          PIC_HIP(ctx, hipMemcpyAsync(dst[j], src[j], n * sizeof(double),
                                      hipMemcpyDeviceToDevice, ctx->stream));
      } else {
        k_p32_to_d<<<nb(n), kBlock, 0, ctx->stream>>>(p32(ctx, s, ctx->cur[s]), n, c.dx,
                                                      tape_p[s]);
        PIC_LAUNCH_CHECK(ctx);
      }
    }
  }
  if ((rc = pic_renew_b_half(ctx))) return rc;
  if ((rc = pic_halo_b(ctx))) return rc;
  for (int s = 0; s < 2; ++s)
    if ((rc = push_fused(ctx, s))) return rc;
  if ((rc = pic_renew_b_half(ctx))) return rc;
  if ((rc = pic_halo_b(ctx))) return rc;
  if ((rc = pic_fold_current(ctx))) return rc;
  if ((rc = pic_renew_e(ctx))) return rc;
  if ((rc = pic_halo_e(ctx))) return rc;
  ctx->c_sort++;
  ctx->c_add++;
  ctx->step++;
  return 0;
}

double sum_partials(pic_ctx *ctx, const double *dev, int count, int stride, int comp,
                    int *rc) {
  std::vector<double> h((size_t)count * stride);
  hipError_t e = hipMemcpyAsync(h.data(), dev, h.size() * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    *rc = set_error(ctx, PIC_E_HIP, "adjoint reduction: %s", hipGetErrorString(e));
    return 0;
  }
  double s = 0;
  for (int b = 0; b < count; ++b) s += h[(size_t)b * stride + comp];
  return s;
}

}  // namespace

void adj_free(pic_ctx *ctx) {
  AdjState *A = state(ctx);
  if (!A) return;
  free_ckpts(A);
  if (A->G) (void)hipFree(A->G);
  if (A->b0) (void)hipFree(A->b0);
  if (A->lam) (void)hipFree(A->lam);
  if (A->f6) (void)hipFree(A->f6);
  if (A->tmp) (void)hipFree(A->tmp);
  if (A->red) (void)hipFree(A->red);
  delete A;
  ctx->adj = nullptr;
}

}  // namespace pic

using namespace pic;

extern "C" int pic_adj_forward(pic_ctx *ctx, const double params[4], int nsteps,
                               int ckpt_every, double radius, int nquad,
                               double *loss) {
  if (!ctx || !params || !loss || nsteps <= 0 || ckpt_every <= 0 || nquad <= 0 ||
      nquad > 256 * kBlock || !(radius > 0))
    return PIC_E_ARG;
  // the sphere's gather stencils reach centre +- (R + 1) nodes: R <= centre - 2
  const Geom &g = ctx->geom;
  const int cmin = std::min(std::min(g.NX, g.NY), g.NZ) / 2 + 2;
  if (radius > cmin - 2)
    return set_error(ctx, PIC_E_ARG, "adjoint: sphere radius %g > %d cells", radius,
                     cmin - 2);
  int rc = ctx->cfg.precision == PIC_FP64 ? 0 : f32::settle_sorts(ctx);
  if (rc) return rc;
  rc = ensure_state(ctx);
  if (rc) return rc;
  AdjState *A = state(ctx);
  free_ckpts(A);
  for (int a = 0; a < 4; ++a) A->p[a] = params[a];
  A->nsteps = nsteps;
  A->ckpt = ckpt_every;
  A->R = radius;
  A->nq = nquad;
  A->have_forward = false;
  // parameters: B0 = sum_a M_a G_a + IMF, B = B0, E = 0; V
  const pic_config &c = ctx->cfg;
  const int64_t n = ncell(ctx);
  std::vector<double> G(9 * n), b0(3 * n), zero(n, 0.0);
  PIC_HIP(ctx, hipMemcpy(G.data(), A->G, 9 * n * sizeof(double), hipMemcpyDeviceToHost));
  for (int comp = 0; comp < 3; ++comp)
    for (int64_t i = 0; i < n; ++i)
      b0[comp * n + i] = params[0] * G[(0 + comp) * n + i] + params[1] * G[(3 + comp) * n + i] +
                         params[2] * G[(6 + comp) * n + i] + c.imf[comp];
  PIC_HIP(ctx, hipMemcpy(A->b0, b0.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice));
  for (int comp = 0; comp < 3; ++comp)
    if ((rc = pic_upload_field(ctx, PIC_B0X + comp, b0.data() + comp * n))) return rc;
  for (int comp = 0; comp < 3; ++comp)
    if ((rc = pic_upload_field(ctx, PIC_BX + comp, b0.data() + comp * n))) return rc;
  for (int comp = 0; comp < 3; ++comp)
    if ((rc = pic_upload_field(ctx, PIC_EX + comp, zero.data()))) return rc;
  ctx->cfg.vdrift[2] = params[3];
  ctx->adj_mode = 1;
  ctx->step = 0;
  ctx->c_add = 0;
  ctx->c_sort = 0;
  ctx->inflow_event = 0;
  if ((rc = pic_load_particles(ctx))) return rc;
  if ((rc = pic_begin(ctx))) return rc;
  double L = 0;
  TapeStep ts;
  for (int s = 0; s < nsteps; ++s) {
    if (s % ckpt_every == 0) {
      A->ck.emplace_back();
      if ((rc = save_ckpt(ctx, A->ck.back()))) return rc;
    }
    if ((rc = fwd_step(ctx, ts, nullptr, nullptr, nullptr))) return rc;
    double F;
    if ((rc = force_now(ctx, A, &F))) return rc;
    L += F / nsteps;
  }
  ctx->adj_mode = 0;
  A->loss = L;
  A->have_forward = true;
  *loss = L;
  return 0;
}

extern "C" int pic_adj_backward(pic_ctx *ctx, double dloss, double grad[4]) {
  if (!ctx || !grad) return PIC_E_ARG;
  AdjState *A = state(ctx);
  if (!A || !A->have_forward)
    return set_error(ctx, PIC_E_STATE, "adjoint: pic_adj_forward first");
  const pic_config &c = ctx->cfg;
  const Geom &g = ctx->geom;
  const int64_t n = ncell(ctx);
  const int N = A->nsteps, K = A->ckpt;
  int rc;
  hipStream_t st = ctx->stream;
  PIC_HIP(ctx, hipMemsetAsync(A->lam, 0, 18 * n * sizeof(double), st));
  // particle adjoints (and their copy across a sort), the tape of one
  // segment and the permutations of the sorts in it
  double *lp[2] = {nullptr, nullptr}, *lq[2] = {nullptr, nullptr};
  std::vector<double *> tf(K + 1, nullptr);
  std::vector<double *> tpart(K, nullptr);
  std::vector<uint32_t *> perm(2 * (size_t)K, nullptr);
  std::vector<TapeStep> ts(K);
  const int64_t cap = ctx->cap;
  auto cleanup = [&]() {
    for (int s = 0; s < 2; ++s) {
      if (lp[s]) (void)hipFree(lp[s]);
      if (lq[s]) (void)hipFree(lq[s]);
    }
    for (uint32_t *p : perm)
      if (p) (void)hipFree(p);
    for (double *p : tf)
      if (p) (void)hipFree(p);
    for (double *p : tpart)
      if (p) (void)hipFree(p);
    ctx->adj_mode = 0;
  };
#define TRY(x)          \
  do {                  \
    if ((rc = (x))) {   \
      cleanup();        \
      return rc;        \
    }                   \
  } while (0)
#define TRYH(x) TRY((x) == hipSuccess ? 0 : set_error(ctx, PIC_E_HIP, "adjoint: %s", #x))
  for (int s = 0; s < 2; ++s) {
    TRYH(hipMalloc(&lp[s], 6 * (size_t)cap * sizeof(double)));
    TRYH(hipMemsetAsync(lp[s], 0, 6 * (size_t)cap * sizeof(double), st));
    TRYH(hipMalloc(&lq[s], 6 * (size_t)cap * sizeof(double)));
  }
  for (int j = 0; j <= K; ++j) TRYH(hipMalloc(&tf[j], 6 * n * sizeof(double)));
  for (int j = 0; j < K; ++j) TRYH(hipMalloc(&tpart[j], 12 * (size_t)cap * sizeof(double)));
  auto tape_pd = [&](int j, int s) {
    double *b = tpart[j] + (size_t)s * 6 * cap;
    return PD{b, b + cap, b + 2 * cap, b + 3 * cap, b + 4 * cap, b + 5 * cap};
  };
  const Lam L = lam_view(ctx, A);
  const SigmaProfile sg = sigma(ctx);
  const double tempB = 1. / c.eps0 * c.dt * 0.5 * 0.5, dth = 0.5 * c.dt;
  const double tempE = c.c * c.c * c.dt, sc = 1. / (24. * c.dx);
  const Sphere S = sphere(ctx, A);
  const F6 F = f6_view(ctx, A->f6);
  double gV = 0;
  double *gvd = A->red + 4000;  // device accumulator of the motional-field term
  TRYH(hipMemsetAsync(gvd, 0, sizeof(double), st));
  ctx->adj_mode = 1;
  const int nseg = (int)A->ck.size();
  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int s0 = seg * K, s1 = std::min(N, s0 + K);
    TRY(restore_ckpt(ctx, A->ck[seg]));
    for (int j = 0; j < s1 - s0; ++j) {
      PD tp[2] = {tape_pd(j, 0), tape_pd(j, 1)};
      uint32_t *pm[2] = {nullptr, nullptr};
      if (ctx->c_sort > c.sort_every - 1)  // this step begins with the sort
        for (int s = 0; s < 2; ++s) {
          uint32_t *&b = perm[2 * (size_t)j + s];
          if (!b) TRYH(hipMalloc(&b, (size_t)cap * sizeof(uint32_t)));
          pm[s] = b;
        }
      TRY(fwd_step(ctx, ts[j], tf[j], tp, pm));
    }
    TRY(fields_to_d(ctx, tf[s1 - s0]));
    for (int j = s1 - s0 - 1; j >= 0; --j) {
      // loss term of the state after the step
      k_tape_to_f6<<<nb(n), kBlock, 0, st>>>(n, tf[j + 1], A->b0, F);
      k_force_T<<<nb(A->nq), kBlock, 0, st>>>(g, F, S, dloss / N, L);
      // sr_e, renew_e4_2, sr_current2, sr_b, second renew_b4_2
      k_halo_T_y<<<nb((int64_t)g.SX * g.SZ), kBlock, 0, st>>>(g, L.E[0], L.E[1], L.E[2],
                                                              nullptr, nullptr, nullptr);
      k_halo_T_x<<<nb((int64_t)g.SY * g.SZ), kBlock, 0, st>>>(g, L.E[0], L.E[1], L.E[2],
                                                              nullptr, nullptr, nullptr);
      k_re_T1<<<nb(n), kBlock, 0, st>>>(g, L, sg, tempE, c.mu0, sc);
      k_re_T2<<<nb(n), kBlock, 0, st>>>(g, L);
      for (int sp = 0; sp < 2; ++sp) {
        k_fold_T_y<<<nb((int64_t)g.SX * g.SZ), kBlock, 0, st>>>(g, L.J[sp][0], L.J[sp][1],
                                                                L.J[sp][2]);
        k_fold_T_x<<<nb((int64_t)g.SY * g.SZ), kBlock, 0, st>>>(g, L.J[sp][0], L.J[sp][1],
                                                                L.J[sp][2]);
      }
      k_halo_T_y<<<nb((int64_t)g.SX * g.SZ), kBlock, 0, st>>>(g, L.B[0], L.B[1], L.B[2],
                                                              L.B0[0], L.B0[1], L.B0[2]);
      k_halo_T_x<<<nb((int64_t)g.SY * g.SZ), kBlock, 0, st>>>(g, L.B[0], L.B[1], L.B[2],
                                                              L.B0[0], L.B0[1], L.B0[2]);
      k_rb_T1<<<nb(n), kBlock, 0, st>>>(g, L, sg, tempB, dth, sc);
      k_rb_T2<<<nb(n), kBlock, 0, st>>>(g, L);
      TRY(hipGetLastError() == hipSuccess ? 0 : set_error(ctx, PIC_E_HIP, "adjoint field kernels"));
      // the fields the particle pass read: first half step and sr_b
      k_tape_to_f6<<<nb(n), kBlock, 0, st>>>(n, tf[j], A->b0, F);
      k_rb_fwd<<<nb(n), kBlock, 0, st>>>(g, F, A->b0, A->b0 + n, A->b0 + 2 * n, sg,
                                         tempB, dth, sc);
      k_halo_x_fwd<<<nb((int64_t)g.SY * g.SZ), kBlock, 0, st>>>(g, F.f[3], F.f[4], F.f[5]);
      k_halo_y_fwd<<<nb((int64_t)g.SX * g.SZ), kBlock, 0, st>>>(g, F.f[3], F.f[4], F.f[5]);
      // S (free between the transposed E and B updates) collects the B part
      // of the transposed gather
      TRYH(hipMemsetAsync(L.S[0], 0, 3 * n * sizeof(double), st));
      for (int sp = 0; sp < 2; ++sp) {
        const int64_t np = ts[j].n[sp];
        if (np == 0) continue;
        PassK pk;
        pk.dx = c.dx;
        pk.dth = dth;
        pk.inv3 = 1.0 / (c.dx * c.dx * c.dx);
        pk.qn = c.q * c.weight[sp] * pk.inv3;
        for (int a = 0; a < 3; ++a) {
          pk.blo[a] = c.ballistic_lo[a];
          pk.bhi[a] = c.ballistic_hi[a];
        }
        pk.bk.c = c.c;
        pk.bk.T = 0.5 * c.q * c.dt / (sp == PIC_ION ? c.mass[0] : -c.mass[1]);
        pk.bk.eimf_x = A->p[3] * c.imf[1];
        pk.bk.eimf_y = -A->p[3] * c.imf[0];
        pk.bk.imf_x = c.imf[0];
        pk.bk.imf_y = c.imf[1];
        k_pass_T<<<nb(np), 256, 0, st>>>(g, tape_pd(j, sp), np, F, L, sp, pk, lp[sp], gvd);
      }
      k_add_gather_T<<<nb(n), kBlock, 0, st>>>(n, L);
      TRYH(hipMemsetAsync(L.J[0][0], 0, 6 * n * sizeof(double), st));
      // sr_b and the first renew_b4_2
      k_halo_T_y<<<nb((int64_t)g.SX * g.SZ), kBlock, 0, st>>>(g, L.B[0], L.B[1], L.B[2],
                                                              L.B0[0], L.B0[1], L.B0[2]);
      k_halo_T_x<<<nb((int64_t)g.SY * g.SZ), kBlock, 0, st>>>(g, L.B[0], L.B[1], L.B[2],
                                                              L.B0[0], L.B0[1], L.B0[2]);
      k_rb_T1<<<nb(n), kBlock, 0, st>>>(g, L, sg, tempB, dth, sc);
      k_rb_T2<<<nb(n), kBlock, 0, st>>>(g, L);
      TRY(hipGetLastError() == hipSuccess ? 0 : set_error(ctx, PIC_E_HIP, "adjoint pass kernels"));
      // particles injected at the start of this step: v_z = V + sigma N
      if (ts[j].inflow)
        for (int sp = 0; sp < 2; ++sp) {
          const int64_t cnt = ts[j].hi[sp] - ts[j].lo[sp];
          if (cnt <= 0) continue;
          const int blocks = (int)std::min<int64_t>(1024, div_up(cnt, kBlock));
          k_sum_vz<<<blocks, kBlock, 0, st>>>(lp[sp], ts[j].lo[sp], ts[j].hi[sp], nullptr, 1,
                                              A->red);
          gV += sum_partials(ctx, A->red, blocks, 1, 0, &rc);
          if (rc) {
            cleanup();
            return rc;
          }
        }
      // back across the sort this step began with: the adjoints to the
      // slots the particles held before it
      if (ts[j].sorted)
        for (int sp = 0; sp < 2; ++sp) {
          TRYH(hipMemsetAsync(lq[sp], 0, 6 * (size_t)ts[j].nb[sp] * sizeof(double), st));
          if (ts[j].na[sp] > 0)
            k_unsort_lp<<<nb(ts[j].na[sp]), kBlock, 0, st>>>(
                ts[j].na[sp], perm[2 * (size_t)j + sp], lp[sp], lq[sp]);
          std::swap(lp[sp], lq[sp]);
        }
      TRY(hipGetLastError() == hipSuccess ? 0 : set_error(ctx, PIC_E_HIP, "adjoint unsort"));
    }
  }
  // the particles alive at the first step's start (loaded and injected
  // before it): their v_z carries V as well
  for (int sp = 0; sp < 2; ++sp) {
    const int64_t cnt = ts[0].n[sp];
    if (cnt <= 0) continue;
    const int blocks = (int)std::min<int64_t>(1024, div_up(cnt, kBlock));
    k_sum_vz<<<blocks, kBlock, 0, st>>>(lp[sp], 0, cnt, tape_pd(0, sp).z, 0, A->red);
    gV += sum_partials(ctx, A->red, blocks, 1, 0, &rc);
    if (rc) {
      cleanup();
      return rc;
    }
  }
  {
    std::vector<double> h(1);
    TRYH(hipMemcpyAsync(h.data(), gvd, sizeof(double), hipMemcpyDeviceToHost, st));
    TRYH(hipStreamSynchronize(st));
    gV += h[0];
  }
  {
    const int blocks = 1024;
    k_dot_g<<<blocks, kBlock, 0, st>>>(n, L, A->G, A->red);
    for (int a = 0; a < 3; ++a) {
      grad[a] = sum_partials(ctx, A->red, blocks, 3, a, &rc);
      if (rc) {
        cleanup();
        return rc;
      }
    }
  }
  grad[3] = gV;
  cleanup();
  return 0;
#undef TRY
#undef TRYH
}

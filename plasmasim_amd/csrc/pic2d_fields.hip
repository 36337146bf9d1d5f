// Field routines of the 2D3V step (both numeric forms, templated on the
// plane type): renew_b4 / renew_e4 (2,4) FDTD, sr_grid / sr_current /
// sr_density on one periodic rank, external_current_perp, renew_phi (DST
// Poisson correction) and the field diagnostics.
//
// Every FDTD and Poisson expression keeps the reference's operation order;
// this translation unit is compiled without floating-point contraction, so
// the FP64 instantiations round like the reference's C.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pic2d_internal.h"
#include "pic_dst.h"

namespace pic2d {

template <typename T>
__global__ void k_zero2(T *__restrict__ p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = (T)0;
}

template <typename T>
static int zero(pic2d_ctx *ctx, T *p, int64_t n) {
  if (n <= 0) return 0;
  const int64_t nb = std::min<int64_t>(div_up(n, kBlock), 8192);
  k_zero2<T><<<(unsigned)nb, kBlock, 0, ctx->rt.stream>>>(p, n);
  P2_LAUNCH_CHECK(ctx);
  return 0;
}

// renew_grid (:1348-1389): J (6), n (2) and phi, contiguous planes
int renew_grid(pic2d_ctx *ctx) {
  const int64_t n = (int64_t)(PIC2D_PHI - PIC_JIX + 1) * ctx->g.plane;
  if (ctx->f64) return zero(ctx, plane<double>(ctx, PIC_JIX), n);
  return zero(ctx, plane<float>(ctx, PIC_JIX), n);
}

template <typename T>
struct FieldSet {
  T *ex, *ey, *ez, *bx, *by, *bz;
  const T *b0x, *b0y, *b0z;
  const T *jix, *jiy, *jiz, *jex, *jey, *jez;
};

template <typename T>
static FieldSet<T> field_set(pic2d_ctx *ctx) {
  FieldSet<T> F;
  F.ex = plane<T>(ctx, PIC_EX);
  F.ey = plane<T>(ctx, PIC_EY);
  F.ez = plane<T>(ctx, PIC_EZ);
  F.bx = plane<T>(ctx, PIC_BX);
  F.by = plane<T>(ctx, PIC_BY);
  F.bz = plane<T>(ctx, PIC_BZ);
  F.b0x = plane<T>(ctx, PIC_B0X);
  F.b0y = plane<T>(ctx, PIC_B0Y);
  F.b0z = plane<T>(ctx, PIC_B0Z);
  F.jix = plane<T>(ctx, PIC_JIX);
  F.jiy = plane<T>(ctx, PIC_JIY);
  F.jiz = plane<T>(ctx, PIC_JIZ);
  F.jex = plane<T>(ctx, PIC_JEX);
  F.jey = plane<T>(ctx, PIC_JEY);
  F.jez = plane<T>(ctx, PIC_JEZ);
  return F;
}

// renew_b4 (:1391-1417): one thread per interior node, m fastest
template <typename T>
__global__ void k_renew_b2(Geom2 g, FieldSet<T> F, const T *__restrict__ sig,
                           T dtf, T dx, T e0) {
  const int m = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int k = 2 + blockIdx.y;
  if (m >= g.NZ + 2) return;
  const int64_t i = g.idx(k, m), dk = g.SZ;
  const int64_t S2 = 2 * (int64_t)g.SZ;
  const T one = 1, half = 0.5, c24 = 24, c27 = 27;
  const T *ex = F.ex, *ey = F.ey, *ez = F.ez;
  T s;
  s = sig[(2 * k + 1) * S2 + 2 * m] / e0 * dtf * half * half;
  const T bx = F.b0x[i] + (one - s) / (one + s) * (F.bx[i] - F.b0x[i]) +
               half * dtf / (one + s) * one / c24 *
                   (c27 * (ey[i] - ey[i - 1]) - (ey[i + 1] - ey[i - 2])) / dx;
  s = sig[(2 * k) * S2 + 2 * m] / e0 * dtf * half * half;
  const T by =
      F.b0y[i] + (one - s) / (one + s) * (F.by[i] - F.b0y[i]) -
      dtf * half / (one + s) *
          (one / c24 * (c27 * (ex[i] - ex[i - 1]) - (ex[i + 1] - ex[i - 2])) / dx -
           one / c24 * (c27 * (ez[i] - ez[i - dk]) - (ez[i + dk] - ez[i - 2 * dk])) /
               dx);
  s = sig[(2 * k) * S2 + 2 * m + 1] / e0 * dtf * half * half;
  const T bz = F.b0z[i] + (one - s) / (one + s) * (F.bz[i] - F.b0z[i]) -
               dtf * half / (one + s) * one / c24 *
                   (c27 * (ey[i] - ey[i - dk]) - (ey[i + dk] - ey[i - 2 * dk])) / dx;
  F.bx[i] = bx;
  F.by[i] = by;
  F.bz[i] = bz;
}

// renew_e4 (:1419-1449): curl of B - B0 and Ji - Je
template <typename T>
__global__ void k_renew_e2(Geom2 g, FieldSet<T> F, const T *__restrict__ sig,
                           T dtf, T dx, T mu0, T C) {
  const int m = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int k = 2 + blockIdx.y;
  if (m >= g.NZ + 2) return;
  const int64_t i = g.idx(k, m), dk = g.SZ;
  const int64_t S2 = 2 * (int64_t)g.SZ;
  const T one = 1, half = 0.5, c24 = 24, c27 = 27;
  const T *bx = F.bx, *by = F.by, *bz = F.bz, *b0x = F.b0x, *b0y = F.b0y,
          *b0z = F.b0z;
  T s;
  s = half * mu0 * sig[(2 * k) * S2 + 2 * m + 1] * C * C * dtf;
  const T ex =
      (one - s) / (one + s) * F.ex[i] +
      C * C * dtf / (one + s) *
          (-one / c24 *
               (c27 * (by[i + 1] - b0y[i + 1] - by[i] + b0y[i]) -
                (by[i + 2] - b0y[i + 2] - by[i - 1] + b0y[i - 1])) /
               dx -
           mu0 * (F.jix[i] - F.jex[i]));
  s = half * mu0 * sig[(2 * k + 1) * S2 + 2 * m] * C * C * dtf;
  const T ez =
      (one - s) / (one + s) * F.ez[i] +
      C * C * dtf / (one + s) *
          (one / c24 *
               (c27 * (by[i + dk] - b0y[i + dk] - by[i] + b0y[i]) -
                (by[i + 2 * dk] - b0y[i + 2 * dk] - by[i - dk] + b0y[i - dk])) /
               dx -
           mu0 * (F.jiz[i] - F.jez[i]));
  s = half * mu0 * sig[(2 * k + 1) * S2 + 2 * m + 1] * C * C * dtf;
  const T ey =
      (one - s) / (one + s) * F.ey[i] +
      C * C * dtf / (one + s) *
          (one / c24 *
               (c27 * (bx[i + 1] - b0x[i + 1] - bx[i] + b0x[i]) -
                (bx[i + 2] - b0x[i + 2] - bx[i - 1] + b0x[i - 1])) /
               dx -
           one / c24 *
               (c27 * (bz[i + dk] - b0z[i + dk] - bz[i] + b0z[i]) -
                (bz[i + 2 * dk] - b0z[i + 2 * dk] - bz[i - dk] + b0z[i - dk])) /
               dx -
           mu0 * (F.jiy[i] - F.jey[i]));
  F.ex[i] = ex;
  F.ey[i] = ey;
  F.ez[i] = ez;
}

template <typename T>
static int renew_b_t(pic2d_ctx *ctx, const T *sig) {
  const pic2d_config &c = ctx->cfg;
  dim3 grid((unsigned)div_up(ctx->g.NZ, kBlock), (unsigned)ctx->g.NX);
  k_renew_b2<T><<<grid, kBlock, 0, ctx->rt.stream>>>(
      ctx->g, field_set<T>(ctx), sig, (T)c.dt, (T)c.dx, (T)c.eps0);
  P2_LAUNCH_CHECK(ctx);
  return 0;
}

template <typename T>
static int renew_e_t(pic2d_ctx *ctx, const T *sig) {
  const pic2d_config &c = ctx->cfg;
  dim3 grid((unsigned)div_up(ctx->g.NZ, kBlock), (unsigned)ctx->g.NX);
  k_renew_e2<T><<<grid, kBlock, 0, ctx->rt.stream>>>(
      ctx->g, field_set<T>(ctx), sig, (T)c.dt, (T)c.dx, (T)c.mu0, (T)c.c);
  P2_LAUNCH_CHECK(ctx);
  return 0;
}

int renew_b_half(pic2d_ctx *ctx) {
  return ctx->f64 ? renew_b_t<double>(ctx, ctx->sig64)
                  : renew_b_t<float>(ctx, ctx->sig32);
}

int renew_e(pic2d_ctx *ctx) {
  return ctx->f64 ? renew_e_t<double>(ctx, ctx->sig64)
                  : renew_e_t<float>(ctx, ctx->sig32);
}

// sr_grid (:1021-1071) on one rank: rows 0,1 <- NX,NX+1 and NX+2,NX+3 <-
// 2,3 of every plane (the reference sends whole Grid records)
template <typename T>
__global__ void k_halo2(Geom2 g, T *__restrict__ base, int nplanes) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = 4 * (int64_t)g.SZ;
  if (t >= per * nplanes) return;
  const int p = (int)(t / per);
  const int r = (int)((t % per) / g.SZ), m = (int)(t % g.SZ);
  const int dst = r < 2 ? r : g.NX + r;
  const int src = r < 2 ? g.NX + r : r;
  T *P = base + (int64_t)p * g.plane;
  P[g.idx(dst, m)] = P[g.idx(src, m)];
}

int halo(pic2d_ctx *ctx) {
  const int64_t n = 4 * (int64_t)ctx->g.SZ * PIC2D_NFIELDS;
  const unsigned nb = (unsigned)div_up(n, kBlock);
  if (ctx->f64)
    k_halo2<double><<<nb, kBlock, 0, ctx->rt.stream>>>(
        ctx->g, (double *)ctx->field_mem, PIC2D_NFIELDS);
  else
    k_halo2<float><<<nb, kBlock, 0, ctx->rt.stream>>>(
        ctx->g, (float *)ctx->field_mem, PIC2D_NFIELDS);
  P2_LAUNCH_CHECK(ctx);
  return 0;
}

// sr_current (:1073-1144) / sr_density (:1146-1209) on one rank: one layer
// each way, row 2 += row NX+2 and row NX+1 += row 1
template <typename T>
__global__ void k_fold2(Geom2 g, T *__restrict__ first, int count) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)count * g.SZ) return;
  const int p = (int)(t / g.SZ), m = (int)(t % g.SZ);
  T *P = first + (int64_t)p * g.plane;
  P[g.idx(2, m)] += P[g.idx(g.NX + 2, m)];
  P[g.idx(g.NX + 1, m)] += P[g.idx(1, m)];
}

int fold(pic2d_ctx *ctx, int first, int count) {
  const unsigned nb = (unsigned)div_up((int64_t)count * ctx->g.SZ, kBlock);
  if (ctx->f64)
    k_fold2<double><<<nb, kBlock, 0, ctx->rt.stream>>>(
        ctx->g, plane<double>(ctx, first), count);
  else
    k_fold2<float><<<nb, kBlock, 0, ctx->rt.stream>>>(
        ctx->g, plane<float>(ctx, first), count);
  P2_LAUNCH_CHECK(ctx);
  return 0;
}

// external_current_perp (:2238-2267): assignment of the two line currents
template <typename T>
__global__ void k_set2(T *__restrict__ p, int64_t i0, T v0, int64_t i1, T v1) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    p[i0] = v0;
    p[i1] = v1;
  }
}

int external_current(pic2d_ctx *ctx, int64_t step) {
  const pic2d_config &c = ctx->cfg;
  if (c.ext_k < 0) return 0;
  const double Pi = 3.14159265358979323846;
  const double I = c.ext_current, dx = c.dx;
  double v0, v1;
  if (step < c.ext_ramp) {
    const double w = Pi / (double)c.ext_ramp;
    v0 = -I / dx / dx * (1. + cos(w * (double)step - Pi)) / 2.;
    v1 = I / dx / dx * (1. + cos(w * (double)step - Pi)) / 2.;
  } else {
    v0 = -I / dx / dx;
    v1 = I / dx / dx;
  }
  const int64_t i0 = ctx->g.idx(c.ext_k, c.ext_m[0]);
  const int64_t i1 = ctx->g.idx(c.ext_k, c.ext_m[1]);
  if (ctx->f64)
    k_set2<double><<<1, 64, 0, ctx->rt.stream>>>(plane<double>(ctx, PIC_JIY), i0,
                                                  v0, i1, v1);
  else
    k_set2<float><<<1, 64, 0, ctx->rt.stream>>>(plane<float>(ctx, PIC_JIY), i0,
                                                 (float)v0, i1, (float)v1);
  P2_LAUNCH_CHECK(ctx);
  return 0;
}

// ---------------------------------------------------------------------------
// renew_phi (:2589-2767): source, DST-I along z then x, division by the
// 5-point symbol, inverse DSTs, E += grad phi

template <typename T>
__global__ void k_phi_rhs(Geom2 g, const T *__restrict__ ni,
                          const T *__restrict__ ne, const T *__restrict__ ex,
                          const T *__restrict__ ez, T *__restrict__ rho,
                          int lox, int hix, int loz, int hiz, T q, T e0, T dx) {
  const int mm = blockIdx.x * blockDim.x + threadIdx.x;
  const int kk = blockIdx.y;
  if (mm >= g.NZ) return;
  const int k = kk + 2, m = mm + 2;
  T v = 0;
  if (k >= lox && k < hix && m >= loz && m < hiz) {
    const int64_t i = g.idx(k, m);
    v = -(-q * (ni[i] - ne[i]) / e0 + (ex[i + g.SZ] - ex[i]) / dx +
          (ez[i + 1] - ez[i]) / dx);
  }
  rho[(int64_t)kk * g.NZ + mm] = v;
}

template <typename T>
__global__ void k_phi_apply(Geom2 g, const T *__restrict__ rho,
                            T *__restrict__ ex, T *__restrict__ ez,
                            T *__restrict__ phi, T dx) {
  const int mm = blockIdx.x * blockDim.x + threadIdx.x;
  const int kk = blockIdx.y;
  if (mm >= g.NZ) return;
  const int64_t r = (int64_t)kk * g.NZ + mm;
  const T ph = rho[r];
  const T pxm = kk == 0 ? rho[(int64_t)(g.NX - 1) * g.NZ + mm] : rho[r - g.NZ];
  const T pzm = mm == 0 ? (T)0 : rho[r - 1];
  const int64_t i = g.idx(kk + 2, mm + 2);
  ex[i] += (ph - pxm) / dx;
  phi[i] = ph;
  ez[i] += (ph - pzm) / dx;
}

// twiddles exp(-2 pi i t / L), t < L/2, and the symbol tables
// sin^2(pi k / (2 N)) of both axes, computed on the host in double as the
// reference computes its symbol (:2679-2686)
int init_poisson(pic2d_ctx *ctx) {
  if (ctx->cfg.poisson_every <= 0) return 0;
  const int NX = ctx->g.NX, NZ = ctx->g.NZ;
  const size_t esz = ctx->f64 ? sizeof(double) : sizeof(float);
  P2_HIP(ctx, hipMalloc(&ctx->rho, (size_t)NX * NZ * esz));
  for (int axis = 0; axis < 2; ++axis) {
    const int N = axis == 0 ? NX : NZ;
    size_t bytes = 0, tw_bytes = 0;
    unsigned char *h = ctx->f64 ? pic::dst_axis_tables<double>(N, &bytes, &tw_bytes)
                                : pic::dst_axis_tables<float>(N, &bytes, &tw_bytes);
    if (!h) return err(ctx, PIC_E_NOMEM, "poisson tables");
    void **dst = axis == 0 ? &ctx->tw_x : &ctx->tw_z;
    hipError_t e = hipMalloc(dst, bytes);
    if (e == hipSuccess) e = hipMemcpy(*dst, h, bytes, hipMemcpyHostToDevice);
    free(h);
    P2_HIP(ctx, e);
  }
  return 0;
}

template <typename T>
static int poisson_t(pic2d_ctx *ctx) {
  const pic2d_config &c = ctx->cfg;
  const Geom2 &g = ctx->g;
  const int NX = g.NX, NZ = g.NZ;
  T *rho = (T *)ctx->rho;
  hipStream_t st = ctx->rt.stream;
  dim3 grid((unsigned)div_up(NZ, kBlock), (unsigned)NX);
  k_phi_rhs<T><<<grid, kBlock, 0, st>>>(
      g, plane<T>(ctx, PIC_NI), plane<T>(ctx, PIC_NE), plane<T>(ctx, PIC_EX),
      plane<T>(ctx, PIC_EZ), rho, c.phi_lo[0], c.phi_hi[0], c.phi_lo[1],
      c.phi_hi[1], (T)c.q, (T)c.eps0, (T)c.dx);
  P2_LAUNCH_CHECK(ctx);
  using pic::Cplx;
  using pic::k_dst;
  const Cplx<T> *twx = (const Cplx<T> *)ctx->tw_x;
  const Cplx<T> *twz = (const Cplx<T> *)ctx->tw_z;
  const double *symx = (const double *)((const char *)ctx->tw_x + (size_t)NX * sizeof(Cplx<T>));
  const double *symz = (const double *)((const char *)ctx->tw_z + (size_t)NZ * sizeof(Cplx<T>));
  const double dq = pow(c.dx / 2., 2);
  const size_t ldsz = (size_t)4 * NZ * sizeof(T), ldsx = (size_t)4 * NX * sizeof(T);
  const unsigned bz = (unsigned)div_up(NX, 2), bx = (unsigned)div_up(NZ, 2);
  const int lz = pic::ilog2_ceil(2 * NZ), lx = pic::ilog2_ceil(2 * NX);
  // forward along z (lines k, contiguous), then along x (lines m, stride
  // NZ) with the division, inverse along x, inverse along z
  k_dst<T><<<bz, 256, ldsz, st>>>(rho, NZ, lz, NX, NX, NZ, 0, 1, twz, 0, nullptr,
                                  nullptr, 0.);
  k_dst<T><<<bx, 256, ldsx, st>>>(rho, NX, lx, NZ, NZ, 1, 0, NZ, twx, 1, symz, symx,
                                  dq);
  k_dst<T><<<bx, 256, ldsx, st>>>(rho, NX, lx, NZ, NZ, 1, 0, NZ, twx, 2, nullptr,
                                  nullptr, 0.);
  k_dst<T><<<bz, 256, ldsz, st>>>(rho, NZ, lz, NX, NX, NZ, 0, 1, twz, 2, nullptr,
                                  nullptr, 0.);
  P2_LAUNCH_CHECK(ctx);
  k_phi_apply<T><<<grid, kBlock, 0, st>>>(g, rho, plane<T>(ctx, PIC_EX),
                                          plane<T>(ctx, PIC_EZ),
                                          plane<T>(ctx, PIC2D_PHI), (T)c.dx);
  P2_LAUNCH_CHECK(ctx);
  return 0;
}

int poisson(pic2d_ctx *ctx) {
  if (!ctx->rho) return err(ctx, PIC_E_STATE, "poisson: disabled by poisson_every = 0");
  return ctx->f64 ? poisson_t<double>(ctx) : poisson_t<float>(ctx);
}

// ---------------------------------------------------------------------------
// field diagnostics: one workgroup per interior row, FP64 partial sums
// (|E|^2, |B|^2, |B-B0|^2, ni-ne, gauss^2, divb^2) and maxima (|gauss|,
// |divb|) per row; the host adds the rows in order

template <typename T>
__global__ __launch_bounds__(256) void k_diag_rows(Geom2 g, FieldSet<T> F,
                                                   const T *__restrict__ ni,
                                                   const T *__restrict__ ne,
                                                   double q, double e0,
                                                   double dx,
                                                   double *__restrict__ part) {
  __shared__ double red[8][256];
  const int k = 2 + blockIdx.x;
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int m = 2 + threadIdx.x; m < g.NZ + 2; m += blockDim.x) {
    const int64_t i = g.idx(k, m), dk = g.SZ;
    const double ex = F.ex[i], ey = F.ey[i], ez = F.ez[i];
    const double bx = F.bx[i], by = F.by[i], bz = F.bz[i];
    v[0] += ex * ex + ey * ey + ez * ez;
    v[1] += bx * bx + by * by + bz * bz;
    const double b1x = bx - (double)F.b0x[i], b1y = by - (double)F.b0y[i],
                 b1z = bz - (double)F.b0z[i];
    v[2] += b1x * b1x + b1y * b1y + b1z * b1z;
    const double dn = (double)ni[i] - (double)ne[i];
    v[3] += dn;
    const double r = ((double)F.ex[i + dk] - ex + (double)F.ez[i + 1] - ez) / dx -
                     q * dn / e0;
    v[4] += r * r;
    v[6] = fmax(v[6], fabs(r));
    const double db = (bx - (double)F.b0x[i] - (double)F.bx[i - dk] +
                       (double)F.b0x[i - dk] + bz - (double)F.b0z[i] -
                       (double)F.bz[i - 1] + (double)F.b0z[i - 1]) /
                      dx;
    v[5] += db * db;
    v[7] = fmax(v[7], fabs(db));
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) red[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int c = 0; c < 6; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + s];
      red[6][threadIdx.x] = fmax(red[6][threadIdx.x], red[6][threadIdx.x + s]);
      red[7][threadIdx.x] = fmax(red[7][threadIdx.x], red[7][threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 8) part[(int64_t)blockIdx.x * 8 + threadIdx.x] = red[threadIdx.x][0];
}

template <typename T>
static int diag_t(pic2d_ctx *ctx, pic2d_diag *d) {
  const pic2d_config &c = ctx->cfg;
  const int NX = ctx->g.NX;
  double *part = nullptr;
  P2_HIP(ctx, hipMalloc(&part, (size_t)NX * 8 * sizeof(double)));
  k_diag_rows<T><<<(unsigned)NX, 256, 0, ctx->rt.stream>>>(
      ctx->g, field_set<T>(ctx), plane<T>(ctx, PIC_NI), plane<T>(ctx, PIC_NE),
      c.q, c.eps0, c.dx, part);
  hipError_t e = hipGetLastError();
  double *h = (double *)malloc((size_t)NX * 8 * sizeof(double));
  if (e == hipSuccess && h)
    e = hipMemcpyAsync(h, part, (size_t)NX * 8 * sizeof(double),
                       hipMemcpyDeviceToHost, ctx->rt.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->rt.stream);
  (void)hipFree(part);
  if (!h) return err(ctx, PIC_E_NOMEM, "diagnose");
  if (e != hipSuccess) {
    free(h);
    P2_HIP(ctx, e);
  }
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int r = 0; r < NX; ++r) {
    for (int cc = 0; cc < 6; ++cc) s[cc] += h[r * 8 + cc];
    s[6] = fmax(s[6], h[r * 8 + 6]);
    s[7] = fmax(s[7], h[r * 8 + 7]);
  }
  free(h);
  const double dx2 = c.dx * c.dx;
  d->field_energy_e = 0.5 * c.eps0 * s[0] * dx2;
  d->field_energy_b = 0.5 * s[1] / c.mu0 * dx2;
  d->field_energy_b1 = 0.5 * s[2] / c.mu0 * dx2;
  d->charge = c.q * s[3] * dx2;
  d->gauss_l2 = sqrt(s[4]);
  d->divb_l2 = sqrt(s[5]);
  d->gauss_max = s[6];
  d->divb_max = s[7];
  return 0;
}

int diagnose_fields(pic2d_ctx *ctx, pic2d_diag *d) {
  return ctx->f64 ? diag_t<double>(ctx, d) : diag_t<float>(ctx, d);
}

}  // namespace pic2d

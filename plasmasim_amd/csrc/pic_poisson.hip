// The 3D Dirichlet Poisson correction of E: renew_phi3 (:8315-8665) and
// sr_phi (:1389-1573), for one subdomain, in either numeric form.
//
// The reference gathers the source on rank 0, transposes it between its
// sine transforms with MPI and scatters it back.  Here the source lives in
// one device buffer rho[k][l][m] (interior cells, m contiguous) and each
// sine transform is one batched launch over the lines of one axis (two
// lines per workgroup, a 2n-point complex FFT in LDS, pic_dst.h), so no
// transpose is needed.  Transform order and the symbol's order of
// summation are the reference's.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pic_dst.h"
#include "pic_internal.h"

namespace pic {
namespace {

inline bool f64_form(const pic_ctx *ctx) { return ctx->cfg.precision == PIC_FP64; }

template <typename T>
T *plane(pic_ctx *ctx, int which) {
  return (T *)ctx->field_mem + (int64_t)which * ctx->geom.plane;
}

// source of renew_phi3 (:8336-8353): q (ni - ne)/eps0 minus the (2,4)
// divergence of E, zero outside the margin box
template <typename T>
__global__ void k_phi3_src(Geom g, const T *__restrict__ ex, const T *__restrict__ ey,
                           const T *__restrict__ ez, const T *__restrict__ ni,
                           const T *__restrict__ ne, T *__restrict__ rho, int margin,
                           T q, T e0, T dx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)g.NX * g.NY * g.NZ;
  if (i >= n) return;
  const int mm = (int)(i % g.NZ);
  const int64_t t = i / g.NZ;
  const int ll = (int)(t % g.NY), kk = (int)(t / g.NY);
  T v = 0;
  if (kk > margin && kk < g.NX - margin && ll > margin && ll < g.NY - margin &&
      mm >= margin && mm < g.NZ - margin) {
    const int64_t dk = (int64_t)g.SY * g.SZ, dl = g.SZ;
    const int64_t c = g.idx(kk + 2, ll + 2, mm + 2);
    v = -(-q * (ni[c] - ne[c]) / e0 +
          (-ex[c + 2 * dk] + 27 * ex[c + dk] - 27 * ex[c] + ex[c - dk] -
           ey[c + 2 * dl] + 27 * ey[c + dl] - 27 * ey[c] + ey[c - dl] - ez[c + 2] +
           27 * ez[c + 1] - 27 * ez[c] + ez[c - 1]) /
              (24 * dx));
  }
  rho[i] = v;
}

// phi after renew_phi3 (interior) and sr_phi (x and y guards from their
// periodic images, :1389-1560); the z guards keep renew_grid's zero
template <typename T>
__global__ void k_phi3_store(Geom g, const T *__restrict__ rho, T *__restrict__ phi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)g.SX * g.SY * g.SZ;
  if (i >= n) return;
  const int m = (int)(i % g.SZ);
  const int64_t t = i / g.SZ;
  const int l = (int)(t % g.SY), k = (int)(t / g.SY);
  T v = 0;
  if (m >= 2 && m < g.NZ + 2) {
    const int kk = k < 2 ? k + g.NX - 2 : k >= g.NX + 2 ? k - g.NX - 2 : k - 2;
    const int ll = l < 2 ? l + g.NY - 2 : l >= g.NY + 2 ? l - g.NY - 2 : l - 2;
    v = rho[((int64_t)kk * g.NY + ll) * g.NZ + (m - 2)];
  }
  phi[i] = v;
}

// E += the (2,4) gradient of phi over the interior (:1563-1571)
template <typename T>
__global__ void k_phi3_apply(Geom g, const T *__restrict__ phi, T *__restrict__ ex,
                             T *__restrict__ ey, T *__restrict__ ez, T dx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)g.NX * g.NY * g.NZ;
  if (i >= n) return;
  const int mm = (int)(i % g.NZ);
  const int64_t t = i / g.NZ;
  const int ll = (int)(t % g.NY), kk = (int)(t / g.NY);
  const int64_t dk = (int64_t)g.SY * g.SZ, dl = g.SZ;
  const int64_t c = g.idx(kk + 2, ll + 2, mm + 2);
  ex[c] += (-phi[c + dk] + 27 * phi[c] - 27 * phi[c - dk] + phi[c - 2 * dk]) / (24 * dx);
  ey[c] += (-phi[c + dl] + 27 * phi[c] - 27 * phi[c - dl] + phi[c - 2 * dl]) / (24 * dx);
  ez[c] += (-phi[c + 1] + 27 * phi[c] - 27 * phi[c - 1] + phi[c - 2]) / (24 * dx);
}

template <typename T>
__global__ void k_fill0(T *__restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

template <typename T>
int run(pic_ctx *ctx) {
  const Geom &g = ctx->geom;
  const pic_config &c = ctx->cfg;
  hipStream_t st = ctx->stream;
  T *rho = (T *)ctx->rho;
  const int64_t ni = (int64_t)g.NX * g.NY * g.NZ;
  const int64_t na = (int64_t)g.SX * g.SY * g.SZ;
  k_phi3_src<T><<<(unsigned)div_up(ni, kBlock), kBlock, 0, st>>>(
      g, plane<T>(ctx, PIC_EX), plane<T>(ctx, PIC_EY), plane<T>(ctx, PIC_EZ),
      plane<T>(ctx, PIC_NI), plane<T>(ctx, PIC_NE), rho, c.phi_margin, (T)c.q,
      (T)c.eps0, (T)c.dx);
  PIC_LAUNCH_CHECK(ctx);
  const int N[3] = {g.NX, g.NY, g.NZ};
  const Cplx<T> *tw[3];
  const double *sym[3];
  for (int a = 0; a < 3; ++a) {
    tw[a] = (const Cplx<T> *)ctx->dst_tab[a];
    sym[a] = (const double *)((const char *)ctx->dst_tab[a] + (size_t)N[a] * sizeof(Cplx<T>));
  }
  const double dq = pow(c.dx / 2., 2);
  const int64_t NYZ = (int64_t)g.NY * g.NZ;
  // lines: z (k,l) contiguous; y (k,m) stride NZ; x (l,m) stride NY NZ
  struct Pass {
    int axis, nlines, ninner;
    int64_t sline, souter, selem;
  };
  const Pass pz{2, g.NX * g.NY, g.NX * g.NY, g.NZ, 0, 1};
  const Pass py{1, g.NX * g.NZ, g.NZ, 1, NYZ, g.NZ};
  const Pass px{0, g.NY * g.NZ, g.NY * g.NZ, 1, 0, NYZ};
  auto launch = [&](const Pass &p, int mode) {
    const int n = N[p.axis];
    k_dst<T><<<(unsigned)div_up(p.nlines, 2), 256, (size_t)4 * n * sizeof(T), st>>>(
        rho, n, ilog2_ceil(2 * n), p.nlines, p.ninner, p.sline, p.souter, p.selem,
        tw[p.axis], mode, mode == 1 ? ctx->sym_yz : nullptr,
        mode == 1 ? sym[0] : nullptr, dq);
  };
  // fft_z, fft_y, fft_x, division (:8360-8512), inv_fft_x, inv_fft_y,
  // inv_fft_z (:8517-8653)
  launch(pz, 0);
  launch(py, 0);
  launch(px, 1);
  launch(px, 2);
  launch(py, 2);
  launch(pz, 2);
  PIC_LAUNCH_CHECK(ctx);
  k_phi3_store<T><<<(unsigned)div_up(na, kBlock), kBlock, 0, st>>>(g, rho, (T *)ctx->phi);
  PIC_LAUNCH_CHECK(ctx);
  k_phi3_apply<T><<<(unsigned)div_up(ni, kBlock), kBlock, 0, st>>>(
      g, (const T *)ctx->phi, plane<T>(ctx, PIC_EX), plane<T>(ctx, PIC_EY),
      plane<T>(ctx, PIC_EZ), (T)c.dx);
  PIC_LAUNCH_CHECK(ctx);
  return 0;
}

}  // namespace

int poisson_init(pic_ctx *ctx) {
  const Geom &g = ctx->geom;
  const bool f64 = f64_form(ctx);
  const size_t esz = f64 ? sizeof(double) : sizeof(float);
  PIC_HIP(ctx, hipMalloc(&ctx->rho, (size_t)g.NX * g.NY * g.NZ * esz));
  PIC_HIP(ctx, hipMalloc(&ctx->phi, (size_t)g.SX * g.SY * g.SZ * esz));
  int rc = poisson_zero_phi(ctx);
  if (rc) return rc;
  const int N[3] = {g.NX, g.NY, g.NZ};
  double *symh[3] = {nullptr, nullptr, nullptr};
  for (int a = 0; a < 3; ++a) {
    size_t bytes = 0, tw_bytes = 0;
    unsigned char *h = f64 ? dst_axis_tables<double>(N[a], &bytes, &tw_bytes)
                           : dst_axis_tables<float>(N[a], &bytes, &tw_bytes);
    if (!h) return set_error(ctx, PIC_E_NOMEM, "poisson tables");
    symh[a] = (double *)malloc((size_t)N[a] * sizeof(double));
    if (symh[a]) memcpy(symh[a], h + tw_bytes, (size_t)N[a] * sizeof(double));
    hipError_t e = hipMalloc(&ctx->dst_tab[a], bytes);
    if (e == hipSuccess) e = hipMemcpy(ctx->dst_tab[a], h, bytes, hipMemcpyHostToDevice);
    free(h);
    if (e != hipSuccess || !symh[a]) {
      for (int b = 0; b <= a; ++b) free(symh[b]);
      if (e != hipSuccess)
        return set_error(ctx, PIC_E_HIP, "poisson tables: %s", hipGetErrorString(e));
      return set_error(ctx, PIC_E_NOMEM, "poisson tables");
    }
  }
  // per-line symbol of the x pass: sin^2 of the y index + sin^2 of the z
  // index, summed first as in the reference's A (:8502-8504)
  const size_t nyz = (size_t)g.NY * g.NZ;
  double *h = (double *)malloc(nyz * sizeof(double));
  hipError_t e = h ? hipMalloc(&ctx->sym_yz, nyz * sizeof(double)) : hipErrorOutOfMemory;
  if (h && e == hipSuccess) {
    for (int l = 0; l < g.NY; ++l)
      for (int m = 0; m < g.NZ; ++m) h[(size_t)l * g.NZ + m] = symh[1][l] + symh[2][m];
    e = hipMemcpy(ctx->sym_yz, h, nyz * sizeof(double), hipMemcpyHostToDevice);
  }
  free(h);
  for (int a = 0; a < 3; ++a) free(symh[a]);
  if (e != hipSuccess)
    return set_error(ctx, PIC_E_HIP, "poisson symbol: %s", hipGetErrorString(e));
  return 0;
}

void poisson_free(pic_ctx *ctx) {
  if (ctx->rho) (void)hipFree(ctx->rho);
  if (ctx->phi) (void)hipFree(ctx->phi);
  for (int a = 0; a < 3; ++a)
    if (ctx->dst_tab[a]) (void)hipFree(ctx->dst_tab[a]);
  if (ctx->sym_yz) (void)hipFree(ctx->sym_yz);
  ctx->rho = ctx->phi = nullptr;
  ctx->dst_tab[0] = ctx->dst_tab[1] = ctx->dst_tab[2] = nullptr;
  ctx->sym_yz = nullptr;
}

int poisson_zero_phi(pic_ctx *ctx) {
  const Geom &g = ctx->geom;
  const int64_t n = (int64_t)g.SX * g.SY * g.SZ;
  if (f64_form(ctx))
    k_fill0<double><<<(unsigned)div_up(n, kBlock), kBlock, 0, ctx->stream>>>(
        (double *)ctx->phi, n);
  else
    k_fill0<float><<<(unsigned)div_up(n, kBlock), kBlock, 0, ctx->stream>>>(
        (float *)ctx->phi, n);
  PIC_LAUNCH_CHECK(ctx);
  return 0;
}

int poisson(pic_ctx *ctx) {
  if (f64_form(ctx)) return run<double>(ctx);
  int rc = f32::finish_deposit(ctx);  // the density planes of the source
  if (rc) return rc;
  return run<float>(ctx);
}

int poisson_download_phi(pic_ctx *ctx, double *host) {
  const Geom &g = ctx->geom;
  const size_t n = (size_t)g.SX * g.SY * g.SZ;
  if (!ctx->phi) {
    memset(host, 0, n * sizeof(double));
    return 0;
  }
  PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (f64_form(ctx)) {
    PIC_HIP(ctx, hipMemcpy(host, ctx->phi, n * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
  }
  float *tmp = (float *)malloc(n * sizeof(float));
  if (!tmp) return set_error(ctx, PIC_E_NOMEM, "phi download: host buffer");
  hipError_t e = hipMemcpy(tmp, ctx->phi, n * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    for (size_t i = 0; i < n; ++i) host[i] = (double)tmp[i];
  free(tmp);
  if (e != hipSuccess)
    return set_error(ctx, PIC_E_HIP, "phi download: %s", hipGetErrorString(e));
  return 0;
}

}  // namespace pic

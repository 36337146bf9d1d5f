// C-ABI of the 2D3V step (include/plasmasim2d.h): configuration, lifetime,
// transfers, the routines of fdtd() (PIC2D/2d_pic278_mpi.c :527-1019) and
// the fused step loop.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include <new>

#include "pic2d_internal.h"

using namespace pic2d;

namespace {

const double kPi = 3.14159265358979323846;

bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

int validate(const pic2d_config *c, char *msg, size_t len) {
#define BAD(...)                      \
  do {                                \
    snprintf(msg, len, __VA_ARGS__);  \
    return PIC_E_ARG;                 \
  } while (0)
  if (c->nx < 4 || c->nz < 4) BAD("nx, nz must be >= 4");
  if ((int64_t)(c->nx + 4) * (c->nz + 4) >= ((int64_t)1 << 31))
    BAD("grid too large for 32-bit cell indices");
  if (c->ppc < 0) BAD("ppc must be >= 0");
  if (!(c->dx > 0) || !(c->dt > 0) || !(c->c > 0) || !(c->eps0 > 0) || !(c->mu0 > 0))
    BAD("dx, dt, c, eps0, mu0 must be positive");
  if (!(c->mass[0] > 0) || !(c->mass[1] > 0)) BAD("masses must be positive");
  if (c->precision != PIC_FP64 && c->precision != PIC_FP32) BAD("precision");
  if (c->deposit != PIC_DEPOSIT_ATOMIC)
    BAD("the 2D form has the atomic deposit mode only");
  const int TX = c->tile[0] > 0 ? c->tile[0] : c->nx;
  const int TZ = c->tile[1] > 0 ? c->tile[1] : c->nz;
  if (c->nx % TX || c->nz % TZ) BAD("tiles must divide the interior");
  if ((int64_t)(TX + 6) * (TZ + 6) > 1024 && c->tile[0] > 0)
    BAD("tile (%d, %d) exceeds the LDS window", TX, TZ);
  if (c->poisson_every > 0) {
    if (!pow2(c->nx) || !pow2(c->nz)) BAD("the Poisson solve needs power-of-two nx, nz");
    const int lim = c->precision == PIC_FP64 ? 2048 : 4096;
    if (c->nx > lim || c->nz > lim) BAD("the Poisson solve supports nx, nz <= %d", lim);
    for (int a = 0; a < 2; ++a) {
      const int N = a == 0 ? c->nx : c->nz;
      if (c->phi_lo[a] < 2 || c->phi_hi[a] > N + 2 || c->phi_lo[a] > c->phi_hi[a])
        BAD("phi box must lie in [2, N+2)");
    }
  }
  if (c->ext_k >= 0) {
    if (c->ext_k < 2 || c->ext_k >= c->nx + 2) BAD("ext_k outside the interior");
    for (int j = 0; j < 2; ++j)
      if (c->ext_m[j] < 2 || c->ext_m[j] >= c->nz + 2) BAD("ext_m outside the interior");
    if (c->ext_ramp < 0) BAD("ext_ramp must be >= 0");
  }
  if (c->inflow_every < 0 || c->sort_every < 0 || c->poisson_every < 0)
    BAD("cadences must be >= 0");
  if (c->absorb[0] < 0 || c->absorb[1] < 0) BAD("absorb must be >= 0");
  return 0;
#undef BAD
}

// sigma depth on the 2x lattice, the layer of width A mapped onto the
// reference's 32-cell profile (init_grid :270-283)
double depth2(int i2, int N, int A) {
  double d;
  if (A <= 0) return 0.0;
  if (i2 - 4 <= 2 * A)
    d = (2 * A - (i2 - 4)) / 4.;
  else if (i2 - 4 >= 2 * N - 2 * A)
    d = (i2 - 4 - (2 * N - 2 * A)) / 4.;
  else
    d = 0.;
  return d * (32.0 / A);
}

void free_all(pic2d_ctx *ctx) {
  pic_ctx &rt = ctx->rt;
  tsort_free(ctx);
  if (ctx->field_mem) (void)hipFree(ctx->field_mem);
  if (ctx->sig64) (void)hipFree(ctx->sig64);
  if (ctx->sig32) (void)hipFree(ctx->sig32);
  if (ctx->rho) (void)hipFree(ctx->rho);
  if (ctx->tw_x) (void)hipFree(ctx->tw_x);
  if (ctx->tw_z) (void)hipFree(ctx->tw_z);
  for (int s = 0; s < 2; ++s) {
    for (int b = 0; b < 2; ++b)
      if (ctx->pbuf[s][b]) (void)hipFree(ctx->pbuf[s][b]);
    if (ctx->tile_start[s]) (void)hipFree(ctx->tile_start[s]);
  }
  if (rt.keys) (void)hipFree(rt.keys);
  if (rt.keys_alt) (void)hipFree(rt.keys_alt);
  if (rt.idx) (void)hipFree(rt.idx);
  if (rt.idx_alt) (void)hipFree(rt.idx_alt);
  if (rt.flags) (void)hipFree(rt.flags);
  if (rt.prim_tmp) (void)hipFree(rt.prim_tmp);
  if (rt.dev_counter) (void)hipFree(rt.dev_counter);
  if (rt.host_counter) (void)hipHostFree(rt.host_counter);
  for (int i = 0; i < pic_ctx::kEventPool; ++i)
    for (int j = 0; j < 2; ++j)
      if (rt.tev[j][i]) (void)hipEventDestroy(rt.tev[j][i]);
  if (rt.stream) (void)hipStreamDestroy(rt.stream);
}

// host copy of one field plane, converted from the device type
int download_plane(pic2d_ctx *ctx, int which, double *host) {
  const int64_t n = ctx->g.plane;
  if (ctx->f64) {
    P2_HIP(ctx, hipMemcpyAsync(host, plane<double>(ctx, which), n * sizeof(double),
                               hipMemcpyDeviceToHost, ctx->rt.stream));
    P2_HIP(ctx, hipStreamSynchronize(ctx->rt.stream));
    return 0;
  }
  float *tmp = (float *)malloc(n * sizeof(float));
  if (!tmp) return err(ctx, PIC_E_NOMEM, "download");
  hipError_t e = hipMemcpyAsync(tmp, plane<float>(ctx, which), n * sizeof(float),
                                hipMemcpyDeviceToHost, ctx->rt.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->rt.stream);
  if (e == hipSuccess)
    for (int64_t i = 0; i < n; ++i) host[i] = tmp[i];
  free(tmp);
  P2_HIP(ctx, e);
  return 0;
}

int fill_preset(int which, pic2d_config *c) {
  memset(c, 0, sizeof(*c));
  // normalized units: c = eps0 = mu0 = 1, dx = 1, q = 1, m_i/m_e = 100,
  // omega_pe dt = 0.1 at dt = 0.5 (2D limit of the (2,4) stencil: 0.606)
  c->dx = 1.0;
  c->dt = 0.5;
  c->c = 1.0;
  c->eps0 = 1.0;
  c->mu0 = 1.0;
  c->q = 1.0;
  c->mass[0] = 100.0;
  c->mass[1] = 1.0;
  c->vth[0] = 0.005;
  c->vth[1] = 0.05;
  c->vdrift = 0.1;
  c->precision = PIC_FP64;
  c->deposit = PIC_DEPOSIT_ATOMIC;
  c->poisson_every = 1;
  c->seed = 2000 + which;
  int A = 32, A2 = 16, A3 = 32;  // Absorb_grid, Absorb_grid2, Absorb_grid3
  double L = 64;                 // standoff in cells for the line current
  switch (which) {
    case 0:
      c->nx = 32;
      c->nz = 64;
      c->ppc = 8;
      c->tile[0] = 8;
      c->tile[1] = 16;
      c->inflow_every = 5;
      c->sort_every = 0;
      c->ext_ramp = 20;
      A = 8;
      A2 = 6;
      A3 = 8;
      L = 4;
      break;
    case 1:
      c->nx = 256;
      c->nz = 256;
      c->ppc = 16;
      c->tile[0] = 16;
      c->tile[1] = 16;
      c->inflow_every = 50;
      c->sort_every = 20;
      c->ext_ramp = 200;
      A = 16;
      A2 = 8;
      A3 = 16;
      L = 16;
      break;
    case 3:
      c->nx = 2048;
      c->nz = 2048;
      c->ppc = 100;
      c->tile[0] = 16;
      c->tile[1] = 16;
      c->inflow_every = 250;
      c->sort_every = 20;
      c->ext_ramp = 1000;
      break;
    default:
      return PIC_E_ARG;
  }
  const double n0 = pow(0.1 / c->dt, 2) * c->eps0 * c->mass[1] / (c->q * c->q);
  c->weight[0] = c->weight[1] = n0 / (c->ppc > 0 ? c->ppc : 1);
  c->absorb[0] = 0;  // x is periodic
  c->absorb[1] = A;
  // reference a = 1e-9 S/m at dt = 0.06 us: sigma dt / eps0 = 6.776e-6
  c->sigma_amp = 6.776e-6 / c->dt;
  c->ballistic_lo[0] = -1e300;
  c->ballistic_hi[0] = 1e300;
  c->ballistic_lo[1] = A2 * c->dx;
  c->ballistic_hi[1] = (c->nz + 4 - A2) * c->dx;
  c->phi_lo[0] = 2;
  c->phi_hi[0] = c->nx + 2;
  c->phi_lo[1] = A3;
  c->phi_hi[1] = c->nz + 4 - A3;
  // line currents at the domain centre, 5 cells apart along the flow
  // (external_current_perp :2246-2253); I from the reference's standoff
  // formula L^4 = mu0 4 I^2 R^2 / (pi^2 n m_i V^2) with 2R = 5 dx (init_L
  // :190-199)
  c->ext_k = (c->nx + 4) / 2 + 1;
  c->ext_m[0] = (c->nz + 4) / 2 - 3;
  c->ext_m[1] = (c->nz + 4) / 2 + 2;
  const double R = 2.5 * c->dx;
  c->ext_current = sqrt(pow(L * c->dx, 4) * kPi * kPi * n0 * c->mass[0] *
                        c->vdrift * c->vdrift / (c->mu0 * 4.0 * R * R));
  return 0;
}

}  // namespace

extern "C" size_t pic2d_sizeof_config(void) { return sizeof(pic2d_config); }
extern "C" size_t pic2d_sizeof_diag(void) { return sizeof(pic2d_diag); }

extern "C" int pic2d_config_preset(int which, pic2d_config *out) {
  if (!out) return PIC_E_ARG;
  return fill_preset(which, out);
}

extern "C" int pic2d_create(const pic2d_config *cfg, pic2d_ctx **out) {
  if (!cfg || !out) return PIC_E_ARG;
  *out = nullptr;
  char msg[256];
  int rc = validate(cfg, msg, sizeof(msg));
  if (rc) {
    fprintf(stderr, "pic2d_create: %s\n", msg);
    return rc;
  }
  pic2d_ctx *ctx = new (std::nothrow) pic2d_ctx();
  if (!ctx) return PIC_E_NOMEM;
  ctx->rt.err[0] = 0;
  ctx->cfg = *cfg;
  ctx->f64 = cfg->precision == PIC_FP64;
  Geom2 &g = ctx->g;
  g.NX = cfg->nx;
  g.NZ = cfg->nz;
  g.SX = cfg->nx + 4;
  g.SZ = cfg->nz + 4;
  g.plane = (int64_t)g.SX * g.SZ;
  Tile2 &t = ctx->tg;
  t.TX = cfg->tile[0] > 0 ? cfg->tile[0] : cfg->nx;
  t.TZ = cfg->tile[1] > 0 ? cfg->tile[1] : cfg->nz;
  t.NTX = cfg->nx / t.TX;
  t.NTZ = cfg->nz / t.TZ;
  const size_t esz = ctx->f64 ? sizeof(double) : sizeof(float);
#define CREATE_HIP(call)                                                   \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess) {                                                \
      fprintf(stderr, "pic2d_create: %s: %s\n", #call, hipGetErrorString(e_)); \
      free_all(ctx);                                                       \
      delete ctx;                                                          \
      return e_ == hipErrorOutOfMemory ? PIC_E_NOMEM : PIC_E_HIP;          \
    }                                                                      \
  } while (0)
  CREATE_HIP(hipStreamCreateWithFlags(&ctx->rt.stream, hipStreamNonBlocking));
  CREATE_HIP(hipMalloc(&ctx->field_mem, (size_t)PIC2D_NFIELDS * g.plane * esz));
  CREATE_HIP(hipMemset(ctx->field_mem, 0, (size_t)PIC2D_NFIELDS * g.plane * esz));
  // B0 = B = IMF (init_grid: the line-dipole terms are zeroed, :365, :380)
  {
    const int64_t n = g.plane;
    void *h = malloc(n * esz);
    if (!h) {
      free_all(ctx);
      delete ctx;
      return PIC_E_NOMEM;
    }
    const int pl[6] = {PIC_BX, PIC_BY, PIC_BZ, PIC_B0X, PIC_B0Y, PIC_B0Z};
    for (int j = 0; j < 6; ++j) {
      const double v = cfg->imf[j % 3];
      for (int64_t i = 0; i < n; ++i) {
        if (ctx->f64)
          ((double *)h)[i] = v;
        else
          ((float *)h)[i] = (float)v;
      }
      hipError_t e = hipMemcpy((char *)ctx->field_mem + (size_t)pl[j] * n * esz, h,
                               n * esz, hipMemcpyHostToDevice);
      if (e != hipSuccess) free(h);
      CREATE_HIP(e);
    }
    free(h);
  }
  // sigma on the 2x lattice (init_grid :254-285)
  {
    const int64_t n = (int64_t)4 * g.SX * g.SZ;
    double *h = (double *)malloc(n * sizeof(double));
    if (!h) {
      free_all(ctx);
      delete ctx;
      return PIC_E_NOMEM;
    }
    for (int k = 0; k < 2 * g.SX; ++k) {
      const double x = depth2(k, g.NX, cfg->absorb[0]);
      for (int m = 0; m < 2 * g.SZ; ++m) {
        const double z = depth2(m, g.NZ, cfg->absorb[1]);
        h[(int64_t)k * 2 * g.SZ + m] =
            cfg->sigma_amp * (0.1 * sqrt(pow(x, 4) + pow(z, 4)) +
                              0.01 * sqrt(pow(x, 12) + pow(z, 12)));
      }
    }
    hipError_t e;
    if (ctx->f64) {
      e = hipMalloc(&ctx->sig64, n * sizeof(double));
      if (e == hipSuccess) e = hipMemcpy(ctx->sig64, h, n * sizeof(double), hipMemcpyHostToDevice);
    } else {
      float *f = (float *)h;  // in place: element i is read before written
      for (int64_t i = 0; i < n; ++i) f[i] = (float)h[i];
      e = hipMalloc(&ctx->sig32, n * sizeof(float));
      if (e == hipSuccess) e = hipMemcpy(ctx->sig32, f, n * sizeof(float), hipMemcpyHostToDevice);
    }
    free(h);
    CREATE_HIP(e);
  }
  // particles: two buffers per species (compaction and sort swap them)
  const int64_t ncell = (int64_t)g.NX * g.NZ;
  int64_t cap = cfg->capacity;
  if (cap <= 0)
    cap = ncell * cfg->ppc + ncell * cfg->ppc / 8 + (int64_t)8 * g.NX * cfg->ppc + 4096;
  cap = div_up(cap, 256) * 256;
  if (cap >= ((int64_t)1 << 32) - 1) {
    fprintf(stderr, "pic2d_create: capacity %lld exceeds 32-bit indexing\n",
            (long long)cap);
    free_all(ctx);
    delete ctx;
    return PIC_E_ARG;
  }
  ctx->cap = cap;
  const size_t pbytes = (size_t)cap * (ctx->f64 ? 5 * sizeof(double)
                                                : sizeof(int32_t) + 5 * sizeof(float));
  for (int s = 0; s < 2; ++s)
    for (int b = 0; b < 2; ++b) CREATE_HIP(hipMalloc(&ctx->pbuf[s][b], pbytes));
  pic_ctx &rt = ctx->rt;
  CREATE_HIP(hipMalloc(&rt.keys, cap * sizeof(uint32_t)));
  CREATE_HIP(hipMalloc(&rt.keys_alt, cap * sizeof(uint32_t)));
  CREATE_HIP(hipMalloc(&rt.idx, cap * sizeof(uint32_t)));
  CREATE_HIP(hipMalloc(&rt.flags, cap));
  for (int s = 0; s < 2; ++s) {
    const size_t tb = (size_t)(t.ntiles() + 1) * sizeof(uint32_t);
    CREATE_HIP(hipMalloc(&ctx->tile_start[s], tb));
    CREATE_HIP(hipMemset(ctx->tile_start[s], 0, tb));
  }
  CREATE_HIP(hipMalloc(&rt.dev_counter, 8 * sizeof(int64_t)));
  CREATE_HIP(hipHostMalloc(&rt.host_counter, 8 * sizeof(int64_t)));
  for (int i = 0; i < pic_ctx::kEventPool; ++i)
    for (int j = 0; j < 2; ++j) CREATE_HIP(hipEventCreate(&rt.tev[j][i]));
  CREATE_HIP(hipStreamSynchronize(rt.stream));
#undef CREATE_HIP
  rc = init_poisson(ctx);
  if (!rc) rc = tsort_setup(ctx);
  if (rc) {
    fprintf(stderr, "pic2d_create: %s\n", rt.err);
    free_all(ctx);
    delete ctx;
    return rc;
  }
  // fixed-point unit of the FP32 LDS sums: a power of two with
  // |q w v / dx^2| * scale <= 2^30 for |v| up to 2c
  {
    const double wmax = cfg->weight[0] > cfg->weight[1] ? cfg->weight[0] : cfg->weight[1];
    const double qmax = cfg->q > 1.0 ? cfg->q : 1.0;
    double unit = qmax * (wmax > 0 ? wmax : 1.0) * 2.0 * cfg->c / (cfg->dx * cfg->dx);
    const double nunit = (wmax > 0 ? wmax : 1.0) / (cfg->dx * cfg->dx);
    if (nunit > unit) unit = nunit;
    ctx->fx_scale = ldexp(1.0, (int)floor(30.0 - log2(unit)));
  }
  *out = ctx;
  return 0;
}

extern "C" void pic2d_destroy(pic2d_ctx *ctx) {
  if (!ctx) return;
  if (ctx->rt.stream) (void)hipStreamSynchronize(ctx->rt.stream);
  free_all(ctx);
  delete ctx;
}

extern "C" const char *pic2d_last_error(const pic2d_ctx *ctx) {
  return ctx ? ctx->rt.err : "null context";
}

extern "C" int pic2d_get_config(const pic2d_ctx *ctx, pic2d_config *out) {
  if (!ctx || !out) return PIC_E_ARG;
  *out = ctx->cfg;
  return 0;
}

extern "C" void *pic2d_stream(pic2d_ctx *ctx) { return ctx ? (void *)ctx->rt.stream : nullptr; }

extern "C" int pic2d_sync(pic2d_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  P2_HIP(ctx, hipStreamSynchronize(ctx->rt.stream));
  return 0;
}

extern "C" int pic2d_upload_field(pic2d_ctx *ctx, int which, const double *host) {
  if (!ctx || !host || which < 0 || which >= PIC2D_NFIELDS) return PIC_E_ARG;
  const int64_t n = ctx->g.plane;
  if (ctx->f64) {
    P2_HIP(ctx, hipMemcpy(plane<double>(ctx, which), host, n * sizeof(double),
                          hipMemcpyHostToDevice));
    return 0;
  }
  float *tmp = (float *)malloc(n * sizeof(float));
  if (!tmp) return err(ctx, PIC_E_NOMEM, "upload");
  for (int64_t i = 0; i < n; ++i) tmp[i] = (float)host[i];
  hipError_t e = hipMemcpy(plane<float>(ctx, which), tmp, n * sizeof(float),
                           hipMemcpyHostToDevice);
  free(tmp);
  P2_HIP(ctx, e);
  return 0;
}

extern "C" int pic2d_download_field(pic2d_ctx *ctx, int which, double *host) {
  if (!ctx || !host || which < 0 || which >= PIC2D_NFIELDS) return PIC_E_ARG;
  return download_plane(ctx, which, host);
}

extern "C" int pic2d_upload_particles(pic2d_ctx *ctx, int species, int64_t n,
                                      const double *x, const double *z,
                                      const double *vx, const double *vy,
                                      const double *vz) {
  if (!ctx || species < 0 || species > 1 || n < 0) return PIC_E_ARG;
  if (n > 0 && (!x || !z || !vx || !vy || !vz)) return PIC_E_ARG;
  const double *h[5] = {x, z, vx, vy, vz};
  P2_HIP(ctx, hipStreamSynchronize(ctx->rt.stream));
  return upload_particles(ctx, species, n, h);
}

extern "C" int pic2d_download_particles(pic2d_ctx *ctx, int species, int64_t cap,
                                        int64_t *n, double *x, double *z,
                                        double *vx, double *vy, double *vz) {
  if (!ctx || species < 0 || species > 1 || !n) return PIC_E_ARG;
  double *h[5] = {x, z, vx, vy, vz};
  return download_particles(ctx, species, cap, n, h);
}

extern "C" int pic2d_particle_count(pic2d_ctx *ctx, int species, int64_t *n) {
  if (!ctx || species < 0 || species > 1 || !n) return PIC_E_ARG;
  return count_live(ctx, species, n);
}

extern "C" int pic2d_load_particles(pic2d_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  return load_particles(ctx);
}

extern "C" int pic2d_renew_grid(pic2d_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  pic::timing_begin(&ctx->rt, PIC_T_OTHER);
  int rc = renew_grid(ctx);
  pic::timing_end(&ctx->rt, PIC_T_OTHER);
  return rc;
}

extern "C" int pic2d_inflow(pic2d_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  const uint32_t ev = ++ctx->inflow_event;
  int rc = 0;
  pic::timing_begin(&ctx->rt, PIC_T_INFLOW);
  for (int sp = 0; sp < 2 && !rc; ++sp) rc = inflow(ctx, sp, ev);
  pic::timing_end(&ctx->rt, PIC_T_INFLOW);
  return rc;
}

extern "C" int pic2d_renew_b_half(pic2d_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  pic::timing_begin(&ctx->rt, PIC_T_FIELD_B);
  int rc = renew_b_half(ctx);
  pic::timing_end(&ctx->rt, PIC_T_FIELD_B);
  return rc;
}

extern "C" int pic2d_halo(pic2d_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  pic::timing_begin(&ctx->rt, PIC_T_OTHER);
  int rc = halo(ctx);
  pic::timing_end(&ctx->rt, PIC_T_OTHER);
  return rc;
}

extern "C" int pic2d_push_deposit_j(pic2d_ctx *ctx, int species) {
  if (!ctx || species < 0 || species > 1) return PIC_E_ARG;
  pic::timing_begin(&ctx->rt, PIC_T_PUSH);
  int rc = push(ctx, species, 1, 0, 0, 0);
  pic::timing_end(&ctx->rt, PIC_T_PUSH);
  return rc;
}

extern "C" int pic2d_fold_current(pic2d_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  pic::timing_begin(&ctx->rt, PIC_T_OTHER);
  int rc = fold(ctx, PIC_JIX, 6);
  pic::timing_end(&ctx->rt, PIC_T_OTHER);
  return rc;
}

extern "C" int pic2d_external_current(pic2d_ctx *ctx, int64_t step) {
  if (!ctx || step < 0) return PIC_E_ARG;
  return external_current(ctx, step);
}

extern "C" int pic2d_drift_deposit_n(pic2d_ctx *ctx, int species) {
  if (!ctx || species < 0 || species > 1) return PIC_E_ARG;
  pic::timing_begin(&ctx->rt, PIC_T_PUSH);
  int rc = push(ctx, species, 0, 1, 0, 0);
  pic::timing_end(&ctx->rt, PIC_T_PUSH);
  return rc;
}

extern "C" int pic2d_fold_density(pic2d_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  pic::timing_begin(&ctx->rt, PIC_T_OTHER);
  int rc = fold(ctx, PIC_NI, 2);
  pic::timing_end(&ctx->rt, PIC_T_OTHER);
  return rc;
}

extern "C" int pic2d_renew_e(pic2d_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  pic::timing_begin(&ctx->rt, PIC_T_FIELD_E);
  int rc = renew_e(ctx);
  pic::timing_end(&ctx->rt, PIC_T_FIELD_E);
  return rc;
}

extern "C" int pic2d_poisson(pic2d_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  pic::timing_begin(&ctx->rt, PIC_T_FIELD_E);
  int rc = poisson(ctx);
  pic::timing_end(&ctx->rt, PIC_T_FIELD_E);
  return rc;
}

extern "C" int pic2d_particle_boundary(pic2d_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  int rc = 0;
  pic::timing_begin(&ctx->rt, PIC_T_OTHER);
  for (int sp = 0; sp < 2 && !rc; ++sp) rc = particle_boundary(ctx, sp);
  pic::timing_end(&ctx->rt, PIC_T_OTHER);
  return rc;
}

extern "C" int pic2d_sort(pic2d_ctx *ctx, int species) {
  if (!ctx || species < 0 || species > 1) return PIC_E_ARG;
  pic::timing_begin(&ctx->rt, PIC_T_SORT);
  int rc = sort(ctx, species);
  pic::timing_end(&ctx->rt, PIC_T_SORT);
  return rc;
}

extern "C" int pic2d_sort_stats(const pic2d_ctx *ctx, int64_t *tile, int64_t *radix) {
  if (!ctx) return PIC_E_ARG;
  if (tile) *tile = ctx->sort_tile_count;
  if (radix) *radix = ctx->sort_radix_count;
  return 0;
}

// pre-loop sequence (:584-626)
extern "C" int pic2d_begin(pic2d_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  int rc;
  if ((rc = renew_grid(ctx))) return rc;
  for (int sp = 0; sp < 2; ++sp)
    if ((rc = push(ctx, sp, 0, 0, 0, 1))) return rc;
  if ((rc = fold(ctx, PIC_NI, 2))) return rc;
  if ((rc = halo(ctx))) return rc;
  if (ctx->cfg.poisson_every > 0 && (rc = poisson(ctx))) return rc;
  if ((rc = halo(ctx))) return rc;
  ctx->step = 0;
  ctx->c_add = 0;
  ctx->c_sort = 0;
  return 0;
}

// the loop body of fdtd (:630-1014), particle loops A and B and the
// exchange fused into one pass per species.  The density fold moves ahead
// of the second renew_b4 + sr_grid: those read no density, and the sr_grid
// copy of the density guards happens again after renew_e4, so every value a
// later routine reads is the reference's.
extern "C" int pic2d_step(pic2d_ctx *ctx, int nsteps) {
  if (!ctx || nsteps < 0) return PIC_E_ARG;
  const pic2d_config &c = ctx->cfg;
  int rc;
  // an error leaves no sort permutation waiting for a pass
#define TRY(x)                              \
  do {                                      \
    if ((rc = (x))) {                       \
      (void)settle_sorts(ctx);              \
      return rc;                            \
    }                                       \
  } while (0)
  for (int s = 0; s < nsteps; ++s) {
    TRY(pic2d_renew_grid(ctx));
    if (c.inflow_every > 0 && ctx->c_add >= c.inflow_every) {
      TRY(pic2d_inflow(ctx));
      ctx->c_add = 0;
    }
    if (c.sort_every > 0 && ctx->c_sort >= c.sort_every) {
      // the FP32 form's tile sort leaves its permutation to the species'
      // fused pass below, which reads the particles through it (no copy);
      // only field work lies between them
      for (int sp = 0; sp < 2; ++sp) {
        pic::timing_begin(&ctx->rt, PIC_T_SORT);
        rc = ctx->f64 ? sort(ctx, sp) : sort_for_pass(ctx, sp);
        pic::timing_end(&ctx->rt, PIC_T_SORT);
        TRY(rc);
      }
      ctx->c_sort = 0;
    }
    TRY(pic2d_renew_b_half(ctx));
    TRY(pic2d_halo(ctx));
    for (int sp = 0; sp < 2; ++sp) {
      const int tclass = ctx->pending[sp] ? PIC_T_PUSH_SORT : PIC_T_PUSH;
      pic::timing_begin(&ctx->rt, tclass);
      rc = push(ctx, sp, 1, 1, 1, 0);
      pic::timing_end(&ctx->rt, tclass);
      TRY(rc);
    }
    TRY(pic2d_fold_current(ctx));
    TRY(pic2d_external_current(ctx, ctx->step));
    TRY(pic2d_fold_density(ctx));
    TRY(pic2d_renew_b_half(ctx));
    TRY(pic2d_halo(ctx));
    TRY(pic2d_renew_e(ctx));
    TRY(pic2d_halo(ctx));
    if (c.poisson_every > 0 && ctx->step % c.poisson_every == 0) {
      TRY(pic2d_poisson(ctx));
      TRY(pic2d_halo(ctx));
    }
    ctx->step++;
    ctx->c_add++;
    ctx->c_sort++;
  }
#undef TRY
  return 0;
}

extern "C" int pic2d_step_index(const pic2d_ctx *ctx, int64_t *step) {
  if (!ctx || !step) return PIC_E_ARG;
  *step = ctx->step;
  return 0;
}

extern "C" int pic2d_diagnose(pic2d_ctx *ctx, pic2d_diag *out) {
  if (!ctx || !out) return PIC_E_ARG;
  memset(out, 0, sizeof(*out));
  int rc = diagnose_fields(ctx, out);
  if (rc) return rc;
  return diagnose_particles(ctx, out);
}

// probe columns of fdtd (:929-1006): the few grid values are copied to the
// host and combined there with the reference's expressions
extern "C" int pic2d_probe(pic2d_ctx *ctx, int k, int m, double out[8]) {
  if (!ctx || !out) return PIC_E_ARG;
  const Geom2 &g = ctx->g;
  if (k < 1 || m < 1 || k > g.SX - 2 || m > g.SZ - 2) return PIC_E_ARG;
  // planes ex..b0z, ni, ne at (k,m), (k-1,m), (k+1,m), (k,m-1), (k,m+1)
  const int pl[11] = {PIC_EX, PIC_EY, PIC_EZ, PIC_BX, PIC_BY, PIC_BZ,
                      PIC_B0X, PIC_B0Y, PIC_B0Z, PIC_NI, PIC_NE};
  const int64_t off[5] = {0, -(int64_t)g.SZ, (int64_t)g.SZ, -1, 1};
  double v[11][5];
  const size_t esz = ctx->f64 ? sizeof(double) : sizeof(float);
  for (int a = 0; a < 11; ++a)
    for (int b = 0; b < 5; ++b) {
      const int64_t i = g.idx(k, m) + off[b];
      const char *src = (const char *)ctx->field_mem + ((int64_t)pl[a] * g.plane + i) * esz;
      if (ctx->f64) {
        P2_HIP(ctx, hipMemcpy(&v[a][b], src, esz, hipMemcpyDeviceToHost));
      } else {
        float f;
        P2_HIP(ctx, hipMemcpy(&f, src, esz, hipMemcpyDeviceToHost));
        v[a][b] = f;
      }
    }
  const pic2d_config &c = ctx->cfg;
  const double dx = c.dx;
  enum { EX, EY, EZ, BX, BY, BZ, B0X, B0Y, B0Z, NI, NE };
  enum { C0, KM, KP, MM, MP };
  for (int a = 0; a < 6; ++a) out[a] = v[a][C0];
  out[6] = fabs(v[BX][C0] - v[B0X][C0] - v[BX][KM] + v[B0X][KM] + v[BZ][C0] -
                v[B0Z][C0] - v[BZ][MM] + v[B0Z][MM]) /
           dx /
           sqrt(pow(v[BX][C0] + v[BX][KM], 2) / 4. + pow(v[BY][C0], 2) +
                pow(v[BZ][C0] + v[BZ][MM], 2) / 4.);
  out[7] = fabs((v[EX][KP] - v[EX][C0] + v[EZ][MP] - v[EZ][C0]) / dx -
                c.q * (v[NI][C0] - v[NE][C0]) / c.eps0) /
           sqrt(pow(v[EX][KP] + v[EX][C0], 2) / 4. + pow(v[EY][C0], 2) +
                pow(v[EZ][MP] + v[EZ][C0], 2) / 4.);
  return 0;
}

extern "C" int pic2d_timing_enable(pic2d_ctx *ctx, int on) {
  if (!ctx) return PIC_E_ARG;
  return pic_timing_enable(&ctx->rt, on);
}

extern "C" int pic2d_timing_get(pic2d_ctx *ctx, int cls, double *ms,
                                int64_t *launches) {
  if (!ctx) return PIC_E_ARG;
  return pic_timing_get(&ctx->rt, cls, ms, launches);
}

extern "C" int pic2d_timing_reset(pic2d_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  return pic_timing_reset(&ctx->rt);
}

// C-ABI entry points: context lifetime, data movement, the step routines in
// fdtd() order, the step loop and kernel timing.  Host code only; kernels
// live in pic_fp64.hip / pic_fp32.hip.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include "pic_internal.h"
#include "pic_tiles.h"

using namespace pic;

namespace pic {

int set_error(pic_ctx *ctx, int code, const char *fmt, ...) {
  if (ctx) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(ctx->err, sizeof(ctx->err), fmt, ap);
    va_end(ap);
  }
  return code;
}

StepConst step_const(const pic_ctx *ctx, int sp) {
  const pic_config &c = ctx->cfg;
  StepConst k;
  k.dx = c.dx;
  k.dt = c.dt;
  k.c = c.c;
  k.q = c.q;
  k.mass = sp == PIC_ION ? c.mass[0] : -c.mass[1];
  k.weight = c.weight[sp];
  k.vdrift_z = c.vdrift[2];
  k.imf_x = c.imf[0];
  k.imf_y = c.imf[1];
  for (int a = 0; a < 3; ++a) {
    k.blo[a] = c.ballistic_lo[a];
    k.bhi[a] = c.ballistic_hi[a];
  }
  k.periodic[0] = c.periodic[0];
  k.periodic[1] = c.periodic[1];
  return k;
}

// mode 1 times every class, mode 2 the particle passes only (each event
// record is one more packet in the stream: the passes alone keep the
// step's other launches back to back)
static bool timed(const pic_ctx *ctx, int which) {
  return ctx->timing == 1 ||
         (ctx->timing == 2 && (which == PIC_T_PUSH || which == PIC_T_PUSH_SORT));
}

void timing_begin(pic_ctx *ctx, int which) {
  if (!timed(ctx, which)) return;
  if (ctx->tpending == pic_ctx::kEventPool) timing_flush(ctx);
  (void)hipEventRecord(ctx->tev[0][ctx->tpending], ctx->stream);
}

void timing_end(pic_ctx *ctx, int which) {
  if (!timed(ctx, which)) return;
  (void)hipEventRecord(ctx->tev[1][ctx->tpending], ctx->stream);
  ctx->tclass[ctx->tpending] = which;
  ctx->tpending++;
}

int timing_flush(pic_ctx *ctx) {
  if (ctx->tpending == 0) return 0;
  PIC_HIP(ctx, hipEventSynchronize(ctx->tev[1][ctx->tpending - 1]));
  for (int i = 0; i < ctx->tpending; ++i) {
    float ms = 0.f;
    PIC_HIP(ctx, hipEventElapsedTime(&ms, ctx->tev[0][i], ctx->tev[1][i]));
    ctx->t_ms[ctx->tclass[i]] += ms;
    ctx->t_n[ctx->tclass[i]] += 1;
  }
  ctx->tpending = 0;
  return 0;
}

}  // namespace pic

// ---------------------------------------------------------------------------
// configuration

static void fill_common(pic_config *c, int n, int ppc, double mi, double V,
                        double standoff, double alpha) {
  memset(c, 0, sizeof(*c));
  c->nx = c->ny = c->nz = n;
  c->ppc = ppc;
  c->dx = 1.0;
  // The (2,4) stencil's 3D leapfrog limit is c dt/dx <= 6/(7 sqrt 3) =
  // 0.4949; dt = 0.45 keeps the presets stable.
  c->dt = 0.45;
  c->c = 1.0;
  c->eps0 = 1.0;
  c->mu0 = 1.0;
  c->q = 1.0;
  c->mass[0] = mi;
  c->mass[1] = 1.0;
  // omega_pe dt = 0.1: n0 = (0.1/dt)^2 eps0 m_e / q^2
  const double wpe = 0.1 / c->dt;
  const double n0 = wpe * wpe * c->eps0 * c->mass[1] / (c->q * c->q);
  c->weight[0] = c->weight[1] = n0 / ppc;
  c->vth[0] = 0.005;
  c->vth[1] = 0.05;
  c->vdrift[2] = V;
  // standoff L = (mu0 M^2 / (8 pi^2 n m_i V^2))^(1/6) (reference :226)
  const double Pi = 3.14159265358979323846;
  c->dipole_m =
      sqrt(pow(standoff, 6.0) * 8.0 * Pi * Pi * n0 * mi * V * V / c->mu0);
  c->dipole_alpha = alpha;
  c->dipole_soft = 2.0;
  c->dipole_rmax = 210.0;
  const int A = n / 8 > 2 ? n / 8 : 2;
  c->absorb[2] = A;
  // reference a = 1e-6 S/m at dt = 1.5 ns: sigma dt / eps0 = 1.694e-4
  c->sigma_amp = 1.6941e-4 / c->dt;
  const int bw = A + 2;
  for (int a = 0; a < 2; ++a) {
    c->ballistic_lo[a] = -1.0e30;
    c->ballistic_hi[a] = 1.0e30;
  }
  c->ballistic_lo[2] = (2 + bw) * c->dx;
  c->ballistic_hi[2] = (n + 2 - bw) * c->dx;
  c->periodic[0] = c->periodic[1] = 1;
  c->inflow_every = 5;
  c->sort_every = 20;
  c->refill_far = 1;
  c->precision = PIC_FP32;
  c->deposit = PIC_DEPOSIT_ATOMIC;
  c->tile[0] = c->tile[1] = 4;
  c->tile[2] = 8;
  c->poisson_every = 0;  // PHI() is compiled out in the reference (:107)
  c->phi_margin = bw;    // Absorb_grid3 = Absorb_grid2, the ballistic width
}

extern "C" size_t pic_sizeof_config(void) { return sizeof(pic_config); }
extern "C" size_t pic_sizeof_diag(void) { return sizeof(pic_diag); }

extern "C" int pic_config_preset(int preset, pic_config *cfg) {
  if (!cfg) return PIC_E_ARG;
  const double Pi = 3.14159265358979323846;
  switch (preset) {
    case 0:  // tiny grid for tests
      fill_common(cfg, 8, 4, 100.0, 0.1, 3.0, Pi / 2);
      cfg->nz = 16;
      cfg->absorb[2] = 2;
      cfg->ballistic_lo[2] = 6.0;
      cfg->ballistic_hi[2] = 14.0;
      cfg->precision = PIC_FP64;
      cfg->tile[0] = cfg->tile[1] = cfg->tile[2] = 0;  // reference order
      break;
    case 1:  // 32^3 x 16 ppc, standoff 8
      fill_common(cfg, 32, 16, 100.0, 0.1, 8.0, Pi / 2);
      cfg->precision = PIC_FP64;
      cfg->tile[0] = cfg->tile[1] = cfg->tile[2] = 0;  // reference order
      break;
    case 2:  // 128^3 x 64 ppc, standoff 32
      fill_common(cfg, 128, 64, 100.0, 0.1, 32.0, Pi / 2);
      // 8 x 4 x 4 sort tiles: of the shapes whose LDS window fits (4 x 4 x 8,
      // 4 x 4 x 4, 2 x 4 x 8, 2 x 2 x 16, 8 x 4 x 4) the fastest FP32 pass
      // on this workload, 2 % ahead of 4 x 4 x 8
      cfg->tile[0] = 8;
      cfg->tile[1] = cfg->tile[2] = 4;
      break;
    case 4:  // 64^3 x 32 ppc (adjoint workload)
      fill_common(cfg, 64, 32, 100.0, 0.1, 16.0, Pi / 2);
      break;
    case 5:  // 256^3 x 32 ppc, m_i/m_e = 400, drift 0.05, tilt 30 deg
      fill_common(cfg, 256, 32, 400.0, 0.05, 32.0, Pi / 3);
      break;
    default:
      return PIC_E_ARG;
  }
  cfg->seed = 1000 + (uint64_t)preset;
  return 0;
}

static int validate(const pic_config *c, char *err, size_t n) {
  if (c->nx < 4 || c->ny < 4 || c->nz < 8) {
    snprintf(err, n, "grid too small (%d,%d,%d)", c->nx, c->ny, c->nz);
    return PIC_E_ARG;
  }
  const int64_t cells = (int64_t)(c->nx + 4) * (c->ny + 4) * (c->nz + 4);
  if (cells > ((int64_t)1 << 31)) {
    snprintf(err, n, "grid too large");
    return PIC_E_ARG;
  }
  if (c->ppc < 1 || !(c->dx > 0) || !(c->dt > 0) || !(c->c > 0)) {
    snprintf(err, n, "ppc, dx, dt and c must be positive");
    return PIC_E_ARG;
  }
  // a particle must not cross more than one cell per half step (|v| < c)
  if (c->c * c->dt >= c->dx) {
    snprintf(err, n, "c*dt must be below dx");
    return PIC_E_ARG;
  }
  if (c->precision != PIC_FP64 && c->precision != PIC_FP32) {
    snprintf(err, n, "precision must be PIC_FP64 or PIC_FP32");
    return PIC_E_ARG;
  }
  if (c->deposit != PIC_DEPOSIT_ATOMIC && c->deposit != PIC_DEPOSIT_DETERMINISTIC) {
    snprintf(err, n, "unknown deposit mode");
    return PIC_E_ARG;
  }
  if (c->precision == PIC_FP32 &&
      (c->nx + 4 > 1023 || c->ny + 4 > 1023 || c->nz + 4 > 1023)) {
    snprintf(err, n, "FP32 form packs cells in 10 bits per axis: n <= 1019");
    return PIC_E_ARG;
  }
  if (c->inflow_every < 1 || c->sort_every < 1) {
    snprintf(err, n, "cadences must be >= 1");
    return PIC_E_ARG;
  }
  if (c->poisson_every < 0 || c->phi_margin < 0) {
    snprintf(err, n, "poisson_every and phi_margin must be >= 0");
    return PIC_E_ARG;
  }
  if (c->poisson_every > 0)
    for (int a = 0; a < 3; ++a) {
      const int N = a == 0 ? c->nx : a == 1 ? c->ny : c->nz;
      if (N < 2 || (N & (N - 1)) != 0) {
        snprintf(err, n, "the Poisson correction needs power-of-two sizes (axis %d: %d)",
                 a, N);
        return PIC_E_ARG;
      }
    }
  for (int a = 0; a < 3; ++a) {
    const int N = a == 0 ? c->nx : a == 1 ? c->ny : c->nz;
    if (c->tile[a] < 0 || (c->tile[a] > 0 && N % c->tile[a] != 0)) {
      snprintf(err, n, "tile[%d] = %d must divide the grid size %d", a,
               c->tile[a], N);
      return PIC_E_ARG;
    }
    if (c->absorb[a] < 0 || 2 * c->absorb[a] > N) {
      snprintf(err, n, "absorber width out of range on axis %d", a);
      return PIC_E_ARG;
    }
  }
  return 0;
}

// B0 of a point dipole at a staggered point (external_current7 :4579-4634)
static double dipole_b0(const pic_config *c, double x, double y, double z,
                        int comp) {
  double r = sqrt(x * x + y * y + z * z);
  const double Pi = 3.14159265358979323846;
  const double sa = sin(c->dipole_alpha), ca = cos(c->dipole_alpha);
  if (c->dipole_soft > 0.0 && r < c->dipole_soft * c->dx) {
    const double s = c->dipole_soft * c->dx / r;
    x *= s;
    y *= s;
    z *= s;
    r = sqrt(x * x + y * y + z * z);
  }
  const double M = c->dipole_m;
  double v;
  if (comp == 0)
    v = c->mu0 * M * (3. * x * y * sa + 3. * x * z * ca) / (4. * Pi * pow(r, 5));
  else if (comp == 1)
    v = c->mu0 * M * (3. * y * y * sa + 3. * y * z * ca - r * r * sa) /
        (4. * Pi * pow(r, 5));
  else
    v = c->mu0 * M * (3. * y * z * sa + 3. * z * z * ca - r * r * ca) /
        (4. * Pi * pow(r, 5));
  return v + c->imf[comp];
}

// sigma depth on the 2x lattice (init_grid :332-366); the layer of width A
// is mapped onto the reference's 32-cell profile
static double sigma_depth(int i2, int N, int A) {
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

static int host_fields_b0(const pic_ctx *ctx, double *b0) {
  const pic_config *c = &ctx->cfg;
  const Geom &g = ctx->geom;
  const int64_t nc = (int64_t)g.SX * g.SY * g.SZ;
  const double dx = c->dx;
  const int cx = (g.NX + 4) / 2, cy = (g.NY + 4) / 2, cz = (g.NZ + 4) / 2;
  for (int k = 0; k < g.SX; ++k)
    for (int l = 0; l < g.SY; ++l)
      for (int m = 0; m < g.SZ; ++m) {
        const int64_t i = g.idx(k, l, m);
        const double x = (k - cx) * dx, y = (l - cy) * dx, z = (m - cz) * dx;
        const double r = sqrt(x * x + y * y + z * z);
        if (r < c->dipole_rmax * dx) {
          b0[i] = dipole_b0(c, (k - cx + 0.5) * dx, (l - cy) * dx,
                            (m - cz) * dx, 0);
          b0[nc + i] = dipole_b0(c, (k - cx) * dx, (l - cy + 0.5) * dx,
                                 (m - cz) * dx, 1);
          b0[2 * nc + i] = dipole_b0(c, (k - cx) * dx, (l - cy) * dx,
                                     (m - cz + 0.5) * dx, 2);
        } else {
          b0[i] = c->imf[0];
          b0[nc + i] = c->imf[1];
          b0[2 * nc + i] = c->imf[2];
        }
      }
  return 0;
}

static void free_all(pic_ctx *ctx) {
  poisson_free(ctx);
  f64::tsort_free(ctx);
  if (ctx->field_mem) (void)hipFree(ctx->field_mem);
  if (ctx->sig_mem) (void)hipFree(ctx->sig_mem);
  for (int s = 0; s < 2; ++s)
    for (int b = 0; b < 2; ++b)
      if (ctx->pbuf[s][b]) (void)hipFree(ctx->pbuf[s][b]);
  if (ctx->keys) (void)hipFree(ctx->keys);
  if (ctx->keys_alt) (void)hipFree(ctx->keys_alt);
  if (ctx->idx) (void)hipFree(ctx->idx);
  if (ctx->idx_alt) (void)hipFree(ctx->idx_alt);
  if (ctx->flags) (void)hipFree(ctx->flags);
  if (ctx->prim_tmp) (void)hipFree(ctx->prim_tmp);
  if (ctx->dev_counter) (void)hipFree(ctx->dev_counter);
  for (int sp = 0; sp < 2; ++sp)
    if (ctx->tile_start[sp]) (void)hipFree(ctx->tile_start[sp]);
  if (ctx->host_counter) (void)hipHostFree(ctx->host_counter);
  if (ctx->diag_partial) (void)hipFree(ctx->diag_partial);
  if (ctx->thrust_acc) (void)hipFree(ctx->thrust_acc);
  if (ctx->thrust_part) (void)hipFree(ctx->thrust_part);
  if (ctx->thrust_bprev) (void)hipFree(ctx->thrust_bprev);
  for (int i = 0; i < pic_ctx::kEventPool; ++i)
    for (int j = 0; j < 2; ++j)
      if (ctx->tev[j][i]) (void)hipEventDestroy(ctx->tev[j][i]);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  free(ctx->b0_host);
  if (ctx->cfg.precision == PIC_FP32) f32::free_state(ctx);
  adj_free(ctx);
}

// ---------------------------------------------------------------------------
// lifetime

extern "C" int pic_create(const pic_config *cfg, pic_ctx **out) {
  if (!cfg || !out) return PIC_E_ARG;
  *out = nullptr;
  char err[256];
  int rc = validate(cfg, err, sizeof(err));
  if (rc) {
    fprintf(stderr, "pic_create: %s\n", err);
    return rc;
  }
  pic_ctx *ctx = new (std::nothrow) pic_ctx();
  if (!ctx) return PIC_E_NOMEM;
  ctx->err[0] = 0;
  ctx->cfg = *cfg;
  Geom &g = ctx->geom;
  g.NX = cfg->nx;
  g.NY = cfg->ny;
  g.NZ = cfg->nz;
  g.SX = g.NX + 4;
  g.SY = g.NY + 4;
  g.SZ = g.NZ + 4;
  const int64_t nc = (int64_t)g.SX * g.SY * g.SZ;
  g.plane = div_up(nc, 64) * 64;
  const bool f64 = cfg->precision == PIC_FP64;
  const size_t esz = f64 ? sizeof(double) : sizeof(float);

#define CREATE_HIP(call)                                             \
  do {                                                               \
    hipError_t e_ = (call);                                          \
    if (e_ != hipSuccess) {                                          \
      fprintf(stderr, "pic_create: %s: %s\n", #call,                 \
              hipGetErrorString(e_));                                \
      free_all(ctx);                                                 \
      delete ctx;                                                    \
      return e_ == hipErrorOutOfMemory ? PIC_E_NOMEM : PIC_E_HIP;    \
    }                                                                \
  } while (0)

  CREATE_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  // this device's static scratch for the allocation-free entry points
  if ((rc = noalloc_init(ctx))) {
    free_all(ctx);
    delete ctx;
    return rc;
  }
  ctx->field_bytes = (size_t)PIC_NFIELDS * g.plane * esz;
  CREATE_HIP(hipMalloc(&ctx->field_mem, ctx->field_bytes));
  CREATE_HIP(hipMemsetAsync(ctx->field_mem, 0, ctx->field_bytes, ctx->stream));

  // sigma ingredients: depth^4 and depth^12 per axis on the 2x lattice
  const int smax = 2 * (g.SX > g.SY ? (g.SX > g.SZ ? g.SX : g.SZ)
                                    : (g.SY > g.SZ ? g.SY : g.SZ));
  ctx->sig_stride = smax;
  double *sig = (double *)calloc((size_t)6 * smax, sizeof(double));
  if (!sig) {
    free_all(ctx);
    delete ctx;
    return PIC_E_NOMEM;
  }
  const int Ns[3] = {g.NX, g.NY, g.NZ}, Ss[3] = {g.SX, g.SY, g.SZ};
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < 2 * Ss[a]; ++i) {
      const double d = sigma_depth(i, Ns[a], cfg->absorb[a]);
      sig[(2 * a) * smax + i] = pow(d, 4);
      sig[(2 * a + 1) * smax + i] = pow(d, 12);
    }
  CREATE_HIP(hipMalloc(&ctx->sig_mem, (size_t)6 * smax * sizeof(double)));
  CREATE_HIP(hipMemcpy(ctx->sig_mem, sig, (size_t)6 * smax * sizeof(double),
                       hipMemcpyHostToDevice));
  free(sig);

  // static dipole field, B = B0 at start (init_grid + external_current7)
  ctx->b0_host = (double *)malloc((size_t)3 * nc * sizeof(double));
  if (!ctx->b0_host) {
    free_all(ctx);
    delete ctx;
    return PIC_E_NOMEM;
  }
  host_fields_b0(ctx, ctx->b0_host);
  CREATE_HIP(hipStreamSynchronize(ctx->stream));
  for (int comp = 0; comp < 3; ++comp) {
    const double *src = ctx->b0_host + comp * nc;
    if (f64) {
      double *base = (double *)ctx->field_mem;
      CREATE_HIP(hipMemcpy(base + (PIC_B0X + comp) * g.plane, src,
                           nc * sizeof(double), hipMemcpyHostToDevice));
      CREATE_HIP(hipMemcpy(base + (PIC_BX + comp) * g.plane, src,
                           nc * sizeof(double), hipMemcpyHostToDevice));
    } else {
      float *tmp = (float *)malloc(nc * sizeof(float));
      if (!tmp) {
        free_all(ctx);
        delete ctx;
        return PIC_E_NOMEM;
      }
      for (int64_t i = 0; i < nc; ++i) tmp[i] = (float)src[i];
      float *base = (float *)ctx->field_mem;
      hipError_t e = hipMemcpy(base + (PIC_B0X + comp) * g.plane, tmp,
                               nc * sizeof(float), hipMemcpyHostToDevice);
      free(tmp);
      CREATE_HIP(e);
      // B planes hold B - B0 = 0 in the FP32 form
    }
  }

  // particle storage
  const int64_t ncell_in = (int64_t)g.NX * g.NY * g.NZ;
  int64_t cap = cfg->capacity;
  if (cap <= 0)
    cap = ncell_in * cfg->ppc + ncell_in * cfg->ppc / 4 +
          (int64_t)4 * g.NX * g.NY * 2 * cfg->ppc + 4096;
  cap = div_up(cap, 256) * 256;
  if (cap >= ((int64_t)1 << 32) - 1) {
    fprintf(stderr, "pic_create: capacity %lld exceeds 32-bit indexing\n",
            (long long)cap);
    free_all(ctx);
    delete ctx;
    return PIC_E_ARG;
  }
  ctx->cap = cap;
  const size_t pbytes = (size_t)cap * (f64 ? 6 * sizeof(double)
                                           : sizeof(int32_t) + 6 * sizeof(float));
  for (int s = 0; s < 2; ++s)
    for (int b = 0; b < 2; ++b) CREATE_HIP(hipMalloc(&ctx->pbuf[s][b], pbytes));
  CREATE_HIP(hipMalloc(&ctx->keys, cap * sizeof(uint32_t)));
  CREATE_HIP(hipMalloc(&ctx->keys_alt, cap * sizeof(uint32_t)));
  CREATE_HIP(hipMalloc(&ctx->idx, cap * sizeof(uint32_t)));
  CREATE_HIP(hipMalloc(&ctx->idx_alt, cap * sizeof(uint32_t)));
  CREATE_HIP(hipMalloc(&ctx->flags, cap));
  {
    const int nt = tile_geom(ctx).ntiles();
    for (int sp = 0; sp < 2; ++sp) {
      CREATE_HIP(hipMalloc(&ctx->tile_start[sp], (size_t)(nt + 1) * sizeof(uint32_t)));
      CREATE_HIP(hipMemset(ctx->tile_start[sp], 0, (size_t)(nt + 1) * sizeof(uint32_t)));
    }
  }
  CREATE_HIP(hipMalloc(&ctx->dev_counter, 8 * sizeof(int64_t)));
  CREATE_HIP(hipHostMalloc(&ctx->host_counter, 8 * sizeof(int64_t)));
  for (int i = 0; i < pic_ctx::kEventPool; ++i)
    for (int j = 0; j < 2; ++j) CREATE_HIP(hipEventCreate(&ctx->tev[j][i]));
  CREATE_HIP(hipStreamSynchronize(ctx->stream));
#undef CREATE_HIP
  if (cfg->poisson_every > 0) {
    int rc2 = poisson_init(ctx);
    if (rc2) {
      fprintf(stderr, "pic_create: %s\n", ctx->err);
      free_all(ctx);
      delete ctx;
      return rc2;
    }
  }
  {
    int rc2 = f64 ? f64::tsort_setup(ctx) : f32::init_state(ctx);
    if (rc2) {
      fprintf(stderr, "pic_create: %s\n", ctx->err);
      free_all(ctx);
      delete ctx;
      return rc2;
    }
  }
  *out = ctx;
  return 0;
}

extern "C" void pic_destroy(pic_ctx *ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  free_all(ctx);
  delete ctx;
}

extern "C" const char *pic_last_error(const pic_ctx *ctx) {
  return ctx ? ctx->err : "null context";
}

extern "C" int pic_get_config(const pic_ctx *ctx, pic_config *cfg) {
  if (!ctx || !cfg) return PIC_E_ARG;
  *cfg = ctx->cfg;
  return 0;
}

extern "C" void *pic_stream(pic_ctx *ctx) { return ctx ? ctx->stream : nullptr; }

extern "C" int pic_sync(pic_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// data movement

static inline bool is_f64(const pic_ctx *ctx) {
  return ctx->cfg.precision == PIC_FP64;
}

extern "C" int pic_upload_field(pic_ctx *ctx, int which, const double *host) {
  if (!ctx || !host || which < 0 || which >= PIC_NFIELDS) return PIC_E_ARG;
  const Geom &g = ctx->geom;
  const int64_t nc = (int64_t)g.SX * g.SY * g.SZ;
  if (is_f64(ctx)) {
    double *dst = (double *)ctx->field_mem + which * g.plane;
    PIC_HIP(ctx, hipMemcpyAsync(dst, host, nc * sizeof(double),
                                hipMemcpyHostToDevice, ctx->stream));
    PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
  }
  PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  float *tmp = (float *)malloc(nc * sizeof(float));
  if (!tmp) return set_error(ctx, PIC_E_NOMEM, "upload: host buffer");
  if (which >= PIC_BX && which <= PIC_BZ) {
    const double *b0 = ctx->b0_host + (which - PIC_BX) * nc;
    for (int64_t i = 0; i < nc; ++i) tmp[i] = (float)(host[i] - b0[i]);
  } else {
    for (int64_t i = 0; i < nc; ++i) tmp[i] = (float)host[i];
    if (which >= PIC_B0X && which <= PIC_B0Z)
      memcpy(ctx->b0_host + (which - PIC_B0X) * nc, host, nc * sizeof(double));
  }
  float *dst = (float *)ctx->field_mem + which * g.plane;
  hipError_t e = hipMemcpy(dst, tmp, nc * sizeof(float), hipMemcpyHostToDevice);
  free(tmp);
  if (e != hipSuccess)
    return set_error(ctx, PIC_E_HIP, "upload: %s", hipGetErrorString(e));
  if ((which >= PIC_BX && which <= PIC_BZ) || (which >= PIC_B0X && which <= PIC_B0Z))
    return f32::refresh_total_b(ctx);
  return 0;
}

extern "C" int pic_download_field(pic_ctx *ctx, int which, double *host) {
  if (!ctx || !host || which < 0 || which > PIC_PHI) return PIC_E_ARG;
  const Geom &g = ctx->geom;
  const int64_t nc = (int64_t)g.SX * g.SY * g.SZ;
  if (which == PIC_PHI) return poisson_download_phi(ctx, host);
  if (!is_f64(ctx)) {
    int rc = f32::finish_deposit(ctx);
    if (rc) return rc;
  }
  PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (is_f64(ctx)) {
    const double *src = (const double *)ctx->field_mem + which * g.plane;
    PIC_HIP(ctx, hipMemcpy(host, src, nc * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
  }
  float *tmp = (float *)malloc(nc * sizeof(float));
  if (!tmp) return set_error(ctx, PIC_E_NOMEM, "download: host buffer");
  const float *src = (const float *)ctx->field_mem + which * g.plane;
  hipError_t e = hipMemcpy(tmp, src, nc * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess) {
    if (which >= PIC_BX && which <= PIC_BZ) {
      const double *b0 = ctx->b0_host + (which - PIC_BX) * nc;
      for (int64_t i = 0; i < nc; ++i) host[i] = b0[i] + (double)tmp[i];
    } else {
      for (int64_t i = 0; i < nc; ++i) host[i] = (double)tmp[i];
    }
  }
  free(tmp);
  if (e != hipSuccess)
    return set_error(ctx, PIC_E_HIP, "download: %s", hipGetErrorString(e));
  return 0;
}

extern "C" int pic_upload_particles(pic_ctx *ctx, int species, int64_t n,
                                    const double *x, const double *y,
                                    const double *z, const double *vx,
                                    const double *vy, const double *vz) {
  if (!ctx || species < 0 || species > 1 || n < 0) return PIC_E_ARG;
  if (n > 0 && (!x || !y || !z || !vx || !vy || !vz)) return PIC_E_ARG;
  if (n > ctx->cap)
    return set_error(ctx, PIC_E_CAPACITY, "upload: %lld > capacity %lld",
                     (long long)n, (long long)ctx->cap);
  const double *h[6] = {x, y, z, vx, vy, vz};
  if (!is_f64(ctx)) return f32::upload_particles(ctx, species, n, h);
  Particles64 P = p64(ctx, species, ctx->cur[species]);
  double *d[6] = {P.x, P.y, P.z, P.vx, P.vy, P.vz};
  for (int a = 0; a < 6; ++a)
    if (n > 0)
      PIC_HIP(ctx, hipMemcpyAsync(d[a], h[a], n * sizeof(double),
                                  hipMemcpyHostToDevice, ctx->stream));
  PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->n[species] = n;
  ctx->sorted_end[species] = 0;
  return 0;
}

extern "C" int pic_download_particles(pic_ctx *ctx, int species, int64_t cap,
                                      int64_t *n, double *x, double *y,
                                      double *z, double *vx, double *vy,
                                      double *vz) {
  if (!ctx || !n || species < 0 || species > 1) return PIC_E_ARG;
  double *h[6] = {x, y, z, vx, vy, vz};
  if (!is_f64(ctx)) return f32::download_particles(ctx, species, cap, n, h);
  const int64_t ns = ctx->n[species];
  Particles64 P = p64(ctx, species, ctx->cur[species]);
  double *d[6] = {P.x, P.y, P.z, P.vx, P.vy, P.vz};
  double *tmp = (double *)malloc((size_t)(ns > 0 ? ns : 1) * 6 * sizeof(double));
  if (!tmp) return set_error(ctx, PIC_E_NOMEM, "download: host buffer");
  hipError_t e = hipStreamSynchronize(ctx->stream);
  for (int a = 0; a < 6 && e == hipSuccess; ++a)
    if (ns > 0)
      e = hipMemcpy(tmp + a * ns, d[a], ns * sizeof(double),
                    hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    free(tmp);
    return set_error(ctx, PIC_E_HIP, "download: %s", hipGetErrorString(e));
  }
  int64_t live = 0;
  for (int64_t i = 0; i < ns; ++i)
    if (!isnan(tmp[2 * ns + i])) live++;
  *n = live;
  if (cap < live) {
    free(tmp);
    return set_error(ctx, PIC_E_CAPACITY, "download: %lld live > cap %lld",
                     (long long)live, (long long)cap);
  }
  int64_t w = 0;
  for (int64_t i = 0; i < ns; ++i) {
    if (isnan(tmp[2 * ns + i])) continue;
    for (int a = 0; a < 6; ++a) h[a][w] = tmp[a * ns + i];
    w++;
  }
  free(tmp);
  return 0;
}

extern "C" int pic_particle_count(pic_ctx *ctx, int species, int64_t *n) {
  if (!ctx || !n || species < 0 || species > 1) return PIC_E_ARG;
  return is_f64(ctx) ? f64::count_live(ctx, species, n)
                     : f32::count_live(ctx, species, n);
}

// ---------------------------------------------------------------------------
// step routines

#define DISPATCH(fn, ...) \
  (is_f64(ctx) ? f64::fn(__VA_ARGS__) : f32::fn(__VA_ARGS__))

extern "C" int pic_load_particles(pic_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  return DISPATCH(load_particles, ctx);
}

extern "C" int pic_renew_grid(pic_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  timing_begin(ctx, PIC_T_OTHER);
  int rc = DISPATCH(renew_grid, ctx);
  if (!rc && ctx->phi) rc = poisson_zero_phi(ctx);  // renew_grid zeroes phi (:2552)
  timing_end(ctx, PIC_T_OTHER);
  return rc;
}

extern "C" int pic_poisson(pic_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  if (!ctx->phi)
    return set_error(ctx, PIC_E_STATE, "poisson: disabled (poisson_every = 0)");
  timing_begin(ctx, PIC_T_FIELD_E);
  int rc = poisson(ctx);
  timing_end(ctx, PIC_T_FIELD_E);
  if (rc) return rc;
  return pic_halo_e(ctx);
}

extern "C" int pic_renew_b_half(pic_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  int rc;
  if (!is_f64(ctx) && (rc = f32::finish_deposit(ctx))) return rc;
  timing_begin(ctx, PIC_T_FIELD_B);
  rc = DISPATCH(renew_b_half, ctx);
  timing_end(ctx, PIC_T_FIELD_B);
  return rc;
}

extern "C" int pic_halo_b(pic_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  timing_begin(ctx, PIC_T_OTHER);
  int rc = DISPATCH(halo, ctx, PIC_BX);
  timing_end(ctx, PIC_T_OTHER);
  return rc;
}

extern "C" int pic_halo_e(pic_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  timing_begin(ctx, PIC_T_OTHER);
  int rc = DISPATCH(halo, ctx, PIC_EX);
  timing_end(ctx, PIC_T_OTHER);
  return rc;
}

extern "C" int pic_push_deposit_j(pic_ctx *ctx, int species) {
  if (!ctx || species < 0 || species > 1) return PIC_E_ARG;
  timing_begin(ctx, PIC_T_PUSH);
  int rc = DISPATCH(push, ctx, species, 1, 0, 0);
  timing_end(ctx, PIC_T_PUSH);
  return rc;
}

extern "C" int pic_drift_deposit_n(pic_ctx *ctx, int species) {
  if (!ctx || species < 0 || species > 1) return PIC_E_ARG;
  timing_begin(ctx, PIC_T_PUSH);
  int rc = DISPATCH(push, ctx, species, 0, 1, 0);
  timing_end(ctx, PIC_T_PUSH);
  return rc;
}

extern "C" int pic_fold_current(pic_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  int rc;
  if (!is_f64(ctx) && (rc = f32::finish_deposit(ctx))) return rc;
  timing_begin(ctx, PIC_T_OTHER);
  rc = DISPATCH(fold_current, ctx);
  timing_end(ctx, PIC_T_OTHER);
  return rc;
}

extern "C" int pic_renew_e(pic_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  int rc;
  if (!is_f64(ctx) && (rc = f32::finish_deposit(ctx))) return rc;
  timing_begin(ctx, PIC_T_FIELD_E);
  rc = DISPATCH(renew_e, ctx);
  timing_end(ctx, PIC_T_FIELD_E);
  return rc;
}

extern "C" int pic_particle_boundary(pic_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  int rc = 0;
  timing_begin(ctx, PIC_T_OTHER);
  for (int sp = 0; sp < 2 && !rc; ++sp) rc = DISPATCH(particle_boundary, ctx, sp);
  timing_end(ctx, PIC_T_OTHER);
  return rc;
}

extern "C" int pic_inflow(pic_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  ctx->inflow_event++;
  int rc = 0;
  timing_begin(ctx, PIC_T_INFLOW);
  for (int sp = 0; sp < 2 && !rc; ++sp)
    rc = DISPATCH(inflow, ctx, sp, ctx->inflow_event);
  timing_end(ctx, PIC_T_INFLOW);
  return rc;
}

extern "C" int pic_sort(pic_ctx *ctx, int species) {
  if (!ctx || species < 0 || species > 1) return PIC_E_ARG;
  timing_begin(ctx, PIC_T_SORT);
  int rc = DISPATCH(sort, ctx, species);
  timing_end(ctx, PIC_T_SORT);
  return rc;
}

extern "C" int pic_diagnose(pic_ctx *ctx, pic_diag *out) {
  if (!ctx || !out) return PIC_E_ARG;
  int rc;
  if (!is_f64(ctx) && (rc = f32::finish_deposit(ctx))) return rc;
  return DISPATCH(diagnose, ctx, out);
}

extern "C" int pic_begin(pic_ctx *ctx) { return pic_inflow(ctx); }

extern "C" int pic_sort_stats(const pic_ctx *ctx, int64_t *tile, int64_t *radix) {
  if (!ctx) return PIC_E_ARG;
  if (tile) *tile = ctx->sort_tile_count;
  if (radix) *radix = ctx->sort_radix_count;
  return 0;
}

extern "C" int pic_get_counters(const pic_ctx *ctx, int64_t *step,
                                int32_t *c_add, int32_t *c_sort) {
  if (!ctx) return PIC_E_ARG;
  if (step) *step = ctx->step;
  if (c_add) *c_add = ctx->c_add;
  if (c_sort) *c_sort = ctx->c_sort;
  return 0;
}

extern "C" int pic_set_counters(pic_ctx *ctx, int64_t step, int32_t c_add,
                                int32_t c_sort) {
  if (!ctx || step < 0 || c_add < 0 || c_sort < 0) return PIC_E_ARG;
  ctx->step = step;
  ctx->c_add = c_add;
  ctx->c_sort = c_sort;
  return 0;
}

namespace pic {

// ---------------------------------------------------------------------------
// thrust_fave (:5124-5285) accumulated over pic_step's steps

// acc[q] += qn * (sum over the nb workgroups' partials of quantity q), the
// workgroups summed in a fixed order
constexpr int kFinish = 1024;
__global__ __launch_bounds__(kFinish) void k_thrust_finish(const double *__restrict__ part,
                                                           int nb, double qn, double *acc) {
  __shared__ double sh[6][kFinish];
  double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += kFinish)
    for (int q = 0; q < 6; ++q) v[q] += part[(int64_t)b * 6 + q];
  for (int q = 0; q < 6; ++q) sh[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int st = kFinish / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int q = 0; q < 6; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 6) acc[threadIdx.x] += qn * sh[threadIdx.x][0];
}

// the force parts of the diagnostics partials (particle quantities 8..13,
// already multiplied by q_s w dx^3) of both species added to acc
__global__ __launch_bounds__(64) void k_thrust_from_diag(const double *__restrict__ pf,
                                                         double *acc) {
  const int t = threadIdx.x;
  if (t >= 12) return;
  const int sp = t / 6, q = t % 6;
  const double *pp =
      pf + (int64_t)kDiagNBF * kDiagNQF + (int64_t)sp * kDiagNBP * kDiagNQP;
  double v = 0.0;
  for (int b = 0; b < kDiagNBP; ++b) v += pp[(int64_t)b * kDiagNQP + 8 + q];
  acc[t] += v;
}

__global__ void k_fill_zero(double *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

int thrust_add(pic_ctx *ctx, int sp, int nb, double qn) {
  k_thrust_finish<<<1, kFinish, 0, ctx->stream>>>(ctx->thrust_part, nb, qn,
                                                  ctx->thrust_acc + 6 * sp);
  PIC_LAUNCH_CHECK(ctx);
  return 0;
}

// the pending force from the current state, by the diagnostics' own particle
// gather (the state is still that of the end of the last step)
static int thrust_settle(pic_ctx *ctx) {
  if (!ctx->thrust_pending) return 0;
  int rc = is_f64(ctx) ? f64::diagnose_partials(ctx, ctx->diag_partial, nullptr)
                       : f32::diagnose_partials(ctx, ctx->diag_partial, nullptr);
  if (rc) return rc;
  k_thrust_from_diag<<<1, 64, 0, ctx->stream>>>(ctx->diag_partial, ctx->thrust_acc);
  PIC_LAUNCH_CHECK(ctx);
  ctx->thrust_pending = 0;
  ctx->thrust_steps++;
  return 0;
}

// keep the end-of-step total B for the fused force of the next pass: the
// FP64 B planes into thrust_bprev, the FP32 form's total-B planes into the
// upper half of their own 6-plane block
static int thrust_keep_b(pic_ctx *ctx) {
  const int64_t pl = ctx->geom.plane;
  if (is_f64(ctx)) {
    PIC_HIP(ctx, hipMemcpyAsync(ctx->thrust_bprev, fields<double>(ctx).f[PIC_BX],
                                3 * pl * sizeof(double), hipMemcpyDeviceToDevice,
                                ctx->stream));
  } else {
    float *bt = f32::total_b(ctx);
    PIC_HIP(ctx, hipMemcpyAsync(bt + 3 * pl, bt, 3 * pl * sizeof(float),
                                hipMemcpyDeviceToDevice, ctx->stream));
  }
  return 0;
}

}  // namespace pic

// No reference counterpart as an entry point: switches on thrust_fave's
// per-step accumulation (:926, :5124-5285) inside pic_step.
extern "C" int pic_thrust_ave_enable(pic_ctx *ctx, int on) {
  if (!ctx) return PIC_E_ARG;
  ctx->thrust_pending = 0;
  ctx->thrust_now = 0;
  ctx->thrust_steps = 0;
  ctx->thrust_on = 0;
  if (!on) return 0;
  const Geom &g = ctx->geom;
  const size_t esz = is_f64(ctx) ? sizeof(double) : sizeof(float);
  const TileGeom tg = tile_geom(ctx);
  // workgroups of one pass: the tiles and the tail chunks of up to the
  // capacity's slots
  const int64_t slots = (int64_t)tg.ntiles() + div_up(ctx->cap, kTailChunk) + 1;
  if (!ctx->thrust_acc) PIC_HIP(ctx, hipMalloc(&ctx->thrust_acc, 16 * sizeof(double)));
  if (ctx->thrust_part_slots < slots) {
    if (ctx->thrust_part) (void)hipFree(ctx->thrust_part);
    ctx->thrust_part = nullptr;
    ctx->thrust_part_slots = 0;
    PIC_HIP(ctx, hipMalloc(&ctx->thrust_part, (size_t)slots * 6 * sizeof(double)));
    ctx->thrust_part_slots = slots;
  }
  if (is_f64(ctx) && !ctx->thrust_bprev)
    PIC_HIP(ctx, hipMalloc(&ctx->thrust_bprev, 3 * (size_t)g.plane * esz));
  if (ctx->diag_partial_cap < kDiagPartials) {
    if (ctx->diag_partial) (void)hipFree(ctx->diag_partial);
    ctx->diag_partial = nullptr;
    ctx->diag_partial_cap = 0;
    PIC_HIP(ctx, hipMalloc(&ctx->diag_partial, kDiagPartials * sizeof(double)));
    ctx->diag_partial_cap = kDiagPartials;
  }
  k_fill_zero<<<1, 64, 0, ctx->stream>>>(ctx->thrust_acc, 16);
  PIC_LAUNCH_CHECK(ctx);
  ctx->thrust_on = 1;
  return 0;
}

extern "C" int pic_thrust_ave_get(pic_ctx *ctx, double F[27], int64_t *nsteps,
                                  int reset) {
  if (!ctx || !F) return PIC_E_ARG;
  if (!ctx->thrust_on) return set_error(ctx, PIC_E_STATE, "thrust_ave: not enabled");
  int rc;
  if (!is_f64(ctx) && (rc = f32::finish_deposit(ctx))) return rc;
  if ((rc = thrust_settle(ctx))) return rc;
  double a[12];
  PIC_HIP(ctx, hipMemcpyAsync(a, ctx->thrust_acc, sizeof a, hipMemcpyDeviceToHost,
                              ctx->stream));
  PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  // F[0..2] total Lorentz force, F[3..8] ion E and v x B parts, F[9..14]
  // electron parts; F[15..26] (particles with flag < 0) stay 0: flag is 1
  // for every live particle (:510, :6058)
  for (int i = 0; i < 27; ++i) F[i] = 0.0;
  for (int j = 0; j < 3; ++j) {
    F[3 + j] = a[j];
    F[6 + j] = a[3 + j];
    F[9 + j] = a[6 + j];
    F[12 + j] = a[9 + j];
    F[j] = ((a[j] + a[3 + j]) + a[6 + j]) + a[9 + j];
  }
  if (nsteps) *nsteps = ctx->thrust_steps;
  if (reset) {
    k_fill_zero<<<1, 64, 0, ctx->stream>>>(ctx->thrust_acc, 16);
    PIC_LAUNCH_CHECK(ctx);
    ctx->thrust_steps = 0;
  }
  return 0;
}

namespace pic {
int push_fused(pic_ctx *ctx, int sp) {
  timing_begin(ctx, PIC_T_PUSH);
  int rc = DISPATCH(push, ctx, sp, 1, 1, 1);
  timing_end(ctx, PIC_T_PUSH);
  return rc;
}
}  // namespace pic

// fdtd() step loop (:687-949) with loops A and B and the particle boundary
// fused into one pass per species.
extern "C" int pic_step(pic_ctx *ctx, int nsteps) {
  if (!ctx || nsteps < 0) return PIC_E_ARG;
  const pic_config &c = ctx->cfg;
  const bool f64 = is_f64(ctx);
  int rc;
  // an error leaves no sort permutation waiting for a pass
#define TRY(x)                                  \
  do {                                          \
    if ((rc = (x))) {                           \
      if (!f64) (void)f32::settle_sorts(ctx);   \
      return rc;                                \
    }                                           \
  } while (0)
  // The FP32 form's sort leaves its permutation to the step's fused pass,
  // which reads each particle through it (sort_for_pass): no copy of the
  // particles.  When the sort and the inflow fall on the same step the
  // inflow runs first, so that nothing but field work lies between the sort
  // and the pass.  The inflow draws each particle from the counter of its
  // reference-order index and the sort only permutes, so either order leaves
  // the same particles; only their slots differ.
  const auto sort_both = [&]() -> int {
    ctx->c_sort = 0;
    for (int sp = 0; sp < 2; ++sp) {
      timing_begin(ctx, PIC_T_SORT);
      const int r = f64 ? f64::sort(ctx, sp) : f32::sort_for_pass(ctx, sp);
      timing_end(ctx, PIC_T_SORT);
      if (r) return r;
    }
    return 0;
  };
  for (int s = 0; s < nsteps; ++s) {
    TRY(pic_renew_grid(ctx));
    const bool sort_due = ctx->c_sort > c.sort_every - 1;
    const bool add_due = ctx->c_add > c.inflow_every - 1;
    const bool inflow_first = !f64 && sort_due && add_due;
    if (sort_due && !inflow_first) TRY(sort_both());
    if (add_due) {
      // the inflow removes z < 4 dx (:6019-6021); where thrust_fave's box
      // reaches below that (nz < 20) the pending force is taken first
      if (ctx->thrust_pending && c.nz * 0.2 < 4.0) TRY(thrust_settle(ctx));
      TRY(pic_inflow(ctx));
      ctx->c_add = 0;
    }
    if (inflow_first) TRY(sort_both());
    if (ctx->thrust_pending) {
      TRY(thrust_keep_b(ctx));
      ctx->thrust_now = 1;
    }
    TRY(pic_renew_b_half(ctx));
    TRY(pic_halo_b(ctx));
    for (int sp = 0; sp < 2; ++sp) {
      const int tclass = !f64 && f32::pending(ctx, sp) ? PIC_T_PUSH_SORT : PIC_T_PUSH;
      timing_begin(ctx, tclass);
      rc = DISPATCH(push, ctx, sp, 1, 1, 1);
      timing_end(ctx, tclass);
      if (rc) {
        ctx->thrust_now = 0;
        if (!f64) (void)f32::settle_sorts(ctx);
        return rc;
      }
    }
    if (ctx->thrust_now) {
      ctx->thrust_now = 0;
      ctx->thrust_pending = 0;
      ctx->thrust_steps++;
    }
    TRY(pic_renew_b_half(ctx));
    TRY(pic_halo_b(ctx));
    TRY(pic_fold_current(ctx));
    TRY(pic_renew_e(ctx));
    TRY(pic_halo_e(ctx));
    if (c.poisson_every > 0 && ctx->step % c.poisson_every == 0)
      TRY(pic_poisson(ctx));  // PHI(renew_phi3; sr_phi; sr_e) :901-906
    ctx->c_sort++;
    ctx->c_add++;
    ctx->step++;
    if (ctx->thrust_on) ctx->thrust_pending = 1;
  }
#undef TRY
  return 0;
}

// ---------------------------------------------------------------------------
// timing

extern "C" int pic_timing_enable(pic_ctx *ctx, int on) {
  if (!ctx || on < 0 || on > 2) return PIC_E_ARG;
  // events recorded under the old mode are collected first
  int rc = timing_flush(ctx);
  if (rc) return rc;
  ctx->timing = on;
  return 0;
}

extern "C" int pic_timing_get(pic_ctx *ctx, int which, double *ms,
                              int64_t *launches) {
  if (!ctx || which < 0 || which >= PIC_T_COUNT) return PIC_E_ARG;
  int rc = timing_flush(ctx);
  if (rc) return rc;
  if (ms) *ms = ctx->t_ms[which];
  if (launches) *launches = ctx->t_n[which];
  return 0;
}

extern "C" int pic_timing_reset(pic_ctx *ctx) {
  if (!ctx) return PIC_E_ARG;
  int rc = timing_flush(ctx);
  if (rc) return rc;
  for (int i = 0; i < PIC_T_COUNT; ++i) {
    ctx->t_ms[i] = 0;
    ctx->t_n[i] = 0;
  }
  return 0;
}

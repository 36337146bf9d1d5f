/*
 * pic_oracle.c - CPU oracle (FP64) for the 3D EM-PIC step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed CPU
 * baseline.  It restates, on one rank, the routines of ashida-lab/PLASMAsim
 * PIC/3d_pic458_domain.c that the step calls, with the same expression
 * structure so that IEEE FP64 arithmetic matches operation for operation.
 * Build with -ffp-contract=off (no fused multiply-add contraction).
 *
 * Deliberate differences from the reference, shared with the library:
 *   - One particle list per species (the reference's THREAD_N lists are an
 *     OpenMP artefact, :99); the list is a contiguous array.
 *   - Particles that cross a periodic face are shifted in place instead of
 *     being unlinked and appended at the list end by the MPI transport
 *     (:2041-2531); the particle multiset is identical, only list order
 *     differs, which no routine except the stable sort observes.
 *   - Random numbers come from Philox4x32-10 keyed by (seed, species, event,
 *     particle index), so CPU and GPU draw identical streams; the reference's
 *     GSL generators (:436-446) are not available here.
 *   - Loading places exactly ppc particles per cell (the reference's list
 *     construction leaves one extra particle per thread list, :449-545).
 *
 * Pinning: every routine is checked on the CPU against known answers and
 * closed forms (tests/test_oracle_pinning.py), including the softened dipole
 * core (r < dipole_soft, a config-level extension: the point-dipole field at
 * radius dipole_soft on the same ray) and the diagnostics (exact divergence
 * of cubic fields, uniform-field energies and Lorentz forces).  Unpinned:
 * none beyond the random streams, which no reference test or fixture fixes.
 */
#include "pic_oracle_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;

int orc_threads(int n) {
  if (n > 0) g_threads = n;
  return g_threads;
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11), the published round function. */

static inline void mulhilo32(uint32_t a, uint32_t b, uint32_t *hi,
                             uint32_t *lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo32(0xD2511F53u, c0, &hi0, &lo0);
    mulhilo32(0xCD9E8D57u, c2, &hi1, &lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* 53-bit uniform in [0,1) from two words */
static inline double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) *
         (1.0 / 9007199254740992.0);
}

/* One macro-particle from the stream (seed, species, event, index):
 * position uniform in the cell whose lower corner is (cx,cy,cz) in length
 * units, velocity mu + sigma*N(0,1) per component (ran_multivar_gaussian
 * with a diagonal covariance, :7956-7985; Cholesky of diag(s^2) is diag(s)).
 * Normals by Box-Muller. */
void orc_draw_particle(uint64_t seed, int sp, uint32_t event, uint64_t idx,
                       double cx, double cy, double cz, double dx,
                       double vth, const double mu[3], Particle *p) {
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t w[4][4];
  for (int d = 0; d < 4; d++) {
    uint32_t ctr[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), event,
                       (uint32_t)(4 * sp + d)};
    orc_philox4x32_10(ctr, key, w[d]);
  }
  double ux = u53(w[0][0], w[0][1]);
  double uy = u53(w[0][2], w[0][3]);
  double uz = u53(w[1][0], w[1][1]);
  double g0 = u53(w[1][2], w[1][3]);
  double g1 = u53(w[2][0], w[2][1]);
  double g2 = u53(w[2][2], w[2][3]);
  double g3 = u53(w[3][0], w[3][1]);
  p->x = cx + ux * dx;
  p->y = cy + uy * dx;
  p->z = cz + uz * dx;
  const double two_pi = 6.283185307179586;
  double r0 = sqrt(-2.0 * log(1.0 - g0));
  double r1 = sqrt(-2.0 * log(1.0 - g2));
  double n0 = r0 * cos(two_pi * g1);
  double n1 = r0 * sin(two_pi * g1);
  double n2 = r1 * cos(two_pi * g3);
  p->vx = vth * n0 + mu[0];
  p->vy = vth * n1 + mu[1];
  p->vz = vth * n2 + mu[2];
}

/* ------------------------------------------------------------------ */
/* particle lists */

int orc_plist_reserve(PList *L, int64_t cap) {
  if (cap <= L->cap) return 0;
  Particle *np = (Particle *)realloc(L->p, (size_t)cap * sizeof(Particle));
  if (!np) return -1;
  L->p = np;
  L->cap = cap;
  return 0;
}

/* ------------------------------------------------------------------ */
/* set-up: init_grid (:291-410) with external_current7 (:4579-4634) */

static void dipole_b0(const pic_config *c, double x, double y, double z,
                      int comp, double *out) {
  /* point dipole mu0 M/(4 pi) (3 (m.r) r - r^2 m) / r^5, moment
   * M (0, sin a, cos a); the reference's expression per component. */
  double r = sqrt(x * x + y * y + z * z);
  const double Pi = 3.14159265358979323846;
  const double sa = sin(c->dipole_alpha), ca = cos(c->dipole_alpha);
  if (c->dipole_soft > 0.0 && r < c->dipole_soft * c->dx) {
    /* softened core: the field at radius dipole_soft along the same ray */
    double s = c->dipole_soft * c->dx / r;
    x *= s;
    y *= s;
    z *= s;
    r = sqrt(x * x + y * y + z * z);
  }
  double M = c->dipole_m;
  double v;
  if (comp == 0)
    v = c->mu0 * M * (3. * x * y * sa + 3. * x * z * ca) / (4. * Pi * pow(r, 5));
  else if (comp == 1)
    v = c->mu0 * M * (3. * y * y * sa + 3. * y * z * ca - r * r * sa) /
        (4. * Pi * pow(r, 5));
  else
    v = c->mu0 * M * (3. * y * z * sa + 3. * z * z * ca - r * r * ca) /
        (4. * Pi * pow(r, 5));
  *out = v + c->imf[comp];
}

/* sigma depth along one axis on the 2x lattice (init_grid :332-366),
 * with the layer width mapped onto the reference's 32-cell profile. */
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

orc_ctx *orc_create(const pic_config *cfg) {
  orc_ctx *o = (orc_ctx *)calloc(1, sizeof(orc_ctx));
  if (!o) return NULL;
  o->cfg = *cfg;
  o->NX = cfg->nx;
  o->NY = cfg->ny;
  o->NZ = cfg->nz;
  o->SX = cfg->nx + 4;
  o->SY = cfg->ny + 4;
  o->SZ = cfg->nz + 4;
  size_t ncell = (size_t)o->SX * o->SY * o->SZ;
  o->grid = (Grid *)calloc(ncell, sizeof(Grid));
  o->sigma = (double *)calloc(ncell * 8, sizeof(double));
  o->nthreads = g_threads;
  o->acc = (Acc *)calloc(ncell * (size_t)o->nthreads, sizeof(Acc));
  if (!o->grid || !o->sigma || !o->acc) {
    orc_destroy(o);
    return NULL;
  }
  const pic_config *c = cfg;
  const double dx = c->dx;
  /* external_current7: B0 at each component's staggered point, inside
   * r < rmax measured at the cell corner (:4587-4592). */
  const int cx = (o->NX + 4) / 2, cy = (o->NY + 4) / 2, cz = (o->NZ + 4) / 2;
  for (int k = 0; k < o->SX; k++)
    for (int l = 0; l < o->SY; l++)
      for (int m = 0; m < o->SZ; m++) {
        Grid *g = &G(o, k, l, m);
        double x = (k - cx) * dx, y = (l - cy) * dx, z = (m - cz) * dx;
        double r = sqrt(x * x + y * y + z * z);
        if (r < c->dipole_rmax * dx) {
          dipole_b0(c, (k - cx + 0.5) * dx, (l - cy) * dx, (m - cz) * dx, 0,
                    &g->b0x);
          dipole_b0(c, (k - cx) * dx, (l - cy + 0.5) * dx, (m - cz) * dx, 1,
                    &g->b0y);
          dipole_b0(c, (k - cx) * dx, (l - cy) * dx, (m - cz + 0.5) * dx, 2,
                    &g->b0z);
        } else {
          g->b0x = c->imf[0];
          g->b0y = c->imf[1];
          g->b0z = c->imf[2];
        }
        g->bx = g->b0x;
        g->by = g->b0y;
        g->bz = g->b0z;
      }
  /* sigma profile (:332-366) */
  for (int k = 0; k < 2 * o->SX; k++) {
    double xd = sigma_depth(k, o->NX, c->absorb[0]);
    for (int l = 0; l < 2 * o->SY; l++) {
      double yd = sigma_depth(l, o->NY, c->absorb[1]);
      for (int m = 0; m < 2 * o->SZ; m++) {
        double zd = sigma_depth(m, o->NZ, c->absorb[2]);
        SIG(o, k, l, m) =
            c->sigma_amp * (0.1 * sqrt(pow(xd, 4) + pow(yd, 4) + pow(zd, 4)) +
                            0.01 * sqrt(pow(xd, 12) + pow(yd, 12) + pow(zd, 12)));
      }
    }
  }
  return o;
}

void orc_destroy(orc_ctx *o) {
  if (!o) return;
  free(o->grid);
  free(o->sigma);
  free(o->acc);
  free(o->sp[0].p);
  free(o->sp[1].p);
  free(o->adjG);
  free(o);
}

double orc_sigma(orc_ctx *o, int k2, int l2, int m2) {
  return SIG(o, k2, l2, m2);
}

/* ------------------------------------------------------------------ */
/* data movement */

static double *field_ptr(Grid *g, int which) {
  switch (which) {
    case PIC_EX: return &g->ex;
    case PIC_EY: return &g->ey;
    case PIC_EZ: return &g->ez;
    case PIC_BX: return &g->bx;
    case PIC_BY: return &g->by;
    case PIC_BZ: return &g->bz;
    case PIC_B0X: return &g->b0x;
    case PIC_B0Y: return &g->b0y;
    case PIC_B0Z: return &g->b0z;
    case PIC_JIX: return &g->jix0;
    case PIC_JIY: return &g->jiy0;
    case PIC_JIZ: return &g->jiz0;
    case PIC_JEX: return &g->jex0;
    case PIC_JEY: return &g->jey0;
    case PIC_JEZ: return &g->jez0;
    case PIC_NI: return &g->ni;
    case PIC_NE: return &g->ne;
  }
  return NULL;
}

int orc_upload_field(orc_ctx *o, int which, const double *host) {
  if (which < 0 || which >= PIC_NFIELDS) return PIC_E_ARG;
  size_t n = (size_t)o->SX * o->SY * o->SZ;
  for (size_t i = 0; i < n; i++) *field_ptr(&o->grid[i], which) = host[i];
  return 0;
}

int orc_download_field(orc_ctx *o, int which, double *host) {
  if (which < 0 || which > PIC_PHI) return PIC_E_ARG;
  size_t n = (size_t)o->SX * o->SY * o->SZ;
  for (size_t i = 0; i < n; i++)
    host[i] = which == PIC_PHI ? o->grid[i].phi : *field_ptr(&o->grid[i], which);
  return 0;
}

int orc_upload_particles(orc_ctx *o, int sp, int64_t n, const double *x,
                         const double *y, const double *z, const double *vx,
                         const double *vy, const double *vz) {
  if (sp < 0 || sp > 1 || n < 0) return PIC_E_ARG;
  PList *L = &o->sp[sp];
  if (orc_plist_reserve(L, n > 0 ? n : 1)) return PIC_E_NOMEM;
  for (int64_t i = 0; i < n; i++) {
    Particle p = {x[i], y[i], z[i], vx[i], vy[i], vz[i]};
    L->p[i] = p;
  }
  L->n = n;
  return 0;
}

int orc_download_particles(orc_ctx *o, int sp, int64_t cap, int64_t *n,
                           double *x, double *y, double *z, double *vx,
                           double *vy, double *vz) {
  if (sp < 0 || sp > 1) return PIC_E_ARG;
  PList *L = &o->sp[sp];
  *n = L->n;
  if (cap < L->n) return PIC_E_CAPACITY;
  for (int64_t i = 0; i < L->n; i++) {
    x[i] = L->p[i].x;
    y[i] = L->p[i].y;
    z[i] = L->p[i].z;
    vx[i] = L->p[i].vx;
    vy[i] = L->p[i].vy;
    vz[i] = L->p[i].vz;
  }
  return 0;
}

int64_t orc_particle_count(orc_ctx *o, int sp) { return o->sp[sp].n; }

/* ------------------------------------------------------------------ */
/* init_particle_flat (:412-625): ppc per interior cell, cells in (k,l,m)
 * order with m fastest, position uniform in the cell, drifting Maxwellian.
 * Particle index p = cell*ppc + j keys the Philox stream (event 0). */
int orc_load_particles(orc_ctx *o) {
  const pic_config *c = &o->cfg;
  const int64_t ncell = (int64_t)o->NX * o->NY * o->NZ;
  const int64_t n = ncell * c->ppc;
  for (int sp = 0; sp < 2; sp++) {
    PList *L = &o->sp[sp];
    if (orc_plist_reserve(L, n > 0 ? n : 1)) return PIC_E_NOMEM;
    L->n = n;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t i = 0; i < n; i++) {
      int64_t cell = i / c->ppc;
      int m = (int)(cell % o->NZ);
      int l = (int)((cell / o->NZ) % o->NY);
      int k = (int)(cell / ((int64_t)o->NZ * o->NY));
      orc_draw_particle(c->seed, sp, 0u, (uint64_t)i, (k + 2) * c->dx,
                    (l + 2) * c->dx, (m + 2) * c->dx, c->dx, c->vth[sp],
                    c->vdrift, &L->p[i]);
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* renew_grid (:2533-2578) */
int orc_renew_grid(orc_ctx *o) {
  size_t n = (size_t)o->SX * o->SY * o->SZ;
  for (size_t i = 0; i < n; i++) {
    Grid *g = &o->grid[i];
    g->jix0 = g->jiy0 = g->jiz0 = 0.;
    g->jex0 = g->jey0 = g->jez0 = 0.;
    g->ni = g->ne = 0.;
    g->phi = 0.;
  }
  return 0;
}

/* renew_b4_2 (:2580-2611): B <- B0 + f (B - B0) - g (27 D1 - D3)/24/dx */
int orc_renew_b_half(orc_ctx *o) {
  const pic_config *c = &o->cfg;
  const double dt = c->dt, dx = c->dx;
  const double temp = 1. / c->eps0 * dt * 0.5 * 0.5;
  const double dth = 0.5 * dt;
  const double temp_24 = 1. / 24.;
  const double inv_dx = 1. / dx;
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int k = 2; k < o->NX + 2; k++) {
    for (int l = 2; l < o->NY + 2; l++) {
      for (int m = 2; m < o->NZ + 2; m++) {
        Grid *g = &G(o, k, l, m);
        double s;
        s = SIG(o, 2 * k + 1, 2 * l, 2 * m);
        g->bx = g->b0x + (1. - s * temp) / (1. + s * temp) * (g->bx - g->b0x) -
                dth / (1. + s * temp) * temp_24 *
                    (27. * (G(o, k, l, m).ez - G(o, k, l - 1, m).ez -
                            G(o, k, l, m).ey + G(o, k, l, m - 1).ey) -
                     (G(o, k, l + 1, m).ez - G(o, k, l - 2, m).ez -
                      G(o, k, l, m + 1).ey + G(o, k, l, m - 2).ey)) *
                    inv_dx;
        s = SIG(o, 2 * k, 2 * l + 1, 2 * m);
        g->by = g->b0y + (1. - s * temp) / (1. + s * temp) * (g->by - g->b0y) -
                dth / (1. + s * temp) * temp_24 *
                    (27. * (G(o, k, l, m).ex - G(o, k, l, m - 1).ex -
                            G(o, k, l, m).ez + G(o, k - 1, l, m).ez) -
                     (G(o, k, l, m + 1).ex - G(o, k, l, m - 2).ex -
                      G(o, k + 1, l, m).ez + G(o, k - 2, l, m).ez)) *
                    inv_dx;
        s = SIG(o, 2 * k, 2 * l, 2 * m + 1);
        g->bz = g->b0z + (1. - s * temp) / (1. + s * temp) * (g->bz - g->b0z) -
                dth / (1. + s * temp) * temp_24 *
                    (27. * (G(o, k, l, m).ey - G(o, k - 1, l, m).ey -
                            G(o, k, l, m).ex + G(o, k, l - 1, m).ex) -
                     (G(o, k + 1, l, m).ey - G(o, k - 2, l, m).ey -
                      G(o, k, l + 1, m).ex + G(o, k, l - 2, m).ex)) *
                    inv_dx;
      }
    }
  }
  return 0;
}

/* renew_e4_2 (:2613-2652) */
int orc_renew_e(orc_ctx *o) {
  const pic_config *c = &o->cfg;
  const double dt = c->dt, dx = c->dx, mu0 = c->mu0;
  const double temp = c->c * c->c * dt;
  const double temp_24 = 1. / 24.;
  const double inv_dx = 1. / dx;
#define B1X(k, l, m) (G(o, k, l, m).bx - G(o, k, l, m).b0x)
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int k = 2; k < o->NX + 2; k++) {
    for (int l = 2; l < o->NY + 2; l++) {
      for (int m = 2; m < o->NZ + 2; m++) {
        Grid *g = &G(o, k, l, m);
        double s;
        s = SIG(o, 2 * k, 2 * l + 1, 2 * m + 1);
        g->ex = (1. - 0.5 * mu0 * s * temp) / (1. + 0.5 * mu0 * s * temp) * g->ex +
                temp / (1. + 0.5 * mu0 * s * temp) *
                    (temp_24 *
                         (27. * ((G(o, k, l + 1, m).bz - G(o, k, l + 1, m).b0z -
                                  G(o, k, l, m).bz + G(o, k, l, m).b0z) -
                                 (G(o, k, l, m + 1).by - G(o, k, l, m + 1).b0y -
                                  G(o, k, l, m).by + G(o, k, l, m).b0y)) -
                          ((G(o, k, l + 2, m).bz - G(o, k, l + 2, m).b0z -
                            G(o, k, l - 1, m).bz + G(o, k, l - 1, m).b0z) -
                           (G(o, k, l, m + 2).by - G(o, k, l, m + 2).b0y -
                            G(o, k, l, m - 1).by + G(o, k, l, m - 1).b0y))) *
                         inv_dx -
                     mu0 * (g->jix0 - g->jex0));
        s = SIG(o, 2 * k + 1, 2 * l, 2 * m + 1);
        g->ey = (1. - 0.5 * mu0 * s * temp) / (1. + 0.5 * mu0 * s * temp) * g->ey +
                temp / (1. + 0.5 * mu0 * s * temp) *
                    (temp_24 *
                         (27. * ((G(o, k, l, m + 1).bx - G(o, k, l, m + 1).b0x -
                                  G(o, k, l, m).bx + G(o, k, l, m).b0x) -
                                 (G(o, k + 1, l, m).bz - G(o, k + 1, l, m).b0z -
                                  G(o, k, l, m).bz + G(o, k, l, m).b0z)) -
                          ((G(o, k, l, m + 2).bx - G(o, k, l, m + 2).b0x -
                            G(o, k, l, m - 1).bx + G(o, k, l, m - 1).b0x) -
                           (G(o, k + 2, l, m).bz - G(o, k + 2, l, m).b0z -
                            G(o, k - 1, l, m).bz + G(o, k - 1, l, m).b0z))) *
                         inv_dx -
                     mu0 * (g->jiy0 - g->jey0));
        s = SIG(o, 2 * k + 1, 2 * l + 1, 2 * m);
        g->ez = (1. - 0.5 * mu0 * s * temp) / (1. + 0.5 * mu0 * s * temp) * g->ez +
                temp / (1. + 0.5 * mu0 * s * temp) *
                    (temp_24 *
                         (27. * ((G(o, k + 1, l, m).by - G(o, k + 1, l, m).b0y -
                                  G(o, k, l, m).by + G(o, k, l, m).b0y) -
                                 (G(o, k, l + 1, m).bx - G(o, k, l + 1, m).b0x -
                                  G(o, k, l, m).bx + G(o, k, l, m).b0x)) -
                          ((G(o, k + 2, l, m).by - G(o, k + 2, l, m).b0y -
                            G(o, k - 1, l, m).by + G(o, k - 1, l, m).b0y) -
                           (G(o, k, l + 2, m).bx - G(o, k, l + 2, m).b0x -
                            G(o, k, l - 1, m).bx + G(o, k, l - 1, m).b0x))) *
                         inv_dx -
                     mu0 * (g->jiz0 - g->jez0));
      }
    }
  }
#undef B1X
  return 0;
}

/* sr_b / sr_e on one rank (:1001-1387): x faces over all (l,m), then
 * y faces over all (k,m); guards 0,1 <- N,N+1 and N+2,N+3 <- 2,3. */
static void halo3(orc_ctx *o, size_t off0, size_t off1, size_t off2) {
  const int NX = o->NX, NY = o->NY;
#define F(k, l, m, off) (*(double *)((char *)&G(o, k, l, m) + (off)))
  size_t offs[3] = {off0, off1, off2};
  for (int c = 0; c < 3; c++) {
    size_t off = offs[c];
    for (int l = 0; l < o->SY; l++)
      for (int m = 0; m < o->SZ; m++) {
        F(0, l, m, off) = F(NX, l, m, off);
        F(1, l, m, off) = F(NX + 1, l, m, off);
        F(NX + 2, l, m, off) = F(2, l, m, off);
        F(NX + 3, l, m, off) = F(3, l, m, off);
      }
  }
  for (int c = 0; c < 3; c++) {
    size_t off = offs[c];
    for (int k = 0; k < o->SX; k++)
      for (int m = 0; m < o->SZ; m++) {
        F(k, 0, m, off) = F(k, NY, m, off);
        F(k, 1, m, off) = F(k, NY + 1, m, off);
        F(k, NY + 2, m, off) = F(k, 2, m, off);
        F(k, NY + 3, m, off) = F(k, 3, m, off);
      }
  }
#undef F
}

int orc_halo_b(orc_ctx *o) {
  halo3(o, offsetof(Grid, bx), offsetof(Grid, by), offsetof(Grid, bz));
  return 0;
}

int orc_halo_e(orc_ctx *o) {
  halo3(o, offsetof(Grid, ex), offsetof(Grid, ey), offsetof(Grid, ez));
  return 0;
}

/* sr_current2 (:1805-2039) on one rank: 2,3 += N+2,N+3 then N,N+1 += 0,1
 * along x over all (l,m); the same along y over all (k,m).  z guards are not
 * folded (the reference never exchanges along z). */
int orc_fold_current(orc_ctx *o) {
  const int NX = o->NX, NY = o->NY;
  static const size_t offs[8] = {
      offsetof(Grid, jix0), offsetof(Grid, jiy0), offsetof(Grid, jiz0),
      offsetof(Grid, jex0), offsetof(Grid, jey0), offsetof(Grid, jez0),
      offsetof(Grid, ni),   offsetof(Grid, ne)};
#define F(k, l, m, off) (*(double *)((char *)&G(o, k, l, m) + (off)))
  for (int l = 0; l < o->SY; l++)
    for (int m = 0; m < o->SZ; m++)
      for (int c = 0; c < 8; c++) {
        F(2, l, m, offs[c]) += F(NX + 2, l, m, offs[c]);
        F(3, l, m, offs[c]) += F(NX + 3, l, m, offs[c]);
      }
  for (int l = 0; l < o->SY; l++)
    for (int m = 0; m < o->SZ; m++)
      for (int c = 0; c < 8; c++) {
        F(NX, l, m, offs[c]) += F(0, l, m, offs[c]);
        F(NX + 1, l, m, offs[c]) += F(1, l, m, offs[c]);
      }
  for (int k = 0; k < o->SX; k++)
    for (int m = 0; m < o->SZ; m++)
      for (int c = 0; c < 8; c++) {
        F(k, 2, m, offs[c]) += F(k, NY + 2, m, offs[c]);
        F(k, 3, m, offs[c]) += F(k, NY + 3, m, offs[c]);
      }
  for (int k = 0; k < o->SX; k++)
    for (int m = 0; m < o->SZ; m++)
      for (int c = 0; c < 8; c++) {
        F(k, NY, m, offs[c]) += F(k, 0, m, offs[c]);
        F(k, NY + 1, m, offs[c]) += F(k, 1, m, offs[c]);
      }
#undef F
  return 0;
}

/* ------------------------------------------------------------------ */
/* shape_func_v2 (:3017-3326): trilinear gather on six staggered lattices */

/* weights along one axis: integer-located component (base (int)(x/dx)) */
#define WINT(xv, k, up, lo)          \
  k = (int)((xv) / dx);              \
  up = ((k + 1) * dx - (xv));        \
  lo = ((xv) - k * dx)
/* half-located component (base (int)(x/dx-0.5)) */
#define WHALF(xv, k, up, lo)         \
  k = (int)((xv) / dx - 0.5);        \
  up = (((k + 1) + 0.5) * dx - (xv)); \
  lo = ((xv) - (k + 0.5) * dx)

static inline double gather1(orc_ctx *o, size_t off, int a, int b, int c,
                             double wa0, double wa1, double wb0, double wb1,
                             double wc0, double wc1) {
#define F(k, l, m) (*(double *)((char *)&G(o, k, l, m) + off))
  double p1 = F(a, b, c) * (wa0 * wb0 * wc0) + F(a, b, c + 1) * (wa0 * wb0 * wc1);
  double p2 = F(a, b + 1, c) * (wa0 * wb1 * wc0) +
              F(a, b + 1, c + 1) * (wa0 * wb1 * wc1);
  double p3 = F(a + 1, b, c) * (wa1 * wb0 * wc0) +
              F(a + 1, b, c + 1) * (wa1 * wb0 * wc1);
  double p4 = F(a + 1, b + 1, c) * (wa1 * wb1 * wc0) +
              F(a + 1, b + 1, c + 1) * (wa1 * wb1 * wc1);
#undef F
  return p1 + p2 + p3 + p4;
}

void orc_gather(orc_ctx *o, double x, double y, double z, double out[6]) {
  const double dx = o->cfg.dx;
  const double inv_dx3 = 1. / (dx * dx * dx);
  int ki, kh, li, lh, mi, mh;
  double kiu, kil, khu, khl, liu, lil, lhu, lhl, miu, mil, mhu, mhl;
  WINT(x, ki, kiu, kil);
  WHALF(x, kh, khu, khl);
  WINT(y, li, liu, lil);
  WHALF(y, lh, lhu, lhl);
  WINT(z, mi, miu, mil);
  WHALF(z, mh, mhu, mhl);
  /* Ex (k, l+1/2, m+1/2) */
  out[0] = gather1(o, offsetof(Grid, ex), ki, lh, mh, kiu, kil, lhu, lhl, mhu,
                   mhl) * inv_dx3;
  /* Ey (k+1/2, l, m+1/2) */
  out[1] = gather1(o, offsetof(Grid, ey), kh, li, mh, khu, khl, liu, lil, mhu,
                   mhl) * inv_dx3;
  /* Ez (k+1/2, l+1/2, m) */
  out[2] = gather1(o, offsetof(Grid, ez), kh, lh, mi, khu, khl, lhu, lhl, miu,
                   mil) * inv_dx3;
  /* Bx (k+1/2, l, m) */
  out[3] = gather1(o, offsetof(Grid, bx), kh, li, mi, khu, khl, liu, lil, miu,
                   mil) * inv_dx3;
  /* By (k, l+1/2, m) */
  out[4] = gather1(o, offsetof(Grid, by), ki, lh, mi, kiu, kil, lhu, lhl, miu,
                   mil) * inv_dx3;
  /* Bz (k, l, m+1/2) */
  out[5] = gather1(o, offsetof(Grid, bz), ki, li, mh, kiu, kil, liu, lil, mhu,
                   mhl) * inv_dx3;
}

/* cal_track4 (:2804-2870) given the gathered fields f = (ex..bz):
 * relativistic Boris with B scaled by 1/gamma(u^{n-1/2}), gamma1 -> 0 when
 * |v| >= c (isnan), then x += v dt/2.  m < 0 for electrons (:784). */
void orc_boris(const pic_config *cfg, double m, const double f[6],
               double p[6]) {
  const double C = cfg->c, dt = cfg->dt;
  const double T = 0.5 * cfg->q * dt / m;
  double ex = f[0], ey = f[1], ez = f[2], bx = f[3], by = f[4], bz = f[5];
  double vx = p[3], vy = p[4], vz = p[5];
  ex += cfg->vdrift[2] * cfg->imf[1];
  ey += -cfg->vdrift[2] * cfg->imf[0];
  double gamma1 = C / sqrt(C * C - vx * vx - vy * vy - vz * vz);
  if (isnan(gamma1)) gamma1 = 0.;
  double ux = gamma1 * vx, uy = gamma1 * vy, uz = gamma1 * vz;
  double gamma2 = C / sqrt(C * C + ux * ux + uy * uy + uz * uz);
  double bux = gamma2 * (bx), buy = gamma2 * (by), buz = gamma2 * (bz);
  double S = 2. * T / (1. + T * T * (bux * bux + buy * buy + buz * buz));
  ux += (ex) * T;
  uy += (ey) * T;
  uz += (ez) * T;
  double uox = ux + T * (uy * buz - uz * buy);
  double uoy = uy + T * (uz * bux - ux * buz);
  double uoz = uz + T * (ux * buy - uy * bux);
  double upx = ux + S * (uoy * buz - uoz * buy);
  double upy = uy + S * (uoz * bux - uox * buz);
  double upz = uz + S * (uox * buy - uoy * bux);
  upx += (ex) * T;
  upy += (ey) * T;
  upz += (ez) * T;
  gamma2 = C / sqrt(C * C + upx * upx + upy * upy + upz * upz);
  p[3] = gamma2 * upx;
  p[4] = gamma2 * upy;
  p[5] = gamma2 * upz;
  p[0] += p[3] * dt * 0.5;
  p[1] += p[4] * dt * 0.5;
  p[2] += p[5] * dt * 0.5;
}

/* ------------------------------------------------------------------ */
/* deposits into thread-private accumulators, then a fixed-order reduction
 * into the grid, as grid_thread and the reduction at :869-886 do. */

static void acc_zero(orc_ctx *o) {
  size_t n = (size_t)o->SX * o->SY * o->SZ * o->nthreads;
  memset(o->acc, 0, n * sizeof(Acc));
}

static void acc_reduce(orc_ctx *o, const size_t offs[4], int ncomp) {
  size_t n = (size_t)o->SX * o->SY * o->SZ;
  for (int t = 0; t < o->nthreads; t++) {
    Acc *a = o->acc + (size_t)t * n;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (size_t i = 0; i < n; i++)
      for (int c = 0; c < ncomp; c++)
        *(double *)((char *)&o->grid[i] + offs[c]) += a[i].a[c];
  }
}

#define A(k, l, m) (acc[(((size_t)(k) * o->SY) + (l)) * o->SZ + (m)])

/* shape_func_{ion,electron}0_j_2 (:3397-3553) */
static inline void deposit_j(orc_ctx *o, Acc *acc, const Particle *p,
                             double wgt) {
  const double dx = o->cfg.dx;
  const double dx3 = dx * dx * dx;
  int k = (int)(p->x / dx);
  int l = (int)(p->y / dx - 0.5);
  int m = (int)(p->z / dx - 0.5);
  int kkk = k + 1, lll = l + 1, mmm = m + 1;
  const double qn = o->cfg.q * wgt / dx3;
  double temp = qn * p->vx;
  double kkkp = (kkk * dx - p->x);
  double lllp = ((lll + 0.5) * dx - p->y);
  double mmmp = ((mmm + 0.5) * dx - p->z);
  double pk = (p->x - k * dx);
  double pl = (p->y - (l + 0.5) * dx);
  double pm = (p->z - (m + 0.5) * dx);
  double tempkkkp = temp * kkkp;
  double temppk = temp * pk;
  A(k, l, m).a[0] += (tempkkkp) * (lllp * mmmp);
  A(k, l, mmm).a[0] += (tempkkkp) * (lllp * pm);
  A(k, lll, m).a[0] += (tempkkkp) * (pl * mmmp);
  A(k, lll, mmm).a[0] += (tempkkkp) * (pl * pm);
  A(kkk, l, m).a[0] += (temppk) * (lllp * mmmp);
  A(kkk, l, mmm).a[0] += (temppk) * (lllp * pm);
  A(kkk, lll, m).a[0] += (temppk) * (pl * mmmp);
  A(kkk, lll, mmm).a[0] += (temppk) * (pl * pm);

  k = (int)(p->x / dx - 0.5);
  l = (int)(p->y / dx);
  kkk = k + 1;
  lll = l + 1;
  temp = qn * p->vy;
  kkkp = ((kkk + 0.5) * dx - p->x);
  lllp = (lll * dx - p->y);
  pk = (p->x - (k + 0.5) * dx);
  pl = (p->y - l * dx);
  tempkkkp = temp * kkkp;
  temppk = temp * pk;
  A(k, l, m).a[1] += (tempkkkp) * (lllp * mmmp);
  A(k, l, mmm).a[1] += (tempkkkp) * (lllp * pm);
  A(k, lll, m).a[1] += (tempkkkp) * (pl * mmmp);
  A(k, lll, mmm).a[1] += (tempkkkp) * (pl * pm);
  A(kkk, l, m).a[1] += (temppk) * (lllp * mmmp);
  A(kkk, l, mmm).a[1] += (temppk) * (lllp * pm);
  A(kkk, lll, m).a[1] += (temppk) * (pl * mmmp);
  A(kkk, lll, mmm).a[1] += (temppk) * (pl * pm);

  l = (int)(p->y / dx - 0.5);
  m = (int)(p->z / dx);
  lll = l + 1;
  mmm = m + 1;
  temp = qn * p->vz;
  lllp = ((lll + 0.5) * dx - p->y);
  mmmp = (mmm * dx - p->z);
  pl = (p->y - (l + 0.5) * dx);
  pm = (p->z - m * dx);
  tempkkkp = temp * kkkp;
  temppk = temp * pk;
  A(k, l, m).a[2] += (tempkkkp) * (lllp * mmmp);
  A(k, l, mmm).a[2] += (tempkkkp) * (lllp * pm);
  A(k, lll, m).a[2] += (tempkkkp) * (pl * mmmp);
  A(k, lll, mmm).a[2] += (tempkkkp) * (pl * pm);
  A(kkk, l, m).a[2] += (temppk) * (lllp * mmmp);
  A(kkk, l, mmm).a[2] += (temppk) * (lllp * pm);
  A(kkk, lll, m).a[2] += (temppk) * (pl * mmmp);
  A(kkk, lll, mmm).a[2] += (temppk) * (pl * pm);
}

/* shape_func_{ion,electron}0_n_2 (:3333-3395) */
static inline void deposit_n(orc_ctx *o, Acc *acc, const Particle *p,
                             double wgt) {
  const double dx = o->cfg.dx;
  const double dx3 = dx * dx * dx;
  const int k = (int)(p->x / dx - 0.5);
  const int l = (int)(p->y / dx - 0.5);
  const int m = (int)(p->z / dx - 0.5);
  const int kkk = k + 1, lll = l + 1, mmm = m + 1;
  const double temp = wgt / dx3;
  const double kkkp = ((kkk + 0.5) * dx - p->x);
  const double lllp = ((lll + 0.5) * dx - p->y);
  const double mmmp = ((mmm + 0.5) * dx - p->z);
  const double pk = (p->x - (k + 0.5) * dx);
  const double pl = (p->y - (l + 0.5) * dx);
  const double pm = (p->z - (m + 0.5) * dx);
  const double tempkkkp = temp * kkkp;
  const double temppk = temp * pk;
  A(k, l, m).a[0] += (tempkkkp) * (lllp * mmmp);
  A(k, l, mmm).a[0] += (tempkkkp) * (lllp * pm);
  A(k, lll, m).a[0] += (tempkkkp) * (pl * mmmp);
  A(k, lll, mmm).a[0] += (tempkkkp) * (pl * pm);
  A(kkk, l, m).a[0] += (temppk) * (lllp * mmmp);
  A(kkk, l, mmm).a[0] += (temppk) * (lllp * pm);
  A(kkk, lll, m).a[0] += (temppk) * (pl * mmmp);
  A(kkk, lll, mmm).a[0] += (temppk) * (pl * pm);
}
#undef A

/* deposit gate (:747-749): 0 <= x < (N+4)dx on every axis */
static inline int in_gate(const orc_ctx *o, const Particle *p) {
  const double dx = o->cfg.dx;
  const double d3x = (o->NX + 4) * dx, d3y = (o->NY + 4) * dx,
               d3z = (o->NZ + 4) * dx;
  const double dx0 = 0.;
  return p->x < d3x && p->x >= dx0 && p->y < d3y && p->y >= dx0 &&
         p->z < d3z && p->z >= dx0;
}

/* Array-safety rule shared with the library.  The gather and deposit
 * touch nodes base..base+1 with base (int)(x/dx) or (int)(x/dx-0.5); the
 * reference has no bounds check (:747-749).  Those nodes stay inside the
 * array when -0.5 < x/dx < S-1 on every axis; live particles of the step
 * always satisfy it (|v| < c, dt < dx).  A particle outside it is drifted
 * without field push and deposits nothing, instead of touching memory
 * outside the arrays. */
static inline int in_array(const orc_ctx *o, const Particle *p) {
  const double dx = o->cfg.dx;
  const double xd = p->x / dx, yd = p->y / dx, zd = p->z / dx;
  return xd > -0.5 && xd < o->SX - 1 && yd > -0.5 && yd < o->SY - 1 &&
         zd > -0.5 && zd < o->SZ - 1;
}

/* hot loop A (:731-812) for one species */
int orc_push_deposit_j(orc_ctx *o, int sp) {
  if (sp < 0 || sp > 1) return PIC_E_ARG;
  const pic_config *c = &o->cfg;
  const double dth = c->dt * 0.5;
  const double m = sp == PIC_ION ? c->mass[0] : -c->mass[1];
  const double ax0 = c->ballistic_lo[0], ax1 = c->ballistic_hi[0];
  const double ay0 = c->ballistic_lo[1], ay1 = c->ballistic_hi[1];
  const double az0 = c->ballistic_lo[2], az1 = c->ballistic_hi[2];
  PList *L = &o->sp[sp];
  const size_t ncell = (size_t)o->SX * o->SY * o->SZ;
  acc_zero(o);
  const int T = o->nthreads;
#pragma omp parallel num_threads(T)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    int64_t lo = L->n * t / T, hi = L->n * (t + 1) / T;
    Acc *acc = o->acc + (size_t)t * ncell;
    for (int64_t i = lo; i < hi; i++) {
      Particle *p = &L->p[i];
      if (p->x < ax0 || p->x > ax1 || p->y < ay0 || p->y > ay1 ||
          p->z < az0 || p->z > az1 || !in_array(o, p)) {
        p->x += p->vx * dth;
        p->y += p->vy * dth;
        p->z += p->vz * dth;
      } else {
        double f[6], s[6] = {p->x, p->y, p->z, p->vx, p->vy, p->vz};
        orc_gather(o, p->x, p->y, p->z, f);
        orc_boris(c, m, f, s);
        p->x = s[0];
        p->y = s[1];
        p->z = s[2];
        p->vx = s[3];
        p->vy = s[4];
        p->vz = s[5];
      }
      if (in_gate(o, p) && in_array(o, p)) deposit_j(o, acc, p, c->weight[sp]);
    }
  }
  size_t offs[4];
  if (sp == PIC_ION) {
    offs[0] = offsetof(Grid, jix0);
    offs[1] = offsetof(Grid, jiy0);
    offs[2] = offsetof(Grid, jiz0);
  } else {
    offs[0] = offsetof(Grid, jex0);
    offs[1] = offsetof(Grid, jey0);
    offs[2] = offsetof(Grid, jez0);
  }
  acc_reduce(o, offs, 3);
  return 0;
}

/* hot loop B (:818-867) for one species */
int orc_drift_deposit_n(orc_ctx *o, int sp) {
  if (sp < 0 || sp > 1) return PIC_E_ARG;
  const pic_config *c = &o->cfg;
  const double dth = c->dt * 0.5;
  PList *L = &o->sp[sp];
  const size_t ncell = (size_t)o->SX * o->SY * o->SZ;
  acc_zero(o);
  const int T = o->nthreads;
#pragma omp parallel num_threads(T)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    int64_t lo = L->n * t / T, hi = L->n * (t + 1) / T;
    Acc *acc = o->acc + (size_t)t * ncell;
    for (int64_t i = lo; i < hi; i++) {
      Particle *p = &L->p[i];
      p->x += p->vx * dth;
      p->y += p->vy * dth;
      p->z += p->vz * dth;
      if (in_gate(o, p) && in_array(o, p)) deposit_n(o, acc, p, c->weight[sp]);
    }
  }
  size_t offs[4] = {sp == PIC_ION ? offsetof(Grid, ni) : offsetof(Grid, ne)};
  acc_reduce(o, offs, 1);
  return 0;
}

/* ------------------------------------------------------------------ */
/* sr_particle (:2041-2253) with sr_particle_x/_y (:2255-2531) on one rank:
 * particles with z outside [2dx,(Nz+2)dx) are deleted; the others crossing
 * the x faces are shifted by -/+ Nx dx (x pass), then y likewise.  On a
 * non-periodic axis crossers are dropped, as at the reference's global
 * edges (:2138-2148). */
int orc_particle_boundary(orc_ctx *o) {
  const pic_config *c = &o->cfg;
  const double dx = c->dx;
  const double dx2 = 2. * dx;
  const double gx2 = (o->NX + 2) * dx, gy2 = (o->NY + 2) * dx,
               gz2 = (o->NZ + 2) * dx;
  for (int sp = 0; sp < 2; sp++) {
    PList *L = &o->sp[sp];
    int64_t w = 0;
    for (int64_t i = 0; i < L->n; i++) {
      Particle p = L->p[i];
      if (!(p.z < gz2 && p.z >= dx2)) continue;
      if (p.x >= gx2) {
        if (!c->periodic[0]) continue;
        p.x = p.x - o->NX * dx;
      } else if (p.x < dx2) {
        if (!c->periodic[0]) continue;
        p.x = p.x + o->NX * dx;
      }
      if (p.y >= gy2) {
        if (!c->periodic[1]) continue;
        p.y = p.y - o->NY * dx;
      } else if (p.y < dx2) {
        if (!c->periodic[1]) continue;
        p.y = p.y + o->NY * dx;
      }
      L->p[w++] = p;
    }
    L->n = w;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* add_ion_flat / add_electron_flat (:5993-6817) on one rank, for the z
 * faces: delete z < 4dx or z > Nz dx (:6019-6021), then append ppc
 * particles per cell for the 2-cell slab at the low face (z in
 * [(m+2)dx,(m+3)dx), m = 0,1) and, when refill_far is set, the high face
 * (z in [(m+Nz)dx,(m+Nz+1)dx)) (:6042-6104).  Loop order k, l, m, j as the
 * reference's; the batch index keys the Philox stream of this event. */
int orc_inflow(orc_ctx *o) {
  const pic_config *c = &o->cfg;
  const double dx = c->dx;
  o->inflow_event++;
  const uint32_t ev = o->inflow_event;
  for (int sp = 0; sp < 2; sp++) {
    PList *L = &o->sp[sp];
    int64_t w = 0;
    for (int64_t i = 0; i < L->n; i++) {
      Particle p = L->p[i];
      if (p.z < 4 * dx || p.z > o->NZ * dx) continue;
      L->p[w++] = p;
    }
    L->n = w;
    const int nface = c->refill_far ? 2 : 1;
    const int64_t per_face = (int64_t)o->NX * o->NY * 2 * c->ppc;
    if (orc_plist_reserve(L, L->n + per_face * nface)) return PIC_E_NOMEM;
    int64_t base = L->n;
    for (int face = 0; face < nface; face++) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
      for (int64_t b = 0; b < per_face; b++) {
        int j = (int)(b % c->ppc);
        int m = (int)((b / c->ppc) % 2);
        int l = (int)((b / (2 * c->ppc)) % o->NY);
        int k = (int)(b / ((int64_t)2 * c->ppc * o->NY));
        (void)j;
        int64_t idx = face * per_face + b;
        double zlo = face == 0 ? (m + 2) * dx : (m + o->NZ) * dx;
        orc_draw_particle(c->seed, sp, ev, (uint64_t)idx, (k + 2) * dx,
                      (l + 2) * dx, zlo, dx, c->vth[sp], c->vdrift,
                      &L->p[base + idx]);
      }
    }
    L->n = base + per_face * nface;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* sort key of an interior cell: tile-major, lexicographic (k,l,m) with m
 * fastest inside a tile; with one tile covering the interior this is the
 * reference's ((k*Ny)+l)*Nz+m. */
static int64_t cell_key(const orc_ctx *o, int k, int l, int m) {
  const int TX = o->cfg.tile[0] > 0 ? o->cfg.tile[0] : o->NX;
  const int TY = o->cfg.tile[1] > 0 ? o->cfg.tile[1] : o->NY;
  const int TZ = o->cfg.tile[2] > 0 ? o->cfg.tile[2] : o->NZ;
  const int NTY = o->NY / TY, NTZ = o->NZ / TZ;
  const int64_t tile = ((int64_t)(k / TX) * NTY + l / TY) * NTZ + m / TZ;
  const int64_t in = ((int64_t)(k % TX) * TY + l % TY) * TZ + m % TZ;
  return tile * TX * TY * TZ + in;
}

/* sorting (:7475-7590): key ((int)(x/dx)-2, (int)(y/dx)-2, (int)(z/dx)-2),
 * particles outside [0,N) on any axis are dropped, the rest rebuilt in key
 * order (the reference's lexicographic (k,l,m) order, m fastest, when one
 * tile spans the interior), stable within a cell. */
int orc_sort(orc_ctx *o, int sp) {
  if (sp < 0 || sp > 1) return PIC_E_ARG;
  const double dx = o->cfg.dx;
  PList *L = &o->sp[sp];
  const int64_t ncell = (int64_t)o->NX * o->NY * o->NZ;
  int64_t *count = (int64_t *)calloc((size_t)ncell + 1, sizeof(int64_t));
  int64_t *key = (int64_t *)malloc(sizeof(int64_t) * (size_t)(L->n + 1));
  Particle *out = (Particle *)malloc(sizeof(Particle) * (size_t)(L->n + 1));
  if (!count || !key || !out) {
    free(count);
    free(key);
    free(out);
    return PIC_E_NOMEM;
  }
  for (int64_t i = 0; i < L->n; i++) {
    const Particle *p = &L->p[i];
    int k = (int)(p->x / dx) - 2;
    int l = (int)(p->y / dx) - 2;
    int m = (int)(p->z / dx) - 2;
    if (k >= 0 && k < o->NX && l >= 0 && l < o->NY && m >= 0 && m < o->NZ) {
      key[i] = cell_key(o, k, l, m);
      count[key[i] + 1]++;
    } else {
      key[i] = -1;
    }
  }
  for (int64_t c = 0; c < ncell; c++) count[c + 1] += count[c];
  int64_t total = count[ncell];
  for (int64_t i = 0; i < L->n; i++)
    if (key[i] >= 0) out[count[key[i]]++] = L->p[i];
  free(L->p);
  L->p = out;
  L->n = total;
  L->cap = L->n + 1;
  free(count);
  free(key);
  return 0;
}

/* ------------------------------------------------------------------ */
/* diagnostics: thrust_f (:4940-5122) force and energy sums over the central
 * 20-80 % box, plus whole-interior energies, momentum, charge and the Gauss
 * residual with the stencil of renew_phi3 (:8345-8348). */
int orc_diagnose(orc_ctx *o, pic_diag *out) {
  const pic_config *c = &o->cfg;
  const double dx = c->dx, dx3 = dx * dx * dx;
  memset(out, 0, sizeof(*out));
  double gl2 = 0, gmax = 0;
  for (int k = 2; k < o->NX + 2; k++)
    for (int l = 2; l < o->NY + 2; l++)
      for (int m = 2; m < o->NZ + 2; m++) {
        const Grid *g = &G(o, k, l, m);
        out->field_energy_e +=
            0.5 * (g->ex * g->ex + g->ey * g->ey + g->ez * g->ez) * c->eps0 * dx3;
        out->field_energy_b +=
            0.5 * (g->bx * g->bx + g->by * g->by + g->bz * g->bz) / c->mu0 * dx3;
        double b1x = g->bx - g->b0x, b1y = g->by - g->b0y, b1z = g->bz - g->b0z;
        out->field_energy_b1 +=
            0.5 * (b1x * b1x + b1y * b1y + b1z * b1z) / c->mu0 * dx3;
        out->charge += c->q * (g->ni - g->ne) * dx3;
        double r = -c->q * (g->ni - g->ne) / c->eps0 +
                   (-G(o, k + 2, l, m).ex + 27 * G(o, k + 1, l, m).ex -
                    27 * G(o, k, l, m).ex + G(o, k - 1, l, m).ex -
                    G(o, k, l + 2, m).ey + 27 * G(o, k, l + 1, m).ey -
                    27 * G(o, k, l, m).ey + G(o, k, l - 1, m).ey -
                    G(o, k, l, m + 2).ez + 27 * G(o, k, l, m + 1).ez -
                    27 * G(o, k, l, m).ez + G(o, k, l, m - 1).ez) /
                       (24. * dx);
        gl2 += r * r;
        if (fabs(r) > gmax) gmax = fabs(r);
        /* central box of thrust_f (:4968-4970) */
        if (k - 2 < o->NX * 0.8 && k - 2 > o->NX * 0.2 && l - 2 < o->NY * 0.8 &&
            l - 2 > o->NY * 0.2 && m - 2 < o->NZ * 0.8 && m - 2 > o->NZ * 0.2) {
          out->energy_box[0] +=
              0.5 * (g->ex * g->ex + g->ey * g->ey + g->ez * g->ez) * c->eps0;
          out->energy_box[1] +=
              0.5 * (g->bx * g->bx + g->by * g->by + g->bz * g->bz) / c->mu0;
        }
      }
  out->gauss_l2 = sqrt(gl2);
  out->gauss_max = gmax;
  for (int sp = 0; sp < 2; sp++) {
    const PList *L = &o->sp[sp];
    const double msp = c->mass[sp];
    const double w = c->weight[sp];
    const double sgn = sp == PIC_ION ? 1.0 : -1.0;
    out->count[sp] = L->n;
    for (int64_t i = 0; i < L->n; i++) {
      const Particle *p = &L->p[i];
      double v2 = p->vx * p->vx + p->vy * p->vy + p->vz * p->vz;
      double g = 1.0 / sqrt(1.0 - v2 / (c->c * c->c));
      out->kinetic[sp] += (g - 1.0) * msp * c->c * c->c * w * dx3;
      out->kinetic_nr[sp] += 0.5 * msp * w * dx3 * v2;
      out->momentum[sp][0] += g * msp * p->vx * w * dx3;
      out->momentum[sp][1] += g * msp * p->vy * w * dx3;
      out->momentum[sp][2] += g * msp * p->vz * w * dx3;
      if (p->x < o->NX * dx * 0.8 && p->x > o->NX * dx * 0.2 &&
          p->y < o->NY * dx * 0.8 && p->y > o->NY * dx * 0.2 &&
          p->z < o->NZ * dx * 0.8 && p->z > o->NZ * dx * 0.2) {
        double f[6];
        orc_gather(o, p->x, p->y, p->z, f);
        double ex = f[0], ey = f[1], ez = f[2], bx = f[3], by = f[4], bz = f[5];
        double qn = sgn * c->q * w * dx3;
        out->force[0] += qn * (ex + p->vy * bz - p->vz * by);
        out->force[1] += qn * (ey + p->vz * bx - p->vx * bz);
        out->force[2] += qn * (ez + p->vx * by - p->vy * bx);
        int b = sp == PIC_ION ? 3 : 9;
        out->force[b + 0] += qn * ex;
        out->force[b + 1] += qn * ey;
        out->force[b + 2] += qn * ez;
        out->force[b + 3] += qn * (p->vy * bz - p->vz * by);
        out->force[b + 4] += qn * (p->vz * bx - p->vx * bz);
        out->force[b + 5] += qn * (p->vx * by - p->vy * bx);
        out->energy_box[2 + sp] += 0.5 * msp * w * dx3 * v2;
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* fdtd() (:627-952) on one rank */

int orc_begin(orc_ctx *o) { return orc_inflow(o); }

int orc_step(orc_ctx *o, int nsteps) {
  const pic_config *c = &o->cfg;
  int rc;
  for (int s = 0; s < nsteps; s++) {
    orc_renew_grid(o);
    if (o->c_sort > c->sort_every - 1) {
      o->c_sort = 0;
      if ((rc = orc_sort(o, PIC_ION))) return rc;
      if ((rc = orc_sort(o, PIC_ELECTRON))) return rc;
    }
    if (o->c_add > c->inflow_every - 1) {
      if ((rc = orc_inflow(o))) return rc;
      o->c_add = 0;
    }
    orc_renew_b_half(o);
    orc_halo_b(o);
    orc_push_deposit_j(o, PIC_ION);
    orc_push_deposit_j(o, PIC_ELECTRON);
    orc_renew_b_half(o);
    orc_halo_b(o);
    orc_drift_deposit_n(o, PIC_ION);
    orc_drift_deposit_n(o, PIC_ELECTRON);
    orc_fold_current(o);
    orc_renew_e(o);
    orc_halo_e(o);
    if (c->poisson_every > 0 && o->step % c->poisson_every == 0) /* PHI() :901 */
      if ((rc = orc_poisson(o))) return rc;
    orc_particle_boundary(o);
    o->c_sort++;
    o->c_add++;
    o->step++;
  }
  return 0;
}

/*
 * pic2d_oracle.c - CPU oracle (test infrastructure) for the 2D3V step of
 * ashida-lab/PLASMAsim PIC2D/2d_pic278_mpi.c.
 *
 * FP64 restatement on one rank whose x neighbours are itself: every MPI
 * exchange becomes a local periodic copy or fold along x (sr_grid,
 * sr_current, sr_density, sr_particle, renew_phi's transposes), z is the
 * open flow axis.  Expressions keep the reference's operation order, so the
 * GPU's plain FP64 form can match to rounding.
 *
 * Deliberate differences from the reference (each is also a property the
 * GPU form shares, so parity is defined on identical inputs):
 *   - Random numbers: Philox4x32-10 streams keyed (seed, species, event,
 *     particle index) instead of GSL ranlxd2 (:410-416); normals by
 *     Box-Muller.  Loading and inflow place exactly ppc particles per cell
 *     (the reference draws uniformly over a thread's x range, :453-456).
 *   - One particle list per species, no thread split; particles that cross
 *     the periodic x faces are wrapped in place (the reference packs,
 *     appends and deletes, :812-901), so list order differs, not content.
 *   - Both x directions wrap: the reference's rank 0 and rank p-1 drop at
 *     the global x edges and refill there (:861-870, :3379-3432); one rank
 *     here is an interior rank whose neighbours are itself.
 *   - The Poisson solve (renew_phi :2589-2767) uses a direct DST-I sum
 *     instead of the Numerical-Recipes FFT (:2907-3023); both solve the same
 *     discrete Dirichlet system, so results agree to rounding.
 *   - Probes (:923-1006) are returned by orc2d_probe() for any point instead
 *     of written to files for three fixed ranks.
 *
 * Unpinned parts: none beyond the random streams, which no reference test
 * or fixture fixes.
 */
#include "pic2d_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "pic_oracle.h"

typedef struct {
  double bx, by, bz;
  double b0x, b0y, b0z;
  double ex, ey, ez;
  double jix0, jiy0, jiz0;
  double jex0, jey0, jez0;
  double ni, ne;
  double phi;
} G2; /* the reference's Grid (:102-111) */

typedef struct {
  double a[4]; /* jx, jy, jz, n of one thread (Grid_thread :113-118) */
} Acc2;

typedef struct {
  double x, z, vx, vy, vz;
} P2;

typedef struct {
  P2 *p;
  int64_t n, cap;
} PList2;

struct orc2d_ctx {
  pic2d_config cfg;
  int NX, NZ, SX, SZ;
  G2 *grid;
  double *sigma; /* 2SX x 2SZ (reference sigma :204) */
  PList2 sp[2];
  int nthreads;
  Acc2 *acc; /* nthreads x cells */
  double *rho; /* NX x NZ work plane of renew_phi */
  int64_t step;
  int c_add, c_sort;
  uint32_t inflow_event;
};

#define G(o, k, m) ((o)->grid[(size_t)(k) * (o)->SZ + (m)])
#define SIG(o, k, m) ((o)->sigma[(size_t)(k) * 2 * (o)->SZ + (m)])

static const double Pi = 3.14159265358979323846;

static int nthreads(void) { return orc_threads(0); }

/* ------------------------------------------------------------------ */
/* random streams: the same counter layout as the 3D oracle */

static inline double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) *
         (1.0 / 9007199254740992.0);
}

/* one macro-particle in the cell with lower corner (cx, cz): position
 * uniform in the cell, vx = vth N(0,1), vy = 0 (the reference's vy line is
 * commented out, :458-459), vz = V + vth N(0,1) */
static void draw2(const pic2d_config *c, int sp, uint32_t event, uint64_t idx,
                  double cx, double cz, P2 *p) {
  uint32_t key[2] = {(uint32_t)c->seed, (uint32_t)(c->seed >> 32)};
  uint32_t w[4][4];
  for (int d = 0; d < 4; d++) {
    uint32_t ctr[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), event,
                       (uint32_t)(4 * sp + d)};
    orc_philox4x32_10(ctr, key, w[d]);
  }
  const double dx = c->dx;
  double ux = u53(w[0][0], w[0][1]);
  double uz = u53(w[1][0], w[1][1]);
  double g0 = u53(w[1][2], w[1][3]);
  double g1 = u53(w[2][0], w[2][1]);
  double g2 = u53(w[2][2], w[2][3]);
  double g3 = u53(w[3][0], w[3][1]);
  const double two_pi = 6.283185307179586;
  double r0 = sqrt(-2.0 * log(1.0 - g0));
  double r1 = sqrt(-2.0 * log(1.0 - g2));
  double n0 = r0 * cos(two_pi * g1);
  double n2 = r1 * cos(two_pi * g3);
  p->x = cx + ux * dx;
  p->z = cz + uz * dx;
  p->vx = c->vth[sp] * n0 + 0.0;
  p->vy = 0.0;
  p->vz = c->vth[sp] * n2 + c->vdrift;
}

static int plist_reserve(PList2 *L, int64_t cap) {
  if (cap <= L->cap) return 0;
  P2 *np = (P2 *)realloc(L->p, (size_t)cap * sizeof(P2));
  if (!np) return -1;
  L->p = np;
  L->cap = cap;
  return 0;
}

/* ------------------------------------------------------------------ */
/* init_grid (:224-387) */

/* sigma depth on the 2x lattice along one axis, the layer of width A
 * mapped onto the reference's 32-cell profile (:270-283) */
static double depth2(int i2, int N, int A) {
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

orc2d_ctx *orc2d_create(const pic2d_config *cfg) {
  orc2d_ctx *o = (orc2d_ctx *)calloc(1, sizeof(orc2d_ctx));
  if (!o) return NULL;
  o->cfg = *cfg;
  o->NX = cfg->nx;
  o->NZ = cfg->nz;
  o->SX = cfg->nx + 4;
  o->SZ = cfg->nz + 4;
  const size_t ncell = (size_t)o->SX * o->SZ;
  o->nthreads = nthreads();
  o->grid = (G2 *)calloc(ncell, sizeof(G2));
  o->sigma = (double *)calloc(ncell * 4, sizeof(double));
  o->acc = (Acc2 *)calloc(ncell * (size_t)o->nthreads, sizeof(Acc2));
  o->rho = (double *)calloc((size_t)o->NX * o->NZ, sizeof(double));
  if (!o->grid || !o->sigma || !o->acc || !o->rho) {
    orc2d_destroy(o);
    return NULL;
  }
  /* B0 = IMF: the line-dipole expressions of :353-384 are overwritten by
   * bx = 0, bz = 0 in the reference, so the dipole field comes from
   * external_current_perp alone */
  for (size_t i = 0; i < ncell; i++) {
    G2 *g = &o->grid[i];
    g->b0x = g->bx = cfg->imf[0];
    g->b0y = g->by = cfg->imf[1];
    g->b0z = g->bz = cfg->imf[2];
  }
  for (int k = 0; k < 2 * o->SX; k++) {
    const double x = depth2(k, o->NX, cfg->absorb[0]);
    for (int m = 0; m < 2 * o->SZ; m++) {
      const double z = depth2(m, o->NZ, cfg->absorb[1]);
      SIG(o, k, m) = cfg->sigma_amp * (0.1 * sqrt(pow(x, 4) + pow(z, 4)) +
                                       0.01 * sqrt(pow(x, 12) + pow(z, 12)));
    }
  }
  return o;
}

void orc2d_destroy(orc2d_ctx *o) {
  if (!o) return;
  free(o->grid);
  free(o->sigma);
  free(o->acc);
  free(o->rho);
  free(o->sp[0].p);
  free(o->sp[1].p);
  free(o);
}

double orc2d_sigma(orc2d_ctx *o, int k2, int m2) { return SIG(o, k2, m2); }

/* ------------------------------------------------------------------ */
/* data movement */

static double *fptr(G2 *g, int which) {
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
    case PIC2D_PHI: return &g->phi;
    default: return NULL;
  }
}

int orc2d_upload_field(orc2d_ctx *o, int which, const double *host) {
  if (which < 0 || which >= PIC2D_NFIELDS) return PIC_E_ARG;
  const size_t n = (size_t)o->SX * o->SZ;
  for (size_t i = 0; i < n; i++) *fptr(&o->grid[i], which) = host[i];
  return 0;
}

int orc2d_download_field(orc2d_ctx *o, int which, double *host) {
  if (which < 0 || which >= PIC2D_NFIELDS) return PIC_E_ARG;
  const size_t n = (size_t)o->SX * o->SZ;
  for (size_t i = 0; i < n; i++) host[i] = *fptr(&o->grid[i], which);
  return 0;
}

int orc2d_upload_particles(orc2d_ctx *o, int sp, int64_t n, const double *x,
                           const double *z, const double *vx, const double *vy,
                           const double *vz) {
  if (sp < 0 || sp > 1 || n < 0) return PIC_E_ARG;
  PList2 *L = &o->sp[sp];
  if (plist_reserve(L, n > 0 ? n : 1)) return PIC_E_NOMEM;
  for (int64_t i = 0; i < n; i++) {
    L->p[i].x = x[i];
    L->p[i].z = z[i];
    L->p[i].vx = vx[i];
    L->p[i].vy = vy[i];
    L->p[i].vz = vz[i];
  }
  L->n = n;
  return 0;
}

int orc2d_download_particles(orc2d_ctx *o, int sp, int64_t cap, int64_t *n,
                             double *x, double *z, double *vx, double *vy,
                             double *vz) {
  if (sp < 0 || sp > 1) return PIC_E_ARG;
  PList2 *L = &o->sp[sp];
  if (cap < L->n) return PIC_E_CAPACITY;
  for (int64_t i = 0; i < L->n; i++) {
    x[i] = L->p[i].x;
    z[i] = L->p[i].z;
    vx[i] = L->p[i].vx;
    vy[i] = L->p[i].vy;
    vz[i] = L->p[i].vz;
  }
  *n = L->n;
  return 0;
}

int64_t orc2d_particle_count(orc2d_ctx *o, int sp) {
  return (sp < 0 || sp > 1) ? PIC_E_ARG : o->sp[sp].n;
}

/* ------------------------------------------------------------------ */
/* sort tiles (x, z) in cells; 0 = the whole interior */
static void tiles2(const orc2d_ctx *o, int *TX, int *TZ) {
  *TX = o->cfg.tile[0] > 0 ? o->cfg.tile[0] : o->NX;
  *TZ = o->cfg.tile[1] > 0 ? o->cfg.tile[1] : o->NZ;
}

/* init_particle (:389-525): ppc per interior cell.  Slot i belongs to the
 * cell of sort key i / ppc (tile-major, lexicographic (k,m) inside a tile:
 * the reference's (k,m) order when one tile spans the interior); i keys
 * event 0 of the stream. */
int orc2d_load_particles(orc2d_ctx *o) {
  const pic2d_config *c = &o->cfg;
  const int64_t n = (int64_t)o->NX * o->NZ * c->ppc;
  int TX, TZ;
  tiles2(o, &TX, &TZ);
  const int NTZ = o->NZ / TZ;
  for (int sp = 0; sp < 2; sp++) {
    PList2 *L = &o->sp[sp];
    if (plist_reserve(L, n > 0 ? n : 1)) return PIC_E_NOMEM;
    L->n = n;
#pragma omp parallel for num_threads(nthreads()) schedule(static)
    for (int64_t i = 0; i < n; i++) {
      const int64_t key = i / c->ppc;
      const int tile = (int)(key / (TX * TZ)), in = (int)(key % (TX * TZ));
      const int k = (tile / NTZ) * TX + in / TZ;
      const int m = (tile % NTZ) * TZ + in % TZ;
      draw2(c, sp, 0u, (uint64_t)i, (k + 2) * c->dx, (m + 2) * c->dx,
            &L->p[i]);
    }
  }
  return 0;
}

/* renew_grid (:1348-1389) */
int orc2d_renew_grid(orc2d_ctx *o) {
  const size_t n = (size_t)o->SX * o->SZ;
  for (size_t i = 0; i < n; i++) {
    G2 *g = &o->grid[i];
    g->jix0 = g->jiy0 = g->jiz0 = 0.;
    g->jex0 = g->jey0 = g->jez0 = 0.;
    g->ni = g->ne = 0.;
    g->phi = 0.;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* renew_b4 (:1391-1417), interior k in [2,Nx+2), m in [2,Nz+2) */
int orc2d_renew_b_half(orc2d_ctx *o) {
  const pic2d_config *c = &o->cfg;
  const double dtf = c->dt, dx = c->dx, e0 = c->eps0;
#pragma omp parallel for num_threads(nthreads()) schedule(static)
  for (int k = 2; k < o->NX + 2; k++) {
    for (int m = 2; m < o->NZ + 2; m++) {
      G2 *g = &G(o, k, m);
      double s;
      s = SIG(o, 2 * k + 1, 2 * m) / e0 * dtf * 0.5 * 0.5;
      g->bx = g->b0x + (1. - s) / (1. + s) * (g->bx - g->b0x) +
              0.5 * dtf / (1. + s) * 1. / 24. *
                  (27. * (G(o, k, m).ey - G(o, k, m - 1).ey) -
                   (G(o, k, m + 1).ey - G(o, k, m - 2).ey)) /
                  dx;
      s = SIG(o, 2 * k, 2 * m) / e0 * dtf * 0.5 * 0.5;
      g->by = g->b0y + (1. - s) / (1. + s) * (g->by - g->b0y) -
              dtf * 0.5 / (1. + s) *
                  (1. / 24. *
                       (27. * (G(o, k, m).ex - G(o, k, m - 1).ex) -
                        (G(o, k, m + 1).ex - G(o, k, m - 2).ex)) /
                       dx -
                   1. / 24. *
                       (27. * (G(o, k, m).ez - G(o, k - 1, m).ez) -
                        (G(o, k + 1, m).ez - G(o, k - 2, m).ez)) /
                       dx);
      s = SIG(o, 2 * k, 2 * m + 1) / e0 * dtf * 0.5 * 0.5;
      g->bz = g->b0z + (1. - s) / (1. + s) * (g->bz - g->b0z) -
              dtf * 0.5 / (1. + s) * 1. / 24. *
                  (27. * (G(o, k, m).ey - G(o, k - 1, m).ey) -
                   (G(o, k + 1, m).ey - G(o, k - 2, m).ey)) /
                  dx;
    }
  }
  return 0;
}

/* renew_e4 (:1419-1449); reads B - B0 and Ji - Je */
int orc2d_renew_e(orc2d_ctx *o) {
  const pic2d_config *c = &o->cfg;
  const double dtf = c->dt, dx = c->dx, mu0 = c->mu0, C = c->c;
#pragma omp parallel for num_threads(nthreads()) schedule(static)
  for (int k = 2; k < o->NX + 2; k++) {
    for (int m = 2; m < o->NZ + 2; m++) {
      G2 *g = &G(o, k, m);
      double s;
      s = 0.5 * mu0 * SIG(o, 2 * k, 2 * m + 1) * C * C * dtf;
      g->ex = (1. - s) / (1. + s) * g->ex +
              C * C * dtf / (1. + s) *
                  (-1. / 24. *
                       (27. * (G(o, k, m + 1).by - G(o, k, m + 1).b0y -
                               G(o, k, m).by + G(o, k, m).b0y) -
                        (G(o, k, m + 2).by - G(o, k, m + 2).b0y -
                         G(o, k, m - 1).by + G(o, k, m - 1).b0y)) /
                       dx -
                   mu0 * (g->jix0 - g->jex0));
      s = 0.5 * mu0 * SIG(o, 2 * k + 1, 2 * m) * C * C * dtf;
      g->ez = (1. - s) / (1. + s) * g->ez +
              C * C * dtf / (1. + s) *
                  (1. / 24. *
                       (27. * (G(o, k + 1, m).by - G(o, k + 1, m).b0y -
                               G(o, k, m).by + G(o, k, m).b0y) -
                        (G(o, k + 2, m).by - G(o, k + 2, m).b0y -
                         G(o, k - 1, m).by + G(o, k - 1, m).b0y)) /
                       dx -
                   mu0 * (g->jiz0 - g->jez0));
      s = 0.5 * mu0 * SIG(o, 2 * k + 1, 2 * m + 1) * C * C * dtf;
      g->ey = (1. - s) / (1. + s) * g->ey +
              C * C * dtf / (1. + s) *
                  (1. / 24. *
                       (27. * (G(o, k, m + 1).bx - G(o, k, m + 1).b0x -
                               G(o, k, m).bx + G(o, k, m).b0x) -
                        (G(o, k, m + 2).bx - G(o, k, m + 2).b0x -
                         G(o, k, m - 1).bx + G(o, k, m - 1).b0x)) /
                       dx -
                   1. / 24. *
                       (27. * (G(o, k + 1, m).bz - G(o, k + 1, m).b0z -
                               G(o, k, m).bz + G(o, k, m).b0z) -
                        (G(o, k + 2, m).bz - G(o, k + 2, m).b0z -
                         G(o, k - 1, m).bz + G(o, k - 1, m).b0z)) /
                       dx -
                   mu0 * (g->jiy0 - g->jey0));
    }
  }
  return 0;
}

/* sr_grid (:1021-1071) on one rank: the whole cell record of the two x
 * guard layers from the opposite interior layers */
int orc2d_halo(orc2d_ctx *o) {
  const int NX = o->NX;
  for (int m = 0; m < o->SZ; m++) {
    G(o, 0, m) = G(o, NX, m);
    G(o, 1, m) = G(o, NX + 1, m);
  }
  for (int m = 0; m < o->SZ; m++) {
    G(o, NX + 2, m) = G(o, 2, m);
    G(o, NX + 3, m) = G(o, 3, m);
  }
  return 0;
}

/* sr_current (:1073-1144): one layer each way, 2 += Nx+2, Nx+1 += 1 */
int orc2d_fold_current(orc2d_ctx *o) {
  const int NX = o->NX;
  for (int m = 0; m < o->SZ; m++) {
    G2 *d = &G(o, 2, m), *s = &G(o, NX + 2, m);
    d->jix0 += s->jix0;
    d->jiy0 += s->jiy0;
    d->jiz0 += s->jiz0;
    d->jex0 += s->jex0;
    d->jey0 += s->jey0;
    d->jez0 += s->jez0;
  }
  for (int m = 0; m < o->SZ; m++) {
    G2 *d = &G(o, NX + 1, m), *s = &G(o, 1, m);
    d->jix0 += s->jix0;
    d->jiy0 += s->jiy0;
    d->jiz0 += s->jiz0;
    d->jex0 += s->jex0;
    d->jey0 += s->jey0;
    d->jez0 += s->jez0;
  }
  return 0;
}

/* sr_density (:1146-1209) */
int orc2d_fold_density(orc2d_ctx *o) {
  const int NX = o->NX;
  for (int m = 0; m < o->SZ; m++) {
    G(o, 2, m).ni += G(o, NX + 2, m).ni;
    G(o, 2, m).ne += G(o, NX + 2, m).ne;
  }
  for (int m = 0; m < o->SZ; m++) {
    G(o, NX + 1, m).ni += G(o, 1, m).ni;
    G(o, NX + 1, m).ne += G(o, 1, m).ne;
  }
  return 0;
}

/* external_current_perp (:2238-2267): two line currents -I and +I along y,
 * assigned (not added) after the fold, ramped by (1+cos(w c - pi))/2 */
int orc2d_external_current(orc2d_ctx *o, int64_t step) {
  const pic2d_config *c = &o->cfg;
  const double I = c->ext_current, dx = c->dx;
  if (c->ext_k < 0) return 0;
  if (step < c->ext_ramp) {
    const double w = Pi / (double)c->ext_ramp;
    G(o, c->ext_k, c->ext_m[0]).jiy0 =
        -I / dx / dx * (1. + cos(w * (double)step - Pi)) / 2.;
    G(o, c->ext_k, c->ext_m[1]).jiy0 =
        I / dx / dx * (1. + cos(w * (double)step - Pi)) / 2.;
  } else {
    G(o, c->ext_k, c->ext_m[0]).jiy0 = -I / dx / dx;
    G(o, c->ext_k, c->ext_m[1]).jiy0 = I / dx / dx;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* shape_func_*3 (:1948-2011): bilinear gather on the staggered lattices,
 * the reference's operation order per component */
void orc2d_gather(orc2d_ctx *o, double x, double z, double out[6]) {
  const double dx = o->cfg.dx;
  int k, m;
  double w0, w1, w2, w3;
  /* shape_func_ex_bz3 */
  k = (int)(x / dx);
  m = (int)(z / dx - 0.5);
  w0 = ((k + 1) * dx - x) * ((m + 1.5) * dx - z);
  w1 = (x - k * dx) * ((m + 1.5) * dx - z);
  w2 = ((k + 1) * dx - x) * (z - (m + 0.5) * dx);
  w3 = (x - k * dx) * (z - (m + 0.5) * dx);
  out[0] = (G(o, k, m).ex * w0 + G(o, k + 1, m).ex * w1 +
            G(o, k, m + 1).ex * w2 + G(o, k + 1, m + 1).ex * w3) /
           (dx * dx);
  out[5] = (G(o, k, m).bz * w0 + G(o, k + 1, m).bz * w1 +
            G(o, k, m + 1).bz * w2 + G(o, k + 1, m + 1).bz * w3) /
           (dx * dx);
  /* shape_func_ey3 */
  k = (int)(x / dx - 0.5);
  m = (int)(z / dx - 0.5);
  out[1] = (G(o, k, m).ey * ((k + 1.5) * dx - x) * ((m + 1.5) * dx - z) +
            G(o, k + 1, m).ey * (x - (k + 0.5) * dx) * ((m + 1.5) * dx - z) +
            G(o, k, m + 1).ey * ((k + 1.5) * dx - x) * (z - (m + 0.5) * dx) +
            G(o, k + 1, m + 1).ey * (x - (k + 0.5) * dx) * (z - (m + 0.5) * dx)) /
           (dx * dx);
  /* shape_func_ez_bx3 */
  k = (int)(x / dx - 0.5);
  m = (int)(z / dx);
  w0 = ((k + 1.5) * dx - x) * ((m + 1) * dx - z);
  w1 = (x - (k + 0.5) * dx) * ((m + 1) * dx - z);
  w2 = ((k + 1.5) * dx - x) * (z - m * dx);
  w3 = (x - (k + 0.5) * dx) * (z - m * dx);
  out[2] = (G(o, k, m).ez * w0 + G(o, k + 1, m).ez * w1 +
            G(o, k, m + 1).ez * w2 + G(o, k + 1, m + 1).ez * w3) /
           (dx * dx);
  out[3] = (G(o, k, m).bx * w0 + G(o, k + 1, m).bx * w1 +
            G(o, k, m + 1).bx * w2 + G(o, k + 1, m + 1).bx * w3) /
           (dx * dx);
  /* shape_func_by3 */
  k = (int)(x / dx);
  m = (int)(z / dx);
  out[4] = (G(o, k, m).by * ((k + 1) * dx - x) * ((m + 1) * dx - z) +
            G(o, k + 1, m).by * (x - (k)*dx) * ((m + 1) * dx - z) +
            G(o, k, m + 1).by * ((k + 1) * dx - x) * (z - (m)*dx) +
            G(o, k + 1, m + 1).by * (x - (k)*dx) * (z - (m)*dx)) /
           (dx * dx);
}

/* cal_track_c4 (:1520-1560): non-relativistic Boris, then x, z += v dt/2 */
void orc2d_boris(const pic2d_config *cfg, double m, const double f[6],
                 double p[5]) {
  const double dt = cfg->dt;
  const double T = 0.5 * cfg->q * dt / m;
  double ex = f[0], ey = f[1], ez = f[2], bx = f[3], by = f[4], bz = f[5];
  ex += cfg->vdrift * cfg->imf[1];
  ey += -cfg->vdrift * cfg->imf[0];
  const double S = 2. * T / (1. + T * T * (bx * bx + by * by + bz * bz));
  const double ux = p[2] + ex * T;
  const double uy = p[3] + ey * T;
  const double uz = p[4] + ez * T;
  const double uox = ux + T * (uy * bz - uz * by);
  const double uoy = uy + T * (uz * bx - ux * bz);
  const double uoz = uz + T * (ux * by - uy * bx);
  p[2] = ux + S * (uoy * bz - uoz * by) + ex * T;
  p[3] = uy + S * (uoz * bx - uox * bz) + ey * T;
  p[4] = uz + S * (uox * by - uoy * bx) + ez * T;
  p[0] += p[2] * dt * 0.5;
  p[1] += p[4] * dt * 0.5;
}

/* ------------------------------------------------------------------ */
/* deposits (shape_func_*0_{jx,jy,jz,n} :2018-2144) into thread-private
 * accumulators, reduced in thread order as :722-734 and :790-798 do */

#define A2(k, m) (acc[(size_t)(k) * o->SZ + (m)])

static inline void dep_jy(orc2d_ctx *o, Acc2 *acc, const P2 *p, double qn) {
  const double dx = o->cfg.dx;
  const int k = (int)(p->x / dx - 0.5), m = (int)(p->z / dx - 0.5);
  const double temp = qn * p->vy / (dx * dx);
  A2(k, m).a[1] += temp * ((k + 1.5) * dx - p->x) * ((m + 1.5) * dx - p->z);
  A2(k + 1, m).a[1] += temp * (p->x - (k + 0.5) * dx) * ((m + 1.5) * dx - p->z);
  A2(k, m + 1).a[1] += temp * ((k + 1.5) * dx - p->x) * (p->z - (m + 0.5) * dx);
  A2(k + 1, m + 1).a[1] += temp * (p->x - (k + 0.5) * dx) * (p->z - (m + 0.5) * dx);
}

static inline void dep_jx(orc2d_ctx *o, Acc2 *acc, const P2 *p, double qn) {
  const double dx = o->cfg.dx;
  const int k = (int)(p->x / dx), m = (int)(p->z / dx - 0.5);
  const double temp = qn * p->vx / (dx * dx);
  A2(k, m).a[0] += temp * ((k + 1) * dx - p->x) * ((m + 1.5) * dx - p->z);
  A2(k + 1, m).a[0] += temp * (p->x - (k)*dx) * ((m + 1.5) * dx - p->z);
  A2(k, m + 1).a[0] += temp * ((k + 1) * dx - p->x) * (p->z - (m + 0.5) * dx);
  A2(k + 1, m + 1).a[0] += temp * (p->x - (k)*dx) * (p->z - (m + 0.5) * dx);
}

static inline void dep_jz(orc2d_ctx *o, Acc2 *acc, const P2 *p, double qn) {
  const double dx = o->cfg.dx;
  const int k = (int)(p->x / dx - 0.5), m = (int)(p->z / dx);
  const double temp = qn * p->vz / (dx * dx);
  A2(k, m).a[2] += temp * ((k + 1.5) * dx - p->x) * ((m + 1) * dx - p->z);
  A2(k + 1, m).a[2] += temp * (p->x - (k + 0.5) * dx) * ((m + 1) * dx - p->z);
  A2(k, m + 1).a[2] += temp * ((k + 1.5) * dx - p->x) * (p->z - (m)*dx);
  A2(k + 1, m + 1).a[2] += temp * (p->x - (k + 0.5) * dx) * (p->z - (m)*dx);
}

static inline void dep_n(orc2d_ctx *o, Acc2 *acc, const P2 *p, double n) {
  const double dx = o->cfg.dx;
  const int k = (int)(p->x / dx - 0.5), m = (int)(p->z / dx - 0.5);
  const double temp = n / (dx * dx);
  A2(k, m).a[3] += temp * ((k + 1.5) * dx - p->x) * ((m + 1.5) * dx - p->z);
  A2(k + 1, m).a[3] += temp * (p->x - (k + 0.5) * dx) * ((m + 1.5) * dx - p->z);
  A2(k, m + 1).a[3] += temp * ((k + 1.5) * dx - p->x) * (p->z - (m + 0.5) * dx);
  A2(k + 1, m + 1).a[3] += temp * (p->x - (k + 0.5) * dx) * (p->z - (m + 0.5) * dx);
}
#undef A2

/* the deposit gate of the particle loops: x in [dx,(Nx+3)dx),
 * z in [2dx,(Nz+2)dx) (:678, :766, :592) */
static inline int in_gate(const orc2d_ctx *o, const P2 *p) {
  const double dx = o->cfg.dx;
  return p->x < (o->NX + 3.) * dx && p->x >= dx && p->z < (o->NZ + 2.) * dx &&
         p->z >= 2. * dx;
}

/* safety rule (no reference counterpart, which has no bounds checks):
 * a particle is pushed only where its gather stencil lies inside the
 * array, -0.5 < x/dx < SX-1 and the same in z; elsewhere it only drifts.
 * Particles inside the interior at the start of a step always satisfy it. */
static inline int in_array(const orc2d_ctx *o, const P2 *p) {
  const double xs = p->x / o->cfg.dx, zs = p->z / o->cfg.dx;
  return xs > -0.5 && xs < o->SX - 1 && zs > -0.5 && zs < o->SZ - 1;
}

static void acc_zero(orc2d_ctx *o) {
  memset(o->acc, 0, (size_t)o->SX * o->SZ * o->nthreads * sizeof(Acc2));
}

/* grid.j? += sum over threads (:722-734): c0 = first component, nc = count,
 * dst = the grid offsets of those components */
static void acc_reduce(orc2d_ctx *o, int c0, int nc, const size_t *dst) {
  const size_t n = (size_t)o->SX * o->SZ;
  for (int t = 0; t < o->nthreads; t++) {
    const Acc2 *a = o->acc + (size_t)t * n;
    for (size_t i = 0; i < n; i++)
      for (int c = 0; c < nc; c++)
        *(double *)((char *)&o->grid[i] + dst[c]) += a[i].a[c0 + c];
  }
}

/* hot loop A (:665-720): ballistic drift or cal_track_c4, then the J
 * deposit at x^{n+1/2} */
int orc2d_push_deposit_j(orc2d_ctx *o, int sp) {
  if (sp < 0 || sp > 1) return PIC_E_ARG;
  const pic2d_config *c = &o->cfg;
  const double msp = sp == 0 ? c->mass[0] : -c->mass[1];
  const double qn = c->q * c->weight[sp];
  const double dt = c->dt;
  PList2 *L = &o->sp[sp];
  acc_zero(o);
  const size_t ncell = (size_t)o->SX * o->SZ;
  const int T = o->nthreads;
#pragma omp parallel num_threads(T)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    Acc2 *acc = o->acc + (size_t)t * ncell;
    const int64_t lo = L->n * t / T, hi = L->n * (t + 1) / T;
    for (int64_t i = lo; i < hi; i++) {
      P2 *p = &L->p[i];
      if (p->z < c->ballistic_lo[1] || p->z > c->ballistic_hi[1] ||
          p->x < c->ballistic_lo[0] || p->x > c->ballistic_hi[0] ||
          !in_array(o, p)) {
        p->x += p->vx * dt * 0.5;
        p->z += p->vz * dt * 0.5;
      } else {
        double f[6], q[5] = {p->x, p->z, p->vx, p->vy, p->vz};
        orc2d_gather(o, p->x, p->z, f);
        orc2d_boris(c, msp, f, q);
        p->x = q[0];
        p->z = q[1];
        p->vx = q[2];
        p->vy = q[3];
        p->vz = q[4];
      }
      if (in_gate(o, p)) {
        dep_jy(o, acc, p, qn);
        dep_jx(o, acc, p, qn);
        dep_jz(o, acc, p, qn);
      }
    }
  }
  size_t dst[3];
  if (sp == 0) {
    dst[0] = offsetof(G2, jix0);
    dst[1] = offsetof(G2, jiy0);
    dst[2] = offsetof(G2, jiz0);
  } else {
    dst[0] = offsetof(G2, jex0);
    dst[1] = offsetof(G2, jey0);
    dst[2] = offsetof(G2, jez0);
  }
  acc_reduce(o, 0, 3, dst);
  return 0;
}

static void deposit_n_all(orc2d_ctx *o, int sp, int drift) {
  const pic2d_config *c = &o->cfg;
  const double dt = c->dt, w = c->weight[sp];
  PList2 *L = &o->sp[sp];
  acc_zero(o);
  const size_t ncell = (size_t)o->SX * o->SZ;
  const int T = o->nthreads;
#pragma omp parallel num_threads(T)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    Acc2 *acc = o->acc + (size_t)t * ncell;
    const int64_t lo = L->n * t / T, hi = L->n * (t + 1) / T;
    for (int64_t i = lo; i < hi; i++) {
      P2 *p = &L->p[i];
      if (drift) {
        p->x += p->vx * dt * 0.5;
        p->z += p->vz * dt * 0.5;
      }
      if (in_gate(o, p)) dep_n(o, acc, p, w);
    }
  }
  size_t dst[1] = {sp == 0 ? offsetof(G2, ni) : offsetof(G2, ne)};
  acc_reduce(o, 3, 1, dst);
}

/* hot loop B (:752-788): half drift, density at x^{n+1} */
int orc2d_drift_deposit_n(orc2d_ctx *o, int sp) {
  if (sp < 0 || sp > 1) return PIC_E_ARG;
  deposit_n_all(o, sp, 1);
  return 0;
}

/* ------------------------------------------------------------------ */
/* renew_phi (:2589-2767) */

void orc2d_dst(int n, const double *y, double *out) {
  /* sin(pi j k / n) from the residue (j k) mod 2n, so every table entry is
   * one correctly rounded sine */
  double *tab = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  for (int r = 0; r < 2 * n; r++) tab[r] = sin(Pi * (double)r / (double)n);
  out[0] = 0.0;
  for (int k = 1; k < n; k++) {
    double s = 0.0;
    for (int j = 1; j < n; j++)
      s += y[j] * tab[(int)(((int64_t)j * k) % (2 * n))];
    out[k] = s;
  }
  free(tab);
}

int orc2d_poisson(orc2d_ctx *o) {
  const pic2d_config *c = &o->cfg;
  const int NX = o->NX, NZ = o->NZ;
  const double dx = c->dx, q = c->q, e0 = c->eps0;
  double *rho = o->rho;
  memset(rho, 0, sizeof(double) * (size_t)NX * NZ);
  /* source: q(ni-ne)/eps0 - div E at cell centres inside the box (:2614) */
  for (int k = c->phi_lo[0]; k < c->phi_hi[0]; k++)
    for (int m = c->phi_lo[1]; m < c->phi_hi[1]; m++)
      rho[(size_t)(k - 2) * NZ + (m - 2)] =
          -(-q * (G(o, k, m).ni - G(o, k, m).ne) / e0 +
            (G(o, k + 1, m).ex - G(o, k, m).ex) / dx +
            (G(o, k, m + 1).ez - G(o, k, m).ez) / dx);
  /* forward DST along z per x line, then along x per z mode (fft_z,
   * fft_x); divide by the 5-point symbol; inverse along x, then z */
  const int nmax = NX > NZ ? NX : NZ;
  /* the lines of each pass are independent: one per OpenMP iteration */
#pragma omp parallel num_threads(nthreads())
  {
    double *line = (double *)malloc(sizeof(double) * (size_t)nmax);
    double *tr = (double *)malloc(sizeof(double) * (size_t)nmax);
#pragma omp for schedule(static)
    for (int k = 0; k < NX; k++) {
      orc2d_dst(NZ, rho + (size_t)k * NZ, tr);
      memcpy(rho + (size_t)k * NZ, tr, sizeof(double) * NZ);
    }
#pragma omp for schedule(static)
    for (int m = 0; m < NZ; m++) {
      for (int k = 0; k < NX; k++) line[k] = rho[(size_t)k * NZ + m];
      orc2d_dst(NX, line, tr);
      for (int k = 0; k < NX; k++) {
        const double A = -(pow(sin(Pi * (double)m / (double)NZ / 2.), 2) +
                           pow(sin(Pi * (double)k / (double)NX / 2.), 2)) /
                         pow(dx / 2., 2);
        line[k] = A == 0. ? 0. : tr[k] / A;
      }
      orc2d_dst(NX, line, tr);
      for (int k = 0; k < NX; k++) rho[(size_t)k * NZ + m] = 2. * tr[k] / NX;
    }
#pragma omp for schedule(static)
    for (int k = 0; k < NX; k++) {
      orc2d_dst(NZ, rho + (size_t)k * NZ, tr);
      for (int m = 0; m < NZ; m++) rho[(size_t)k * NZ + m] = 2. * tr[m] / NZ;
    }
    free(line);
    free(tr);
  }
  /* E += grad phi; the x neighbour of k = 0 is the periodic image
   * (:2747-2764) */
  for (int k = 0; k < NX; k++)
    for (int m = 0; m < NZ; m++) {
      const double ph = rho[(size_t)k * NZ + m];
      const double pxm = k == 0 ? rho[(size_t)(NX - 1) * NZ + m]
                                : rho[(size_t)(k - 1) * NZ + m];
      G(o, k + 2, m + 2).ex += (ph - pxm) / dx;
      G(o, k + 2, m + 2).phi = ph;
      const double pzm = m == 0 ? 0. : rho[(size_t)k * NZ + m - 1];
      G(o, k + 2, m + 2).ez += (ph - pzm) / dx;
    }
  return 0;
}

/* ------------------------------------------------------------------ */
/* particle exchange and delete sweep (:812-901) on one rank: particles in
 * the first guard layer wrap by Nx dx, all others outside the interior
 * are deleted */
int orc2d_particle_boundary(orc2d_ctx *o) {
  const double dx = o->cfg.dx;
  const int NX = o->NX, NZ = o->NZ;
  for (int sp = 0; sp < 2; sp++) {
    PList2 *L = &o->sp[sp];
    int64_t w = 0;
    for (int64_t i = 0; i < L->n; i++) {
      P2 p = L->p[i];
      if (p.z < (NZ + 2.) * dx && p.z >= 2. * dx) {
        if (p.x < (NX + 3.) * dx && p.x >= (NX + 2.) * dx)
          p.x = p.x - NX * dx;
        else if (p.x > dx && p.x < 2. * dx)
          p.x = p.x + NX * dx;
      }
      if (p.x < 2. * dx || p.x >= (NX + 2.) * dx || p.z >= (NZ + 2.) * dx ||
          p.z < 2. * dx)
        continue;
      L->p[w++] = p;
    }
    L->n = w;
  }
  return 0;
}

/* add_ion5 / add_electron5 (:3311-3566) for the z faces: delete z < 3dx
 * or z > (Nz+1)dx, then append ppc particles per cell of the first and
 * last interior rows, low face first, cells in k order */
int orc2d_inflow(orc2d_ctx *o) {
  const pic2d_config *c = &o->cfg;
  const double dx = c->dx;
  const uint32_t ev = ++o->inflow_event;
  const int64_t per_face = (int64_t)o->NX * c->ppc;
  for (int sp = 0; sp < 2; sp++) {
    PList2 *L = &o->sp[sp];
    int64_t w = 0;
    for (int64_t i = 0; i < L->n; i++) {
      P2 p = L->p[i];
      if (p.z < 3 * dx || p.z > (o->NZ + 1) * dx) continue;
      L->p[w++] = p;
    }
    L->n = w;
    if (plist_reserve(L, L->n + 2 * per_face)) return PIC_E_NOMEM;
    const int64_t base = L->n;
    for (int face = 0; face < 2; face++)
      for (int64_t b = 0; b < per_face; b++) {
        const int k = (int)(b / c->ppc);
        const int64_t idx = face * per_face + b;
        const double zlo = face == 0 ? 2 * dx : (o->NZ + 1) * dx;
        draw2(c, sp, ev, (uint64_t)idx, (k + 2) * dx, zlo, &L->p[base + idx]);
      }
    L->n = base + 2 * per_face;
  }
  return 0;
}

/* tile-major key, lexicographic (k,m) inside a tile */
static int64_t key2(const orc2d_ctx *o, int k, int m) {
  int TX, TZ;
  tiles2(o, &TX, &TZ);
  const int NTZ = o->NZ / TZ;
  const int64_t tile = (int64_t)(k / TX) * NTZ + m / TZ;
  return tile * TX * TZ + (int64_t)(k % TX) * TZ + m % TZ;
}

/* stable counting sort by cell key; particles outside the interior are
 * dropped (as the 3D sorting does, :7497-7510) */
int orc2d_sort(orc2d_ctx *o, int sp) {
  if (sp < 0 || sp > 1) return PIC_E_ARG;
  const double dx = o->cfg.dx;
  PList2 *L = &o->sp[sp];
  const int64_t ncell = (int64_t)o->NX * o->NZ;
  int64_t *count = (int64_t *)calloc((size_t)ncell + 1, sizeof(int64_t));
  int64_t *key = (int64_t *)malloc(sizeof(int64_t) * (size_t)(L->n + 1));
  P2 *out = (P2 *)malloc(sizeof(P2) * (size_t)(L->n + 1));
  if (!count || !key || !out) {
    free(count);
    free(key);
    free(out);
    return PIC_E_NOMEM;
  }
  for (int64_t i = 0; i < L->n; i++) {
    const int k = (int)(L->p[i].x / dx) - 2, m = (int)(L->p[i].z / dx) - 2;
    if (in_array(o, &L->p[i]) && k >= 0 && k < o->NX && m >= 0 && m < o->NZ) {
      key[i] = key2(o, k, m);
      count[key[i] + 1]++;
    } else {
      key[i] = -1;
    }
  }
  for (int64_t c = 0; c < ncell; c++) count[c + 1] += count[c];
  const int64_t total = count[ncell];
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

/* pre-loop sequence of fdtd (:584-626) */
int orc2d_begin(orc2d_ctx *o) {
  orc2d_renew_grid(o);
  deposit_n_all(o, 0, 0);
  deposit_n_all(o, 1, 0);
  orc2d_fold_density(o);
  orc2d_halo(o);
  if (o->cfg.poisson_every > 0) orc2d_poisson(o);
  orc2d_halo(o);
  o->step = 0;
  o->c_add = 0;
  o->c_sort = 0;
  return 0;
}

/* the loop body of fdtd (:630-1014) */
int orc2d_step(orc2d_ctx *o, int nsteps) {
  const pic2d_config *c = &o->cfg;
  for (int s = 0; s < nsteps; s++) {
    int rc;
    orc2d_renew_grid(o);
    if (c->inflow_every > 0 && o->c_add >= c->inflow_every) {
      if ((rc = orc2d_inflow(o))) return rc;
      o->c_add = 0;
    }
    if (c->sort_every > 0 && o->c_sort >= c->sort_every) {
      if ((rc = orc2d_sort(o, 0)) || (rc = orc2d_sort(o, 1))) return rc;
      o->c_sort = 0;
    }
    orc2d_renew_b_half(o);
    orc2d_halo(o);
    orc2d_push_deposit_j(o, 0);
    orc2d_push_deposit_j(o, 1);
    orc2d_fold_current(o);
    orc2d_external_current(o, o->step);
    orc2d_renew_b_half(o);
    orc2d_halo(o);
    orc2d_drift_deposit_n(o, 0);
    orc2d_drift_deposit_n(o, 1);
    orc2d_fold_density(o);
    orc2d_renew_e(o);
    orc2d_halo(o);
    if (c->poisson_every > 0 && o->step % c->poisson_every == 0) {
      orc2d_poisson(o);
      orc2d_halo(o);
    }
    orc2d_particle_boundary(o);
    o->step++;
    o->c_add++;
    o->c_sort++;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* diagnostics */

int orc2d_diagnose(orc2d_ctx *o, pic2d_diag *d) {
  const pic2d_config *c = &o->cfg;
  const double dx = c->dx, dx2 = dx * dx;
  memset(d, 0, sizeof(*d));
  double se = 0, sb = 0, sb1 = 0, sq = 0, g2 = 0, gm = 0, b2 = 0, bm = 0;
  for (int k = 2; k < o->NX + 2; k++)
    for (int m = 2; m < o->NZ + 2; m++) {
      const G2 *g = &G(o, k, m);
      se += g->ex * g->ex + g->ey * g->ey + g->ez * g->ez;
      sb += g->bx * g->bx + g->by * g->by + g->bz * g->bz;
      const double b1x = g->bx - g->b0x, b1y = g->by - g->b0y,
                   b1z = g->bz - g->b0z;
      sb1 += b1x * b1x + b1y * b1y + b1z * b1z;
      sq += g->ni - g->ne;
      const double r = (G(o, k + 1, m).ex - g->ex + G(o, k, m + 1).ez - g->ez) / dx -
                       c->q * (g->ni - g->ne) / c->eps0;
      g2 += r * r;
      if (fabs(r) > gm) gm = fabs(r);
      const double db = (g->bx - g->b0x - G(o, k - 1, m).bx + G(o, k - 1, m).b0x +
                         g->bz - g->b0z - G(o, k, m - 1).bz + G(o, k, m - 1).b0z) /
                        dx;
      b2 += db * db;
      if (fabs(db) > bm) bm = fabs(db);
    }
  d->field_energy_e = 0.5 * c->eps0 * se * dx2;
  d->field_energy_b = 0.5 * sb / c->mu0 * dx2;
  d->field_energy_b1 = 0.5 * sb1 / c->mu0 * dx2;
  d->charge = c->q * sq * dx2;
  d->gauss_l2 = sqrt(g2);
  d->gauss_max = gm;
  d->divb_l2 = sqrt(b2);
  d->divb_max = bm;
  for (int sp = 0; sp < 2; sp++) {
    const PList2 *L = &o->sp[sp];
    double v2 = 0, px = 0, py = 0, pz = 0;
    for (int64_t i = 0; i < L->n; i++) {
      const P2 *p = &L->p[i];
      v2 += p->vx * p->vx + p->vy * p->vy + p->vz * p->vz;
      px += p->vx;
      py += p->vy;
      pz += p->vz;
    }
    const double mw = c->mass[sp] * c->weight[sp] * dx2;
    d->kinetic[sp] = 0.5 * mw * v2;
    d->momentum[sp][0] = mw * px;
    d->momentum[sp][1] = mw * py;
    d->momentum[sp][2] = mw * pz;
    d->count[sp] = L->n;
  }
  return 0;
}

/* the probe columns of fdtd (:929-948) at (k, m) */
int orc2d_probe(orc2d_ctx *o, int k, int m, double out[8]) {
  const pic2d_config *c = &o->cfg;
  const double dx = c->dx;
  if (k < 1 || m < 1 || k > o->SX - 2 || m > o->SZ - 2) return PIC_E_ARG;
  const G2 *g = &G(o, k, m), *gk = &G(o, k - 1, m), *gm = &G(o, k, m - 1),
           *gkp = &G(o, k + 1, m), *gmp = &G(o, k, m + 1);
  out[0] = g->ex;
  out[1] = g->ey;
  out[2] = g->ez;
  out[3] = g->bx;
  out[4] = g->by;
  out[5] = g->bz;
  out[6] = fabs(g->bx - g->b0x - gk->bx + gk->b0x + g->bz - g->b0z - gm->bz +
                gm->b0z) /
           dx /
           sqrt(pow(g->bx + gk->bx, 2) / 4. + pow(g->by, 2) +
                pow(g->bz + gm->bz, 2) / 4.);
  out[7] = fabs((gkp->ex - g->ex + gmp->ez - g->ez) / dx -
                c->q * (g->ni - g->ne) / c->eps0) /
           sqrt(pow(gkp->ex + g->ex, 2) / 4. + pow(g->ey, 2) +
                pow(gmp->ez + g->ez, 2) / 4.);
  return 0;
}

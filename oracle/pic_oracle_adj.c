/*
 * pic_oracle_adj.c - FP64 discrete adjoint of the 3D step (CPU oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the reference gradient the GPU adjoint is
 * checked against, itself checked by dot-product tests of every transposed
 * operator and by central finite differences of the forward.  The
 * reference has no adjoint; this differentiates the oracle's own step.
 *
 * Parameters p = (Mx, My, Mz, V): the dipole moment vector (the reference's
 * moment M (0, sin a, cos a) of external_current7 :4579-4634, generalised
 * to any direction) and the inflow drift speed V along z (:425, :6005).
 *
 * Loss: the time average over the N steps of the z (flow-direction) flux
 * of the Maxwell stress tensor through a sphere of radius R cells around
 * the dipole, F = sum_q w_q T_zj n_j, T_ij = eps0 (E_i E_j - |E|^2 d_ij/2)
 * + (B_i B_j - |B|^2 d_ij / 2) / mu0 with the total B, on a Fibonacci
 * lattice of nq points with equal weights 4 pi (R dx)^2 / nq; fields at a
 * point by the particle gather (shape_func_v2).
 *
 * Forward model: the oracle's step with the sort disabled; inflow and the
 * particle boundary mark deleted particles dead in place (z = NaN) instead
 * of compacting, so a particle keeps its slot for the whole run.  Every
 * quantity is then a function of the parameters through: B0 (gather, loss,
 * and the guard copies of the total B), the initial and injected
 * velocities v_z = V + sigma N(0,1), and the motional field V x IMF of the
 * push.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include "pic_oracle_internal.h"

#define IDX(o, k, l, m) ((((size_t)(k) * (o)->SY) + (l)) * (o)->SZ + (m))

static const double kPi = 3.14159265358979323846;

/* staggering per component (x, y, z; 1 = half-integer): Ex Ey Ez Bx By Bz,
 * also of Jx Jy Jz (shape_func_v2 :3104-3139, deposits :3397-3553) */
static const int STG[6][3] = {{0, 1, 1}, {1, 0, 1}, {1, 1, 0},
                              {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

typedef struct {
  int b[3];
  double w[3][2];
} Stn;

/* CIC stencil of a component at position p: base node (int)(x/dx - h) and
 * the weights ((b+1+h)dx - x, x - (b+h)dx) per axis */
static void stencil(double dx, const double p[3], const int st[3], Stn *s) {
  for (int a = 0; a < 3; a++) {
    const double h = st[a] ? 0.5 : 0.0;
    const int k = (int)(p[a] / dx - h);
    s->b[a] = k;
    s->w[a][0] = (k + 1 + h) * dx - p[a];
    s->w[a][1] = p[a] - (k + h) * dx;
  }
}

/* sum_n F[n] W_n and sum_n F[n] dW_n/dp (W unnormalised) */
static double st_gather(const orc_ctx *o, const double *F, const Stn *s,
                        double grad[3]) {
  double v = 0;
  grad[0] = grad[1] = grad[2] = 0;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int k = 0; k < 2; k++) {
        const double f = F[IDX(o, s->b[0] + i, s->b[1] + j, s->b[2] + k)];
        const double wx = s->w[0][i], wy = s->w[1][j], wz = s->w[2][k];
        const double sx = i ? 1.0 : -1.0, sy = j ? 1.0 : -1.0, sz = k ? 1.0 : -1.0;
        v += f * wx * wy * wz;
        grad[0] += f * sx * wy * wz;
        grad[1] += f * wx * sy * wz;
        grad[2] += f * wx * wy * sz;
      }
  return v;
}

static void st_scatter(const orc_ctx *o, double *F, const Stn *s, double val) {
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int k = 0; k < 2; k++)
        F[IDX(o, s->b[0] + i, s->b[1] + j, s->b[2] + k)] +=
            val * s->w[0][i] * s->w[1][j] * s->w[2][k];
}

/* ------------------------------------------------------------------ */
/* parameters */

/* field of a point dipole of moment M at (x, y, z) from its centre, the
 * component comp, with the softened core of the forward (the field at the
 * softening radius along the same ray) */
static double moment_field(const pic_config *c, double x, double y, double z,
                           const double M[3], int comp) {
  double r = sqrt(x * x + y * y + z * z);
  if (c->dipole_soft > 0.0 && r < c->dipole_soft * c->dx) {
    const double s = c->dipole_soft * c->dx / r;
    x *= s;
    y *= s;
    z *= s;
    r = sqrt(x * x + y * y + z * z);
  }
  const double rv[3] = {x, y, z};
  const double mr = M[0] * x + M[1] * y + M[2] * z;
  return c->mu0 * (3. * mr * rv[comp] - r * r * M[comp]) / (4. * kPi * pow(r, 5));
}

/* unit-moment fields G[a][comp] at every node, zero beyond dipole_rmax as
 * the forward's B0 (:4587-4592) */
static int adj_unit_fields(orc_ctx *o) {
  if (o->adjG) return 0;
  const size_t nc = (size_t)o->SX * o->SY * o->SZ;
  o->adjG = (double *)calloc(9 * nc, sizeof(double));
  if (!o->adjG) return PIC_E_NOMEM;
  const pic_config *c = &o->cfg;
  const double dx = c->dx;
  const int cx = (o->NX + 4) / 2, cy = (o->NY + 4) / 2, cz = (o->NZ + 4) / 2;
  for (int k = 0; k < o->SX; k++)
    for (int l = 0; l < o->SY; l++)
      for (int m = 0; m < o->SZ; m++) {
        const double x = (k - cx) * dx, y = (l - cy) * dx, z = (m - cz) * dx;
        if (!(sqrt(x * x + y * y + z * z) < c->dipole_rmax * dx)) continue;
        for (int a = 0; a < 3; a++) {
          double M[3] = {0, 0, 0};
          M[a] = 1.0;
          const size_t i = IDX(o, k, l, m);
          o->adjG[(3 * a + 0) * nc + i] =
              moment_field(c, (k - cx + 0.5) * dx, (l - cy) * dx, (m - cz) * dx, M, 0);
          o->adjG[(3 * a + 1) * nc + i] =
              moment_field(c, (k - cx) * dx, (l - cy + 0.5) * dx, (m - cz) * dx, M, 1);
          o->adjG[(3 * a + 2) * nc + i] =
              moment_field(c, (k - cx) * dx, (l - cy) * dx, (m - cz + 0.5) * dx, M, 2);
        }
      }
  return 0;
}

/* B0 = sum_a M_a G_a + IMF, B = B0, E = J = n = 0; V into the config */
static int adj_set_params(orc_ctx *o, const double prm[4]) {
  int rc = adj_unit_fields(o);
  if (rc) return rc;
  o->cfg.vdrift[2] = prm[3];
  const size_t nc = (size_t)o->SX * o->SY * o->SZ;
  const double *G = o->adjG;
  for (size_t i = 0; i < nc; i++) {
    Grid *g = &o->grid[i];
    memset(g, 0, sizeof(*g));
    g->b0x = prm[0] * G[0 * nc + i] + prm[1] * G[3 * nc + i] + prm[2] * G[6 * nc + i] +
             o->cfg.imf[0];
    g->b0y = prm[0] * G[1 * nc + i] + prm[1] * G[4 * nc + i] + prm[2] * G[7 * nc + i] +
             o->cfg.imf[1];
    g->b0z = prm[0] * G[2 * nc + i] + prm[1] * G[5 * nc + i] + prm[2] * G[8 * nc + i] +
             o->cfg.imf[2];
    g->bx = g->b0x;
    g->by = g->b0y;
    g->bz = g->b0z;
  }
  o->sp[0].n = o->sp[1].n = 0;
  o->step = 0;
  o->c_add = 0;
  o->c_sort = 0;
  o->inflow_event = 0;
  return 0;
}

/* ------------------------------------------------------------------ */
/* forward pieces with stable particle slots */

static int adj_inflow(orc_ctx *o, int64_t lo[2], int64_t hi[2]) {
  const pic_config *c = &o->cfg;
  const double dx = c->dx;
  o->inflow_event++;
  const uint32_t ev = o->inflow_event;
  for (int sp = 0; sp < 2; sp++) {
    PList *L = &o->sp[sp];
    for (int64_t i = 0; i < L->n; i++) {
      const double z = L->p[i].z;
      if (z < 4 * dx || z > o->NZ * dx) L->p[i].z = NAN;
    }
    const int nface = c->refill_far ? 2 : 1;
    const int64_t per_face = (int64_t)o->NX * o->NY * 2 * c->ppc;
    if (orc_plist_reserve(L, L->n + per_face * nface)) return PIC_E_NOMEM;
    const int64_t base = L->n;
    for (int face = 0; face < nface; face++)
      for (int64_t b = 0; b < per_face; b++) {
        const int m = (int)((b / c->ppc) % 2);
        const int l = (int)((b / (2 * c->ppc)) % o->NY);
        const int k = (int)(b / ((int64_t)2 * c->ppc * o->NY));
        const int64_t idx = face * per_face + b;
        const double zlo = face == 0 ? (m + 2) * dx : (m + o->NZ) * dx;
        orc_draw_particle(c->seed, sp, ev, (uint64_t)idx, (k + 2) * dx,
                          (l + 2) * dx, zlo, dx, c->vth[sp], c->vdrift,
                          &L->p[base + idx]);
      }
    lo[sp] = base;
    hi[sp] = L->n = base + per_face * nface;
  }
  return 0;
}

/* sr_particle on one rank with dead slots left in place */
static void adj_boundary(orc_ctx *o) {
  const pic_config *c = &o->cfg;
  const double dx = c->dx, dx2 = 2. * dx;
  const double gx2 = (o->NX + 2) * dx, gy2 = (o->NY + 2) * dx,
               gz2 = (o->NZ + 2) * dx;
  for (int sp = 0; sp < 2; sp++) {
    PList *L = &o->sp[sp];
    for (int64_t i = 0; i < L->n; i++) {
      Particle *p = &L->p[i];
      if (isnan(p->z)) continue;
      int dead = !(p->z < gz2 && p->z >= dx2);
      if (!dead) {
        if (p->x >= gx2) {
          if (c->periodic[0]) p->x = p->x - o->NX * dx; else dead = 1;
        } else if (p->x < dx2) {
          if (c->periodic[0]) p->x = p->x + o->NX * dx; else dead = 1;
        }
        if (p->y >= gy2) {
          if (c->periodic[1]) p->y = p->y - o->NY * dx; else dead = 1;
        } else if (p->y < dx2) {
          if (c->periodic[1]) p->y = p->y + o->NY * dx; else dead = 1;
        }
      }
      if (dead) p->z = NAN;
    }
  }
}

/* the loss quadrature: point q and its normal */
static void sphere_point(const orc_ctx *o, double R, int nq, int q, double P[3],
                         double n[3]) {
  const double dx = o->cfg.dx;
  const double z = 1.0 - (2.0 * q + 1.0) / nq;
  const double r = sqrt(1.0 - z * z);
  const double phi = q * kPi * (3.0 - sqrt(5.0));
  n[0] = r * cos(phi);
  n[1] = r * sin(phi);
  n[2] = z;
  const double c[3] = {((o->NX + 4) / 2) * dx, ((o->NY + 4) / 2) * dx,
                       ((o->NZ + 4) / 2) * dx};
  for (int a = 0; a < 3; a++) P[a] = c[a] + R * dx * n[a];
}

double orc_adj_force(orc_ctx *o, double R, int nq) {
  const pic_config *c = &o->cfg;
  const double w = 4.0 * kPi * (R * c->dx) * (R * c->dx) / nq;
  double F = 0;
  for (int q = 0; q < nq; q++) {
    double P[3], n[3], f[6];
    sphere_point(o, R, nq, q, P, n);
    orc_gather(o, P[0], P[1], P[2], f);
    const double *E = f, *B = f + 3;
    const double En = E[0] * n[0] + E[1] * n[1] + E[2] * n[2];
    const double Bn = B[0] * n[0] + B[1] * n[1] + B[2] * n[2];
    const double E2 = E[0] * E[0] + E[1] * E[1] + E[2] * E[2];
    const double B2 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
    F += w * (c->eps0 * (E[2] * En - 0.5 * n[2] * E2) +
              (B[2] * Bn - 0.5 * n[2] * B2) / c->mu0);
  }
  return F;
}

/* ------------------------------------------------------------------ */
/* tape */

typedef struct {
  int64_t n[2];
  Particle *p[2];
  double *f;          /* 6 planes: E, B1 = B - B0 */
  int inflow;
  int64_t lo[2], hi[2];
} Snap;

static int snap_take(orc_ctx *o, Snap *s) {
  const size_t nc = (size_t)o->SX * o->SY * o->SZ;
  s->f = (double *)malloc(6 * nc * sizeof(double));
  if (!s->f) return PIC_E_NOMEM;
  for (size_t i = 0; i < nc; i++) {
    const Grid *g = &o->grid[i];
    s->f[0 * nc + i] = g->ex;
    s->f[1 * nc + i] = g->ey;
    s->f[2 * nc + i] = g->ez;
    s->f[3 * nc + i] = g->bx - g->b0x;
    s->f[4 * nc + i] = g->by - g->b0y;
    s->f[5 * nc + i] = g->bz - g->b0z;
  }
  for (int sp = 0; sp < 2; sp++) {
    s->n[sp] = o->sp[sp].n;
    s->p[sp] = (Particle *)malloc(sizeof(Particle) * (size_t)(s->n[sp] + 1));
    if (!s->p[sp]) return PIC_E_NOMEM;
    memcpy(s->p[sp], o->sp[sp].p, sizeof(Particle) * (size_t)s->n[sp]);
  }
  return 0;
}

/* grid fields from a tape entry (E, B = B0 + B1) */
static void snap_fields(orc_ctx *o, const Snap *s) {
  const size_t nc = (size_t)o->SX * o->SY * o->SZ;
  for (size_t i = 0; i < nc; i++) {
    Grid *g = &o->grid[i];
    g->ex = s->f[0 * nc + i];
    g->ey = s->f[1 * nc + i];
    g->ez = s->f[2 * nc + i];
    g->bx = g->b0x + s->f[3 * nc + i];
    g->by = g->b0y + s->f[4 * nc + i];
    g->bz = g->b0z + s->f[5 * nc + i];
  }
}

static void snap_free(Snap *s) {
  free(s->f);
  free(s->p[0]);
  free(s->p[1]);
}

/* the forward: load, inflow, nsteps; the loss; the tape when asked */
static int adj_forward(orc_ctx *o, const double prm[4], int nsteps, double R,
                       int nq, Snap *tape, double *loss) {
  int rc;
  if ((rc = adj_set_params(o, prm))) return rc;
  if ((rc = orc_load_particles(o))) return rc;
  int64_t lo[2], hi[2];
  if ((rc = adj_inflow(o, lo, hi))) return rc;
  const pic_config *c = &o->cfg;
  double L = 0;
  for (int s = 0; s < nsteps; s++) {
    orc_renew_grid(o);
    int inflow = 0;
    if (o->c_add > c->inflow_every - 1) {
      if ((rc = adj_inflow(o, lo, hi))) return rc;
      o->c_add = 0;
      inflow = 1;
    }
    if (tape) {
      if ((rc = snap_take(o, &tape[s]))) return rc;
      tape[s].inflow = inflow;
      for (int sp = 0; sp < 2; sp++) {
        tape[s].lo[sp] = inflow ? lo[sp] : 0;
        tape[s].hi[sp] = inflow ? hi[sp] : 0;
      }
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
    adj_boundary(o);
    L += orc_adj_force(o, R, nq) / nsteps;
    o->c_add++;
    o->step++;
  }
  if (tape) {
    if ((rc = snap_take(o, &tape[nsteps]))) return rc;
    tape[nsteps].inflow = 0;
  }
  *loss = L;
  return 0;
}

double orc_adj_loss(orc_ctx *o, const double prm[4], int nsteps, double R,
                    int nq) {
  double L = NAN;
  if (adj_forward(o, prm, nsteps, R, nq, NULL, &L)) return NAN;
  return L;
}

/* ------------------------------------------------------------------ */
/* transposed field operators.  All adjoint planes span the whole array. */

typedef struct {
  double *E[3], *B[3], *B0[3], *J[2][3];
} Adj;

static double at(const orc_ctx *o, const double *g, int k, int l, int m) {
  if (k < 0 || l < 0 || m < 0 || k >= o->SX || l >= o->SY || m >= o->SZ) return 0.0;
  return g[IDX(o, k, l, m)];
}

/* (D_a^T g)[j] for D_a f[c] = 27 (f[c] - f[c-a]) - (f[c+a] - f[c-2a]) */
static double dT_back(const orc_ctx *o, const double *g, int k, int l, int m, int a) {
  int d[3] = {0, 0, 0};
  d[a] = 1;
#define G_(s) at(o, g, k + (s) * d[0], l + (s) * d[1], m + (s) * d[2])
  return 27. * G_(0) - 27. * G_(1) - G_(-1) + G_(2);
#undef G_
}

/* (G_a^T g)[j] for G_a f[c] = 27 (f[c+a] - f[c]) - (f[c+2a] - f[c-a]) */
static double dT_fwd(const orc_ctx *o, const double *g, int k, int l, int m, int a) {
  int d[3] = {0, 0, 0};
  d[a] = 1;
#define G_(s) at(o, g, k + (s) * d[0], l + (s) * d[1], m + (s) * d[2])
  return 27. * G_(-1) - 27. * G_(0) - G_(-2) + G_(1);
#undef G_
}

/* renew_b4_2 transposed: lB holds the adjoint of the updated B1; on return
 * the adjoint of the input B1, with -curl^T (h lB) added to lE */
static void renew_b_T(orc_ctx *o, Adj *A) {
  const pic_config *c = &o->cfg;
  const size_t nc = (size_t)o->SX * o->SY * o->SZ;
  const double temp = 1. / c->eps0 * c->dt * 0.5 * 0.5, dth = 0.5 * c->dt;
  const double sc = 1. / (24. * c->dx);
  double *g[3];
  for (int a = 0; a < 3; a++) g[a] = (double *)calloc(nc, sizeof(double));
  for (int k = 2; k < o->NX + 2; k++)
    for (int l = 2; l < o->NY + 2; l++)
      for (int m = 2; m < o->NZ + 2; m++) {
        const size_t i = IDX(o, k, l, m);
        const double s[3] = {SIG(o, 2 * k + 1, 2 * l, 2 * m),
                             SIG(o, 2 * k, 2 * l + 1, 2 * m),
                             SIG(o, 2 * k, 2 * l, 2 * m + 1)};
        for (int a = 0; a < 3; a++) {
          const double f = (1. - s[a] * temp) / (1. + s[a] * temp);
          const double h = dth / (1. + s[a] * temp);
          g[a][i] = h * sc * A->B[a][i];
          A->B[a][i] *= f;
        }
      }
  /* Bx' -= hx sc (D_l Ez - D_m Ey); By' -= hy sc (D_m Ex - D_k Ez);
   * Bz' -= hz sc (D_k Ey - D_l Ex) */
  for (int k = 0; k < o->SX; k++)
    for (int l = 0; l < o->SY; l++)
      for (int m = 0; m < o->SZ; m++) {
        const size_t i = IDX(o, k, l, m);
        A->E[2][i] += -dT_back(o, g[0], k, l, m, 1) + dT_back(o, g[1], k, l, m, 0);
        A->E[1][i] += dT_back(o, g[0], k, l, m, 2) - dT_back(o, g[2], k, l, m, 0);
        A->E[0][i] += -dT_back(o, g[1], k, l, m, 2) + dT_back(o, g[2], k, l, m, 1);
      }
  for (int a = 0; a < 3; a++) free(g[a]);
}

/* renew_e4_2 transposed: lE holds the adjoint of the updated E; on return
 * the adjoint of the input E; curl^T (h lE) added to lB, -mu0 h lE to the
 * ion current adjoint and +mu0 h lE to the electron one */
static void renew_e_T(orc_ctx *o, Adj *A) {
  const pic_config *c = &o->cfg;
  const size_t nc = (size_t)o->SX * o->SY * o->SZ;
  const double temp = c->c * c->c * c->dt, mu0 = c->mu0;
  const double sc = 1. / (24. * c->dx);
  double *u[3];
  for (int a = 0; a < 3; a++) u[a] = (double *)calloc(nc, sizeof(double));
  for (int k = 2; k < o->NX + 2; k++)
    for (int l = 2; l < o->NY + 2; l++)
      for (int m = 2; m < o->NZ + 2; m++) {
        const size_t i = IDX(o, k, l, m);
        const double s[3] = {SIG(o, 2 * k, 2 * l + 1, 2 * m + 1),
                             SIG(o, 2 * k + 1, 2 * l, 2 * m + 1),
                             SIG(o, 2 * k + 1, 2 * l + 1, 2 * m)};
        for (int a = 0; a < 3; a++) {
          const double f = (1. - 0.5 * mu0 * s[a] * temp) / (1. + 0.5 * mu0 * s[a] * temp);
          const double h = temp / (1. + 0.5 * mu0 * s[a] * temp);
          u[a][i] = h * sc * A->E[a][i];
          A->J[0][a][i] += -mu0 * h * A->E[a][i];
          A->J[1][a][i] += mu0 * h * A->E[a][i];
          A->E[a][i] *= f;
        }
      }
  /* Ex' += hx sc (G_l B1z - G_m B1y); Ey' += hy sc (G_m B1x - G_k B1z);
   * Ez' += hz sc (G_k B1y - G_l B1x) */
  for (int k = 0; k < o->SX; k++)
    for (int l = 0; l < o->SY; l++)
      for (int m = 0; m < o->SZ; m++) {
        const size_t i = IDX(o, k, l, m);
        A->B[2][i] += dT_fwd(o, u[0], k, l, m, 1) - dT_fwd(o, u[1], k, l, m, 0);
        A->B[1][i] += -dT_fwd(o, u[0], k, l, m, 2) + dT_fwd(o, u[2], k, l, m, 0);
        A->B[0][i] += dT_fwd(o, u[1], k, l, m, 2) - dT_fwd(o, u[2], k, l, m, 1);
      }
  for (int a = 0; a < 3; a++) free(u[a]);
}

/* one guard copy dst <- src transposed: lam[src] += lam[dst], lam[dst] = 0;
 * for B the copy is of the total field, so B1[dst] = B1[src] + B0[src] -
 * B0[dst] and the B0 adjoint takes +lam at src, -lam at dst */
static void copy_T(double *lam, double *lam0, size_t dst, size_t src) {
  const double v = lam[dst];
  lam[src] += v;
  lam[dst] = 0.0;
  if (lam0) {
    lam0[src] += v;
    lam0[dst] -= v;
  }
}

/* sr_b / sr_e transposed (x faces, then y faces, in the reverse order) */
static void halo_T(orc_ctx *o, double *lam[3], double *lam0[3]) {
  const int NX = o->NX, NY = o->NY;
  for (int c = 0; c < 3; c++) {
    double *L0 = lam0 ? lam0[c] : NULL;
    for (int k = 0; k < o->SX; k++)
      for (int m = 0; m < o->SZ; m++) {
        copy_T(lam[c], L0, IDX(o, k, NY + 3, m), IDX(o, k, 3, m));
        copy_T(lam[c], L0, IDX(o, k, NY + 2, m), IDX(o, k, 2, m));
        copy_T(lam[c], L0, IDX(o, k, 1, m), IDX(o, k, NY + 1, m));
        copy_T(lam[c], L0, IDX(o, k, 0, m), IDX(o, k, NY, m));
      }
    for (int l = 0; l < o->SY; l++)
      for (int m = 0; m < o->SZ; m++) {
        copy_T(lam[c], L0, IDX(o, NX + 3, l, m), IDX(o, 3, l, m));
        copy_T(lam[c], L0, IDX(o, NX + 2, l, m), IDX(o, 2, l, m));
        copy_T(lam[c], L0, IDX(o, 1, l, m), IDX(o, NX + 1, l, m));
        copy_T(lam[c], L0, IDX(o, 0, l, m), IDX(o, NX, l, m));
      }
  }
}

/* sr_current2 transposed: the four fold loops in reverse order; the
 * transpose of a[d] += a[s] is lam[s] += lam[d] */
static void fold_T(orc_ctx *o, double *lam[3]) {
  const int NX = o->NX, NY = o->NY;
  for (int c = 0; c < 3; c++) {
    double *L = lam[c];
    for (int k = 0; k < o->SX; k++)
      for (int m = 0; m < o->SZ; m++) {
        L[IDX(o, k, 0, m)] += L[IDX(o, k, NY, m)];
        L[IDX(o, k, 1, m)] += L[IDX(o, k, NY + 1, m)];
      }
    for (int k = 0; k < o->SX; k++)
      for (int m = 0; m < o->SZ; m++) {
        L[IDX(o, k, NY + 2, m)] += L[IDX(o, k, 2, m)];
        L[IDX(o, k, NY + 3, m)] += L[IDX(o, k, 3, m)];
      }
    for (int l = 0; l < o->SY; l++)
      for (int m = 0; m < o->SZ; m++) {
        L[IDX(o, 0, l, m)] += L[IDX(o, NX, l, m)];
        L[IDX(o, 1, l, m)] += L[IDX(o, NX + 1, l, m)];
      }
    for (int l = 0; l < o->SY; l++)
      for (int m = 0; m < o->SZ; m++) {
        L[IDX(o, NX + 2, l, m)] += L[IDX(o, 2, l, m)];
        L[IDX(o, NX + 3, l, m)] += L[IDX(o, 3, l, m)];
      }
  }
}

/* ------------------------------------------------------------------ */
/* cal_track4 transposed: the relativistic Boris step of orc_boris without
 * its position update.  lv' (in lvp) -> lv, lf (adjoints of E then B), and
 * the V derivative of the motional field */
static void boris_T(const pic_config *cfg, double m, const double f[6],
                    const double v[3], const double lvp[3], double lv[3],
                    double lf[6], double *gV) {
  const double C = cfg->c, T = 0.5 * cfg->q * cfg->dt / m;
  const double E[3] = {f[0] + cfg->vdrift[2] * cfg->imf[1],
                       f[1] - cfg->vdrift[2] * cfg->imf[0], f[2]};
  const double *B = f + 3;
  const double s1 = C * C - (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  const double g1 = s1 > 0 ? C / sqrt(s1) : 0.0;
  double u[3], b[3], um[3], uo[3], up[3], uq[3];
  for (int a = 0; a < 3; a++) u[a] = g1 * v[a];
  const double s2 = C * C + u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const double g2 = C / sqrt(s2);
  for (int a = 0; a < 3; a++) b[a] = g2 * B[a];
  const double D = 1. + T * T * (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  const double S = 2. * T / D;
  for (int a = 0; a < 3; a++) um[a] = u[a] + E[a] * T;
#define CROSS(r, p, q)                  \
  r[0] = p[1] * q[2] - p[2] * q[1];     \
  r[1] = p[2] * q[0] - p[0] * q[2];     \
  r[2] = p[0] * q[1] - p[1] * q[0]
  double t[3];
  CROSS(t, um, b);
  for (int a = 0; a < 3; a++) uo[a] = um[a] + T * t[a];
  CROSS(t, uo, b);
  for (int a = 0; a < 3; a++) up[a] = um[a] + S * t[a];
  for (int a = 0; a < 3; a++) uq[a] = up[a] + E[a] * T;
  const double s3 = C * C + uq[0] * uq[0] + uq[1] * uq[1] + uq[2] * uq[2];
  const double g3 = C / sqrt(s3);
  /* reverse */
  double luq[3], lE[3], lum[3], luo[3], lb[3], lu[3];
  const double dot3 = lvp[0] * uq[0] + lvp[1] * uq[1] + lvp[2] * uq[2];
  for (int a = 0; a < 3; a++) luq[a] = g3 * lvp[a] - C * pow(s3, -1.5) * dot3 * uq[a];
  for (int a = 0; a < 3; a++) lE[a] = T * luq[a];
  /* up = um + S (uo x b) */
  double uoxb[3];
  CROSS(uoxb, uo, b);
  const double lS = luq[0] * uoxb[0] + luq[1] * uoxb[1] + luq[2] * uoxb[2];
  CROSS(t, b, luq);
  for (int a = 0; a < 3; a++) {
    lum[a] = luq[a];
    luo[a] = S * t[a];
  }
  CROSS(t, luq, uo);
  for (int a = 0; a < 3; a++) lb[a] = S * t[a];
  /* uo = um + T (um x b) */
  CROSS(t, b, luo);
  for (int a = 0; a < 3; a++) lum[a] += luo[a] + T * t[a];
  CROSS(t, luo, um);
  for (int a = 0; a < 3; a++) lb[a] += T * t[a];
  /* S = 2T / (1 + T^2 |b|^2) */
  for (int a = 0; a < 3; a++) lb[a] += lS * (-S * S * T * b[a]);
  /* um = u + T E */
  for (int a = 0; a < 3; a++) {
    lu[a] = lum[a];
    lE[a] += T * lum[a];
  }
  /* b = g2 B */
  const double lg2 = lb[0] * B[0] + lb[1] * B[1] + lb[2] * B[2];
  for (int a = 0; a < 3; a++) lf[3 + a] = g2 * lb[a];
  for (int a = 0; a < 3; a++) lu[a] += lg2 * (-C * pow(s2, -1.5)) * u[a];
  /* u = g1 v */
  const double lg1 = lu[0] * v[0] + lu[1] * v[1] + lu[2] * v[2];
  for (int a = 0; a < 3; a++) lv[a] = g1 * lu[a];
  if (s1 > 0)
    for (int a = 0; a < 3; a++) lv[a] += lg1 * C * pow(s1, -1.5) * v[a];
  for (int a = 0; a < 3; a++) lf[a] = lE[a];
  *gV += lE[0] * cfg->imf[1] - lE[1] * cfg->imf[0];
#undef CROSS
}

/* ------------------------------------------------------------------ */
/* the particle pass transposed for one species.  The grid holds the
 * fields the forward pass saw (E, B after the first half step with
 * guards); lp holds the adjoints of the particles after the step and
 * returns those before it. */
static void pass_T(orc_ctx *o, int sp, const Snap *s, double (*lp)[6], Adj *A,
                   double *gV) {
  const pic_config *c = &o->cfg;
  const size_t nc = (size_t)o->SX * o->SY * o->SZ;
  const double dx = c->dx, dth = c->dt * 0.5, inv3 = 1. / (dx * dx * dx);
  const double m = sp == PIC_ION ? c->mass[0] : -c->mass[1];
  const double qn = c->q * c->weight[sp] * inv3;
  double *fld[6];
  /* the fields the gather reads, as planes */
  for (int a = 0; a < 6; a++) fld[a] = (double *)malloc(nc * sizeof(double));
  for (size_t i = 0; i < nc; i++) {
    const Grid *g = &o->grid[i];
    fld[0][i] = g->ex;
    fld[1][i] = g->ey;
    fld[2][i] = g->ez;
    fld[3][i] = g->bx;
    fld[4][i] = g->by;
    fld[5][i] = g->bz;
  }
  for (int64_t i = 0; i < s->n[sp]; i++) {
    const Particle P = s->p[sp][i];
    if (isnan(P.z)) continue;
    /* forward recomputation (orc_push_deposit_j, orc_drift_deposit_n) */
    Particle p1 = P;
    double f[6] = {0, 0, 0, 0, 0, 0};
    const int ball = P.x < c->ballistic_lo[0] || P.x > c->ballistic_hi[0] ||
                     P.y < c->ballistic_lo[1] || P.y > c->ballistic_hi[1] ||
                     P.z < c->ballistic_lo[2] || P.z > c->ballistic_hi[2];
    const double xd = P.x / dx, yd = P.y / dx, zd = P.z / dx;
    const int safe0 = xd > -0.5 && xd < o->SX - 1 && yd > -0.5 && yd < o->SY - 1 &&
                      zd > -0.5 && zd < o->SZ - 1;
    const int pushed = !ball && safe0;
    if (pushed) {
      double st[6] = {P.x, P.y, P.z, P.vx, P.vy, P.vz};
      orc_gather(o, P.x, P.y, P.z, f);
      orc_boris(c, m, f, st);
      p1.x = st[0];
      p1.y = st[1];
      p1.z = st[2];
      p1.vx = st[3];
      p1.vy = st[4];
      p1.vz = st[5];
    } else {
      p1.x += P.vx * dth;
      p1.y += P.vy * dth;
      p1.z += P.vz * dth;
    }
    const double x1d = p1.x / dx, y1d = p1.y / dx, z1d = p1.z / dx;
    const int dep = p1.x < (o->NX + 4) * dx && p1.x >= 0 && p1.y < (o->NY + 4) * dx &&
                    p1.y >= 0 && p1.z < (o->NZ + 4) * dx && p1.z >= 0 &&
                    x1d > -0.5 && x1d < o->SX - 1 && y1d > -0.5 && y1d < o->SY - 1 &&
                    z1d > -0.5 && z1d < o->SZ - 1;
    /* reverse: x2 = x1 + v' dt/2 (wrap is a shift) */
    double lx[3] = {lp[i][0], lp[i][1], lp[i][2]};
    double lvp[3] = {lp[i][3], lp[i][4], lp[i][5]};
    for (int a = 0; a < 3; a++) lvp[a] += lx[a] * dth;
    /* J deposit at x1 with v' */
    if (dep) {
      const double x1[3] = {p1.x, p1.y, p1.z};
      const double vp[3] = {p1.vx, p1.vy, p1.vz};
      for (int cc = 0; cc < 3; cc++) {
        Stn stn;
        double gr[3];
        stencil(dx, x1, STG[cc], &stn);
        const double val = st_gather(o, A->J[sp][cc], &stn, gr);
        lvp[cc] += qn * val;
        for (int a = 0; a < 3; a++) lx[a] += qn * vp[cc] * gr[a];
      }
    }
    /* x1 = x + v' dt/2 */
    for (int a = 0; a < 3; a++) lvp[a] += lx[a] * dth;
    double lv[3];
    if (pushed) {
      const double v0[3] = {P.vx, P.vy, P.vz};
      double lf[6];
      boris_T(c, m, f, v0, lvp, lv, lf, gV);
      const double x0[3] = {P.x, P.y, P.z};
      for (int cc = 0; cc < 6; cc++) {
        Stn stn;
        double gr[3];
        stencil(dx, x0, STG[cc], &stn);
        st_gather(o, fld[cc], &stn, gr);
        for (int a = 0; a < 3; a++) lx[a] += lf[cc] * inv3 * gr[a];
        if (cc < 3) {
          st_scatter(o, A->E[cc], &stn, lf[cc] * inv3);
        } else {
          st_scatter(o, A->B[cc - 3], &stn, lf[cc] * inv3);
          st_scatter(o, A->B0[cc - 3], &stn, lf[cc] * inv3);
        }
      }
    } else {
      for (int a = 0; a < 3; a++) lv[a] = lvp[a];
    }
    for (int a = 0; a < 3; a++) {
      lp[i][a] = lx[a];
      lp[i][3 + a] = lv[a];
    }
  }
  for (int a = 0; a < 6; a++) free(fld[a]);
}

/* d(loss term)/d(fields) of the force at the current grid, scaled */
static void force_T(orc_ctx *o, double R, int nq, double scale, Adj *A) {
  const pic_config *c = &o->cfg;
  const double dx = c->dx, inv3 = 1. / (dx * dx * dx);
  const double w = 4.0 * kPi * (R * dx) * (R * dx) / nq * scale;
  for (int q = 0; q < nq; q++) {
    double P[3], n[3], f[6];
    sphere_point(o, R, nq, q, P, n);
    orc_gather(o, P[0], P[1], P[2], f);
    const double *E = f, *B = f + 3;
    const double En = E[0] * n[0] + E[1] * n[1] + E[2] * n[2];
    const double Bn = B[0] * n[0] + B[1] * n[1] + B[2] * n[2];
    double lf[6];
    for (int a = 0; a < 3; a++) {
      lf[a] = w * c->eps0 * ((a == 2 ? En : 0.0) + E[2] * n[a] - n[2] * E[a]);
      lf[3 + a] = w / c->mu0 * ((a == 2 ? Bn : 0.0) + B[2] * n[a] - n[2] * B[a]);
    }
    for (int cc = 0; cc < 6; cc++) {
      Stn stn;
      stencil(dx, P, STG[cc], &stn);
      if (cc < 3) {
        st_scatter(o, A->E[cc], &stn, lf[cc] * inv3);
      } else {
        st_scatter(o, A->B[cc - 3], &stn, lf[cc] * inv3);
        st_scatter(o, A->B0[cc - 3], &stn, lf[cc] * inv3);
      }
    }
  }
}

static int adj_alloc(orc_ctx *o, Adj *A, double **block) {
  const size_t nc = (size_t)o->SX * o->SY * o->SZ;
  *block = (double *)calloc(15 * nc, sizeof(double));
  if (!*block) return PIC_E_NOMEM;
  double *b = *block;
  for (int a = 0; a < 3; a++) {
    A->E[a] = b + (0 + a) * nc;
    A->B[a] = b + (3 + a) * nc;
    A->B0[a] = b + (6 + a) * nc;
    A->J[0][a] = b + (9 + a) * nc;
    A->J[1][a] = b + (12 + a) * nc;
  }
  return 0;
}

int orc_adj_gradient(orc_ctx *o, const double prm[4], int nsteps, double R,
                     int nq, double *loss, double grad[4]) {
  if (nsteps <= 0 || nq <= 0) return PIC_E_ARG;
  Snap *tape = (Snap *)calloc((size_t)nsteps + 1, sizeof(Snap));
  if (!tape) return PIC_E_NOMEM;
  int rc = adj_forward(o, prm, nsteps, R, nq, tape, loss);
  Adj A;
  double *block = NULL;
  double (*lp[2])[6] = {NULL, NULL};
  const size_t nc = (size_t)o->SX * o->SY * o->SZ;
  double gV = 0;
  if (!rc) rc = adj_alloc(o, &A, &block);
  for (int sp = 0; sp < 2 && !rc; sp++) {
    lp[sp] = calloc((size_t)tape[nsteps].n[sp] + 1, sizeof(double[6]));
    if (!lp[sp]) rc = PIC_E_NOMEM;
  }
  for (int s = nsteps - 1; s >= 0 && !rc; s--) {
    /* loss term of the state after step s */
    snap_fields(o, &tape[s + 1]);
    force_T(o, R, nq, 1.0 / nsteps, &A);
    halo_T(o, A.E, NULL);
    renew_e_T(o, &A);
    fold_T(o, A.J[0]);
    fold_T(o, A.J[1]);
    halo_T(o, A.B, A.B0);
    renew_b_T(o, &A);
    /* the fields the particle pass read */
    snap_fields(o, &tape[s]);
    orc_renew_b_half(o);
    orc_halo_b(o);
    for (int sp = 0; sp < 2; sp++) pass_T(o, sp, &tape[s], lp[sp], &A, &gV);
    for (int sp = 0; sp < 2; sp++)
      for (int a = 0; a < 3; a++) memset(A.J[sp][a], 0, nc * sizeof(double));
    halo_T(o, A.B, A.B0);
    renew_b_T(o, &A);
    /* particles injected at the start of step s: v_z = V + sigma N */
    if (tape[s].inflow)
      for (int sp = 0; sp < 2; sp++)
        for (int64_t i = tape[s].lo[sp]; i < tape[s].hi[sp]; i++) {
          gV += lp[sp][i][5];
          memset(lp[sp][i], 0, sizeof(double[6]));
        }
  }
  if (!rc) {
    /* the loaded and first-injected particles */
    for (int sp = 0; sp < 2; sp++)
      for (int64_t i = 0; i < tape[0].n[sp]; i++)
        if (!isnan(tape[0].p[sp][i].z)) gV += lp[sp][i][5];
    for (int a = 0; a < 3; a++) {
      double g = 0;
      for (int cc = 0; cc < 3; cc++)
        for (size_t i = 0; i < nc; i++) g += A.B0[cc][i] * o->adjG[(3 * a + cc) * nc + i];
      grad[a] = g;
    }
    grad[3] = gV;
  }
  for (int s = 0; s <= nsteps; s++) snap_free(&tape[s]);
  free(tape);
  free(block);
  free(lp[0]);
  free(lp[1]);
  return rc;
}

/* ------------------------------------------------------------------ */
/* dot-product checks of the transposed operators: for each linear map A
 * of the forward and random x, y, <A x, y> against <x, A^T y> (relative
 * difference in out[0..3]: renew_b, renew_e, halo_b with B0, fold) */

static double rnd(uint64_t *s) {
  *s = *s * 6364136223846793005ull + 1442695040888963407ull;
  return ((double)(*s >> 11) / 9007199254740992.0) * 2.0 - 1.0;
}

static double rel(double a, double b) {
  const double d = fabs(a - b), m = fmax(fabs(a), fabs(b));
  return m > 0 ? d / m : d;
}

int orc_adj_selftest(orc_ctx *o, uint64_t seed, double out[4]) {
  const size_t nc = (size_t)o->SX * o->SY * o->SZ;
  Adj A;
  double *block;
  if (adj_alloc(o, &A, &block)) return PIC_E_NOMEM;
  double *x = (double *)malloc(12 * nc * sizeof(double));
  double *y = (double *)malloc(12 * nc * sizeof(double));
  uint64_t st = seed;
  /* renew_b: (E, B1) -> (E, B1') with B0 = 0 (the update is affine in B0
   * only through B - B0) */
  for (size_t i = 0; i < 12 * nc; i++) {
    x[i] = rnd(&st);
    y[i] = rnd(&st);
  }
  for (size_t i = 0; i < nc; i++) {
    Grid *g = &o->grid[i];
    memset(g, 0, sizeof(*g));
    g->ex = x[i];
    g->ey = x[nc + i];
    g->ez = x[2 * nc + i];
    g->bx = x[3 * nc + i];
    g->by = x[4 * nc + i];
    g->bz = x[5 * nc + i];
  }
  orc_renew_b_half(o);
  double lhs = 0, rhs = 0;
  for (size_t i = 0; i < nc; i++) {
    const Grid *g = &o->grid[i];
    lhs += g->ex * y[i] + g->ey * y[nc + i] + g->ez * y[2 * nc + i] +
           g->bx * y[3 * nc + i] + g->by * y[4 * nc + i] + g->bz * y[5 * nc + i];
  }
  memset(block, 0, 15 * nc * sizeof(double));
  for (int a = 0; a < 3; a++) {
    memcpy(A.E[a], y + a * nc, nc * sizeof(double));
    memcpy(A.B[a], y + (3 + a) * nc, nc * sizeof(double));
  }
  renew_b_T(o, &A);
  for (int a = 0; a < 3; a++)
    for (size_t i = 0; i < nc; i++)
      rhs += A.E[a][i] * x[a * nc + i] + A.B[a][i] * x[(3 + a) * nc + i];
  out[0] = rel(lhs, rhs);
  /* renew_e: (E, B1, Ji, Je) -> (E', B1, Ji, Je) */
  for (size_t i = 0; i < nc; i++) {
    Grid *g = &o->grid[i];
    memset(g, 0, sizeof(*g));
    g->ex = x[i];
    g->ey = x[nc + i];
    g->ez = x[2 * nc + i];
    g->bx = x[3 * nc + i];
    g->by = x[4 * nc + i];
    g->bz = x[5 * nc + i];
    g->jix0 = x[6 * nc + i];
    g->jiy0 = x[7 * nc + i];
    g->jiz0 = x[8 * nc + i];
    g->jex0 = x[9 * nc + i];
    g->jey0 = x[10 * nc + i];
    g->jez0 = x[11 * nc + i];
  }
  orc_renew_e(o);
  lhs = rhs = 0;
  for (size_t i = 0; i < nc; i++) {
    const Grid *g = &o->grid[i];
    lhs += g->ex * y[i] + g->ey * y[nc + i] + g->ez * y[2 * nc + i] +
           g->bx * y[3 * nc + i] + g->by * y[4 * nc + i] + g->bz * y[5 * nc + i] +
           g->jix0 * y[6 * nc + i] + g->jiy0 * y[7 * nc + i] + g->jiz0 * y[8 * nc + i] +
           g->jex0 * y[9 * nc + i] + g->jey0 * y[10 * nc + i] + g->jez0 * y[11 * nc + i];
  }
  memset(block, 0, 15 * nc * sizeof(double));
  for (int a = 0; a < 3; a++) {
    memcpy(A.E[a], y + a * nc, nc * sizeof(double));
    memcpy(A.B[a], y + (3 + a) * nc, nc * sizeof(double));
    memcpy(A.J[0][a], y + (6 + a) * nc, nc * sizeof(double));
    memcpy(A.J[1][a], y + (9 + a) * nc, nc * sizeof(double));
  }
  renew_e_T(o, &A);
  for (int a = 0; a < 3; a++)
    for (size_t i = 0; i < nc; i++)
      rhs += A.E[a][i] * x[a * nc + i] + A.B[a][i] * x[(3 + a) * nc + i] +
             A.J[0][a][i] * x[(6 + a) * nc + i] + A.J[1][a][i] * x[(9 + a) * nc + i];
  out[1] = rel(lhs, rhs);
  /* halo_b: (B1, B0) -> B1' = total copy minus B0 */
  for (size_t i = 0; i < nc; i++) {
    Grid *g = &o->grid[i];
    memset(g, 0, sizeof(*g));
    g->b0x = x[6 * nc + i];
    g->b0y = x[7 * nc + i];
    g->b0z = x[8 * nc + i];
    g->bx = g->b0x + x[3 * nc + i];
    g->by = g->b0y + x[4 * nc + i];
    g->bz = g->b0z + x[5 * nc + i];
  }
  orc_halo_b(o);
  lhs = rhs = 0;
  for (size_t i = 0; i < nc; i++) {
    const Grid *g = &o->grid[i];
    lhs += (g->bx - g->b0x) * y[3 * nc + i] + (g->by - g->b0y) * y[4 * nc + i] +
           (g->bz - g->b0z) * y[5 * nc + i];
  }
  memset(block, 0, 15 * nc * sizeof(double));
  for (int a = 0; a < 3; a++) memcpy(A.B[a], y + (3 + a) * nc, nc * sizeof(double));
  halo_T(o, A.B, A.B0);
  for (int a = 0; a < 3; a++)
    for (size_t i = 0; i < nc; i++)
      rhs += A.B[a][i] * x[(3 + a) * nc + i] + A.B0[a][i] * x[(6 + a) * nc + i];
  out[2] = rel(lhs, rhs);
  /* fold of the currents */
  for (size_t i = 0; i < nc; i++) {
    Grid *g = &o->grid[i];
    memset(g, 0, sizeof(*g));
    g->jix0 = x[6 * nc + i];
    g->jiy0 = x[7 * nc + i];
    g->jiz0 = x[8 * nc + i];
  }
  orc_fold_current(o);
  lhs = rhs = 0;
  for (size_t i = 0; i < nc; i++) {
    const Grid *g = &o->grid[i];
    lhs += g->jix0 * y[6 * nc + i] + g->jiy0 * y[7 * nc + i] + g->jiz0 * y[8 * nc + i];
  }
  memset(block, 0, 15 * nc * sizeof(double));
  for (int a = 0; a < 3; a++) memcpy(A.J[0][a], y + (6 + a) * nc, nc * sizeof(double));
  fold_T(o, A.J[0]);
  for (int a = 0; a < 3; a++)
    for (size_t i = 0; i < nc; i++) rhs += A.J[0][a][i] * x[(6 + a) * nc + i];
  out[3] = rel(lhs, rhs);
  free(x);
  free(y);
  free(block);
  return 0;
}

/* the transposed Boris step against central differences of orc_boris:
 * out[0] = max difference over the nine velocity and field derivatives of
 * <lvp, v'(v, f)>, relative to the largest of them */
int orc_adj_boris_check(const pic_config *cfg, double m, const double f[6],
                        const double v[3], const double lvp[3], double out[2]) {
  double lv[3], lf[6], gV = 0;
  boris_T(cfg, m, f, v, lvp, lv, lf, &gV);
  double worst = 0, scale = 0;
  for (int d = 0; d < 9; d++) {
    double fp[6], fm[6], sp[6], sm[6];
    memcpy(fp, f, sizeof(fp));
    memcpy(fm, f, sizeof(fm));
    for (int a = 0; a < 3; a++) sp[3 + a] = sm[3 + a] = v[a];
    sp[0] = sp[1] = sp[2] = sm[0] = sm[1] = sm[2] = 0;
    double h;
    if (d < 3) {
      h = 1e-6 * fmax(fabs(v[d]), 1e-3);
      sp[3 + d] += h;
      sm[3 + d] -= h;
    } else {
      h = 1e-6 * fmax(fabs(f[d - 3]), 1e-3);
      fp[d - 3] += h;
      fm[d - 3] -= h;
    }
    orc_boris(cfg, m, fp, sp);
    orc_boris(cfg, m, fm, sm);
    double fd = 0;
    for (int a = 0; a < 3; a++) fd += lvp[a] * (sp[3 + a] - sm[3 + a]) / (2 * h);
    const double an = d < 3 ? lv[d] : lf[d - 3];
    worst = fmax(worst, fabs(fd - an));
    scale = fmax(scale, fabs(fd));
  }
  out[0] = scale > 0 ? worst / scale : worst;
  out[1] = gV;
  return 0;
}

/*
 * pic_oracle_internal.h - types shared by the CPU oracle's translation
 * units (pic_oracle.c, pic_oracle_adj.c).  TEST INFRASTRUCTURE ONLY.
 */
#ifndef PIC_ORACLE_INTERNAL_H
#define PIC_ORACLE_INTERNAL_H

#include "pic_oracle.h"

/* Grid cell record, as the reference's Grid (:115-123). */
typedef struct {
  double bx, by, bz;
  double b0x, b0y, b0z;
  double ex, ey, ez;
  double jix0, jiy0, jiz0;
  double jex0, jey0, jez0;
  double ni, ne;
  double phi;
} Grid;

/* Thread-private accumulator (reference Grid_thread, :125-129), reduced to
 * the four quantities one deposit call writes. */
typedef struct {
  double a[4];
} Acc;

typedef struct {
  double x, y, z, vx, vy, vz;
} Particle;

typedef struct {
  Particle *p;
  int64_t n, cap;
} PList;

struct orc_ctx {
  pic_config cfg;
  int NX, NY, NZ;          /* interior sizes */
  int SX, SY, SZ;          /* sizes with guards */
  Grid *grid;
  double *sigma;           /* 2SX x 2SY x 2SZ (reference sigma :148) */
  PList sp[2];
  int nthreads;
  Acc *acc;                /* nthreads x cells */
  int64_t step;            /* c */
  int c_add, c_sort;       /* reference c_add, c_thrust */
  uint32_t inflow_event;   /* Philox event id of the next inflow */
  double *adjG;            /* adjoint: B0 of unit moments, 9 planes */
};

#define G(o, k, l, m) ((o)->grid[(((size_t)(k) * (o)->SY) + (l)) * (o)->SZ + (m)])
#define SIG(o, k, l, m) \
  ((o)->sigma[(((size_t)(k) * 2 * (o)->SY) + (l)) * 2 * (o)->SZ + (m)])

/* one macro-particle of the stream (seed, species, event, index) */
void orc_draw_particle(uint64_t seed, int sp, uint32_t event, uint64_t idx,
                       double cx, double cy, double cz, double dx,
                       double vth, const double mu[3], Particle *p);
int orc_plist_reserve(PList *L, int64_t cap);

#endif

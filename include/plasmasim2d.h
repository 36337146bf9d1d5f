/*
 * plasmasim2d.h - C-ABI of the MI355X-native 2D3V electromagnetic PIC step.
 *
 * The 2D3V reference program (ashida-lab/PLASMAsim, PIC2D/2d_pic278_mpi.c)
 * has the same structure as the 3D one: fdtd() (:527-1019) calls a fixed
 * list of C routines on global arrays.  Each entry point below replaces one
 * of them on the device-resident state of a pic2d_ctx.
 *
 * Conventions
 *   - Axes and indices follow the reference: k (x, periodic) and m (z, flow
 *     axis, contiguous), 2 guard cells per side, so field arrays are
 *     (nx+4)*(nz+4) and element (k,m) lives at k*(nz+4)+m.  Fields do not
 *     depend on y; particles carry vy.
 *   - Particle coordinates are in length units with the +2 guard offset
 *     (interior cell k covers [(k+2)dx,(k+3)dx)).
 *   - Staggering (renew_b4/renew_e4 :1391-1449, shape_func_*3 :1948-2011):
 *     Ex,Jx at (k, m+1/2); Ey,Jy at (k+1/2, m+1/2); Ez,Jz at (k+1/2, m);
 *     Bx at (k+1/2, m); By at (k, m); Bz at (k, m+1/2); ni, ne and phi at
 *     (k+1/2, m+1/2).
 *   - Status codes, numeric forms, deposit modes, species and field
 *     selectors are those of plasmasim.h; PIC2D_PHI adds the potential of
 *     the last Poisson correction.
 *   - Every call returns 0 or a negative PIC_E_* code and never exits;
 *     pic2d_last_error() gives the message.  Calls are asynchronous on the
 *     context's HIP stream; downloads, diagnostics and probes synchronize.
 */
#ifndef PLASMASIM2D_H
#define PLASMASIM2D_H

#include <stddef.h>
#include <stdint.h>

#include "plasmasim.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PIC2D_PHI PIC_NFIELDS /* 17: phi of the last Poisson correction */
#define PIC2D_NFIELDS (PIC_NFIELDS + 1)

/*
 * Runtime replacement of the 2D reference's compile-time configuration
 * (:29-80), normalized units for the shipped presets.
 */
typedef struct pic2d_config {
  int32_t nx, nz;        /* interior cells: x periodic, z flow (Grid_Nx/Nz) */
  int32_t ppc;           /* macro-particles per cell per species (Np*Np) */
  double dx, dt;
  double c, eps0, mu0;
  double q;              /* elementary charge; both species deposit with +q */
  double mass[2];        /* ion, electron (positive; the push uses -m_e) */
  double weight[2];      /* density carried by one macro-particle (N0i/N0e) */
  double vth[2];         /* Maxwellian sigma of vx and vz, sqrt(kT/m);
                            vy starts at 0 as in the reference (:459) */
  double vdrift;         /* inflow speed V along +z */
  double imf[3];         /* uniform B0; the push adds E += V*(IMF_y, -IMF_x) */
  double ext_current;    /* line-current amplitude I (external_current_perp) */
  int32_t ext_ramp;      /* steps of the (1+cos)/2 ramp of I (reference 10000) */
  int32_t ext_k;         /* array index k of the two line currents */
  int32_t ext_m[2];      /* array indices m of the -I and +I line currents */
  int32_t absorb[2];     /* sigma-layer width in cells, x and z (0 = none) */
  double sigma_amp;      /* sigma profile amplitude (reference a, :229) */
  double ballistic_lo[2];/* particles with x or z below/above these */
  double ballistic_hi[2];/* coordinates only drift (Absorb_grid2, :669-674) */
  int32_t phi_lo[2];     /* array-index box [lo,hi) of the Poisson source, */
  int32_t phi_hi[2];     /* x and z (Absorb_grid3 margins, :2614-2639)    */
  int32_t poisson_every; /* renew_phi cadence in steps (reference 1, 0 = off) */
  int32_t inflow_every;  /* add_*5 cadence in steps (reference 250) */
  int32_t sort_every;    /* tile sort cadence (0 = never; not in reference) */
  int32_t precision;     /* PIC_FP64 or PIC_FP32 */
  int32_t deposit;       /* PIC_DEPOSIT_ATOMIC or PIC_DEPOSIT_DETERMINISTIC */
  int32_t tile[2];       /* sort tiles in cells (x, z); 0 = whole interior */
  uint64_t seed;         /* Philox4x32-10 key for loading and inflow */
  int64_t capacity;      /* per-species particle capacity; 0 = automatic */
} pic2d_config;

/* Diagnostics, FP64 over the interior (energies per unit length in y). */
typedef struct pic2d_diag {
  double field_energy_e;   /* 1/2 eps0 sum |E|^2 dx^2 */
  double field_energy_b;   /* 1/2 sum |B|^2 / mu0 dx^2 (total B) */
  double field_energy_b1;  /* same for B - B0 */
  double kinetic[2];       /* 1/2 m w dx^2 sum v^2 per species */
  double momentum[2][3];   /* m w dx^2 sum v per species */
  double charge;           /* sum q (ni - ne) dx^2 */
  double gauss_l2;         /* sqrt(sum r^2), r = div E - q(ni-ne)/eps0 at
                              cell centres (the probe stencil, :940-944) */
  double gauss_max;        /* max |r| */
  double divb_l2;          /* sqrt(sum d^2), d = div(B - B0) at nodes */
  double divb_max;
  int64_t count[2];        /* live particles per species */
} pic2d_diag;

typedef struct pic2d_ctx pic2d_ctx;

size_t pic2d_sizeof_config(void);
size_t pic2d_sizeof_diag(void);

/* Presets: 0 = 32 x 64 parity case (8 ppc), 1 = 256 x 256 (16 ppc),
 * 3 = 2048 x 2048 x 100 ppc (838,860,800 particles). */
int pic2d_config_preset(int which, pic2d_config *out);

/* init_grid (:224-387): sigma, B0 = IMF; allocates device state. */
int pic2d_create(const pic2d_config *cfg, pic2d_ctx **out);
void pic2d_destroy(pic2d_ctx *ctx);
const char *pic2d_last_error(const pic2d_ctx *ctx);
int pic2d_get_config(const pic2d_ctx *ctx, pic2d_config *out);
void *pic2d_stream(pic2d_ctx *ctx);
int pic2d_sync(pic2d_ctx *ctx);

/* fields in reference layout [k][m], FP64 on the host */
int pic2d_upload_field(pic2d_ctx *ctx, int which, const double *host);
int pic2d_download_field(pic2d_ctx *ctx, int which, double *host);
/* particles as x, z, vx, vy, vz (FP64 SoA, reference units); downloads
 * return the live particles in storage order */
int pic2d_upload_particles(pic2d_ctx *ctx, int species, int64_t n,
                           const double *x, const double *z, const double *vx,
                           const double *vy, const double *vz);
int pic2d_download_particles(pic2d_ctx *ctx, int species, int64_t cap,
                             int64_t *n, double *x, double *z, double *vx,
                             double *vy, double *vz);
int pic2d_particle_count(pic2d_ctx *ctx, int species, int64_t *n);

/* init_particle (:389-525): ppc per interior cell, drifting Maxwellian */
int pic2d_load_particles(pic2d_ctx *ctx);
/* the pre-loop sequence of fdtd (:584-626): density of all particles,
 * sr_density, sr_grid, renew_phi, sr_grid */
int pic2d_begin(pic2d_ctx *ctx);

/* the routines of the step, in fdtd's order (:630-1014) */
int pic2d_renew_grid(pic2d_ctx *ctx);                  /* :1348-1389 */
int pic2d_inflow(pic2d_ctx *ctx);                      /* add_*5 :3311-3566 */
int pic2d_renew_b_half(pic2d_ctx *ctx);                /* renew_b4 :1391 */
int pic2d_halo(pic2d_ctx *ctx);                        /* sr_grid :1021 */
int pic2d_push_deposit_j(pic2d_ctx *ctx, int species); /* :665-727 */
int pic2d_fold_current(pic2d_ctx *ctx);                /* sr_current :1073 */
int pic2d_external_current(pic2d_ctx *ctx, int64_t step); /* :2238-2267 */
int pic2d_drift_deposit_n(pic2d_ctx *ctx, int species);   /* :752-798 */
int pic2d_fold_density(pic2d_ctx *ctx);                /* sr_density :1146 */
int pic2d_renew_e(pic2d_ctx *ctx);                     /* renew_e4 :1419 */
int pic2d_poisson(pic2d_ctx *ctx);                     /* renew_phi :2589 */
int pic2d_particle_boundary(pic2d_ctx *ctx);           /* :812-901 */
/* per-tile stable sort (no reference counterpart in the 2D loop) */
int pic2d_sort(pic2d_ctx *ctx, int species);
/* how many sorts so far took the tile sort (FP32 form, tiles current since
 * the last sort) and the radix sort (no counterpart in the reference) */
int pic2d_sort_stats(const pic2d_ctx *ctx, int64_t *tile, int64_t *radix);

/* n whole steps of fdtd's loop with the particle loops and the particle
 * boundary fused into one pass per species.  FP32 form: a step's tile sort
 * copies no particles, the species' pass of that step reads them through
 * the sort's permutation (timed as PIC_T_PUSH_SORT). */
int pic2d_step(pic2d_ctx *ctx, int nsteps);
int pic2d_step_index(const pic2d_ctx *ctx, int64_t *step);

int pic2d_diagnose(pic2d_ctx *ctx, pic2d_diag *out);
/* the probe line of fdtd (:924-1006) at array point (k, m): ex, ey, ez, bx,
 * by, bz, |div B|/|B| and |div E - q(ni-ne)/eps0|/|E| */
int pic2d_probe(pic2d_ctx *ctx, int k, int m, double out[8]);

/* per-class kernel timing with HIP events (classes of plasmasim.h) */
int pic2d_timing_enable(pic2d_ctx *ctx, int on);
int pic2d_timing_get(pic2d_ctx *ctx, int cls, double *ms, int64_t *launches);
int pic2d_timing_reset(pic2d_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif

/*
 * pic_oracle.h - CPU oracle for the 3D EM-PIC step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C, FP64 restatement of the
 * reference routines of ashida-lab/PLASMAsim PIC/3d_pic458_domain.c on one
 * rank.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and only as the checker or the timed CPU baseline.  The
 * product library (plasmasim_amd) never calls it.
 *
 * Every function cites the reference lines it follows.  Where the restatement
 * deliberately differs from the reference, the comment says so and why.
 */
#ifndef PIC_ORACLE_H
#define PIC_ORACLE_H

#include <stdint.h>
#include "../include/plasmasim.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

orc_ctx *orc_create(const pic_config *cfg);
void orc_destroy(orc_ctx *o);
int orc_threads(int n); /* OpenMP threads for the oracle (0 = query) */

int orc_upload_field(orc_ctx *o, int which, const double *host);
int orc_download_field(orc_ctx *o, int which, double *host);
int orc_upload_particles(orc_ctx *o, int sp, int64_t n, const double *x,
                         const double *y, const double *z, const double *vx,
                         const double *vy, const double *vz);
int orc_download_particles(orc_ctx *o, int sp, int64_t cap, int64_t *n,
                           double *x, double *y, double *z, double *vx,
                           double *vy, double *vz);
int64_t orc_particle_count(orc_ctx *o, int sp);

int orc_load_particles(orc_ctx *o);
int orc_renew_grid(orc_ctx *o);
int orc_renew_b_half(orc_ctx *o);
int orc_halo_b(orc_ctx *o);
int orc_halo_e(orc_ctx *o);
int orc_push_deposit_j(orc_ctx *o, int sp);
int orc_drift_deposit_n(orc_ctx *o, int sp);
int orc_fold_current(orc_ctx *o);
int orc_renew_e(orc_ctx *o);
int orc_particle_boundary(orc_ctx *o);
int orc_inflow(orc_ctx *o);
int orc_sort(orc_ctx *o, int sp);
int orc_diagnose(orc_ctx *o, pic_diag *out);
int orc_begin(orc_ctx *o);
int orc_step(orc_ctx *o, int nsteps);

/* building blocks exposed for known-answer tests */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);
/* gather at an arbitrary position (shape_func_v2): out = ex,ey,ez,bx,by,bz */
void orc_gather(orc_ctx *o, double x, double y, double z, double out[6]);
/* one cal_track4 step on a single particle with given fields */
void orc_boris(const pic_config *cfg, double m, const double f[6], double p[6]);
/* sigma on the 2x lattice */
double orc_sigma(orc_ctx *o, int k2, int l2, int m2);

/* FP64 discrete adjoint (pic_oracle_adj.c).  p = (Mx, My, Mz, V).  The
 * forward is load, inflow and nsteps steps without the sort; the loss is
 * the step average of the Maxwell-stress z flux through a sphere of R
 * cells (nq quadrature points) around the dipole. */
double orc_adj_loss(orc_ctx *o, const double p[4], int nsteps, double R, int nq);
int orc_adj_gradient(orc_ctx *o, const double p[4], int nsteps, double R,
                     int nq, double *loss, double grad[4]);
/* stress flux through the sphere for the current fields */
double orc_adj_force(orc_ctx *o, double R, int nq);
/* dot-product checks of the transposed field operators */
int orc_adj_selftest(orc_ctx *o, uint64_t seed, double out[4]);

/* FP64-form checkpoints in the format of include/plasmasim.h
 * (pic_checkpoint_save / pic_checkpoint_load); pic_oracle_ckpt.c */
int orc_checkpoint_save(orc_ctx *o, const char *path);
int orc_checkpoint_load(orc_ctx *o, const char *path);

/* output() / output_tecplot() text dumps (format PIC_OUTPUT_*);
 * pic_oracle_out.c */
int orc_output_fields(orc_ctx *o, const char *path, int format, int64_t step);

/* renew_phi3 + sr_phi + E correction + sr_e; pic_oracle_phi3.c */
int orc_poisson(orc_ctx *o);
int orc_adj_boris_check(const pic_config *cfg, double m, const double f[6],
                        const double v[3], const double lvp[3], double out[2]);

#ifdef __cplusplus
}
#endif
#endif

/*
 * pic_run.c - host driver of the 3D EM-PIC step, in C over the C-ABI.
 *
 * Restates the reference's main_func()/fdtd() (ashida-lab/PLASMAsim
 * PIC/3d_pic458_domain.c :256-289, :627-952) on one subdomain: set up the
 * grid and the static dipole (init_grid), load the plasma
 * (init_particle_flat), the pre-loop inflow (:673-674), then the step loop
 * calling one library routine per reference routine in the reference's
 * order, with the sort and inflow cadences and a periodic diagnostics line
 * (thrust_f, every 500 steps in the reference; here every --diag steps).
 *
 * Usage: pic_run [--preset N] [--steps K] [--fp64|--fp32] [--det]
 *                [--diag D] [--fused] [--suspend FILE] [--suspend-every S]
 *                [--resume FILE] [--output-every N] [--tecplot]
 *                [--thrust-every N] [--thrust-ave] [--version V]
 *                [--poisson-every P]
 * --fused calls pic_step() (loops A and B fused into one particle pass)
 * instead of the routine-by-routine loop.  --poisson-every P enables the 3D
 * Poisson correction (renew_phi3, compiled out in the reference) every P
 * steps.
 *
 * Output (fdtd :913-927): --output-every N writes the field dump of
 * output() (:4834-4883; --tecplot: output_tecplot()) to
 * fdtd<V>-0-0-<c>.dat at steps c with c % N == 0; --thrust-every N appends
 * thrust_f's line (:5110-5119: c, F[0..26], E[0..5]) to thrust<V>.txt at
 * those steps; --thrust-ave adds thrust_fave (:5124-5285): F averaged over
 * every step of each N-step window, written to thrust<V>_ave.txt when
 * c % N == 0.  <V> is --version (the reference's version number).
 *
 * Suspend and resume (main_func :270-283, fdtd :918-931): with --resume
 * FILE the run continues from FILE when it exists (instead of loading the
 * plasma), up to --steps steps in total; with --suspend FILE the state is
 * written every --suspend-every steps and at the end.  As in the
 * reference, a file named "finish" in the working directory stops the run
 * at the next diagnostic step (after writing the suspend file).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/plasmasim.h"

#define CHECK(ctx, call)                                                   \
  do {                                                                     \
    int rc_ = (call);                                                      \
    if (rc_) {                                                             \
      fprintf(stderr, "%s failed (%d): %s\n", #call, rc_,                  \
              (ctx) ? pic_last_error(ctx) : "");                           \
      if (ctx) pic_destroy(ctx);                                           \
      return 1;                                                            \
    }                                                                      \
  } while (0)

static double now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* one reference step: fdtd() body :687-949 */
static int step_routines(pic_ctx *ctx, const pic_config *c, int64_t step,
                         int *c_sort, int *c_add) {
  int rc;
  if ((rc = pic_renew_grid(ctx))) return rc;                    /* :688 */
  if (*c_sort > c->sort_every - 1) {                            /* :694 */
    *c_sort = 0;
    if ((rc = pic_sort(ctx, PIC_ION))) return rc;
    if ((rc = pic_sort(ctx, PIC_ELECTRON))) return rc;
  }
  if (*c_add > c->inflow_every - 1) {                           /* :700 */
    if ((rc = pic_inflow(ctx))) return rc;
    *c_add = 0;
  }
  if ((rc = pic_renew_b_half(ctx))) return rc;                  /* :727 */
  if ((rc = pic_halo_b(ctx))) return rc;                        /* :729 */
  if ((rc = pic_push_deposit_j(ctx, PIC_ION))) return rc;       /* :731-812 */
  if ((rc = pic_push_deposit_j(ctx, PIC_ELECTRON))) return rc;
  if ((rc = pic_renew_b_half(ctx))) return rc;                  /* :814 */
  if ((rc = pic_halo_b(ctx))) return rc;                        /* :816 */
  if ((rc = pic_drift_deposit_n(ctx, PIC_ION))) return rc;      /* :818-867 */
  if ((rc = pic_drift_deposit_n(ctx, PIC_ELECTRON))) return rc;
  if ((rc = pic_fold_current(ctx))) return rc;                  /* :893 */
  if ((rc = pic_renew_e(ctx))) return rc;                       /* :895 */
  if ((rc = pic_halo_e(ctx))) return rc;                        /* :899 */
  if (c->poisson_every > 0 && step % c->poisson_every == 0)      /* PHI(), :901 */
    if ((rc = pic_poisson(ctx))) return rc;  /* renew_phi3, sr_phi, sr_e */
  if ((rc = pic_particle_boundary(ctx))) return rc;             /* :909-911 */
  (*c_sort)++;
  (*c_add)++;
  return 0;
}

/* thrust_f's line (:5110-5119): F[15..26] and E[4..5] belong to particles
 * with flag < 0, which the live code never creates (flag is 1, :510,
 * :6058), so they are 0 */
static int thrust_line(const char *path, int64_t c, const double F[27],
                       const double E[6], int with_e) {
  FILE *f = fopen(path, "a");
  if (!f) return PIC_E_IO;
  int ok = fprintf(f, "%lld", (long long)c) > 0;
  for (int i = 0; i < 27; ++i) ok = ok && fprintf(f, " %.3E", F[i]) > 0;
  if (with_e)
    for (int i = 0; i < 6; ++i) ok = ok && fprintf(f, " %.8E", E[i]) > 0;
  ok = ok && fprintf(f, "\n") > 0;
  if (fclose(f) != 0) ok = 0;
  return ok ? 0 : PIC_E_IO;
}

static void forces(const pic_diag *d, double F[27]) {
  for (int i = 0; i < 27; ++i) F[i] = i < 15 ? d->force[i] : 0.0;
}

/* suspend_job (:7629-7675): the counters of the routine loop go with the
 * state */
static int suspend(pic_ctx *ctx, const char *path, int fused, int64_t c,
                   int c_add, int c_sort) {
  int rc;
  if (!fused && (rc = pic_set_counters(ctx, c, c_add, c_sort))) return rc;
  return pic_checkpoint_save(ctx, path);
}

int main(int argc, char **argv) {
  int preset = 1, steps = 100, diag = 50, fused = 0;
  int precision = -1, det = 0, suspend_every = 0;
  int output_every = 0, tecplot = 0, thrust_every = 0, thrust_ave = 0, version = 0;
  int poisson_every = -1;
  const char *suspend_path = NULL, *resume_path = NULL;
  for (int i = 1; i < argc; ++i) {
    if (!strcmp(argv[i], "--preset") && i + 1 < argc) preset = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--steps") && i + 1 < argc) steps = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--diag") && i + 1 < argc) diag = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--fp64")) precision = PIC_FP64;
    else if (!strcmp(argv[i], "--fp32")) precision = PIC_FP32;
    else if (!strcmp(argv[i], "--det")) det = 1;
    else if (!strcmp(argv[i], "--fused")) fused = 1;
    else if (!strcmp(argv[i], "--suspend") && i + 1 < argc) suspend_path = argv[++i];
    else if (!strcmp(argv[i], "--suspend-every") && i + 1 < argc)
      suspend_every = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--resume") && i + 1 < argc) resume_path = argv[++i];
    else if (!strcmp(argv[i], "--output-every") && i + 1 < argc)
      output_every = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--tecplot")) tecplot = 1;
    else if (!strcmp(argv[i], "--thrust-every") && i + 1 < argc)
      thrust_every = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--thrust-ave")) thrust_ave = 1;
    else if (!strcmp(argv[i], "--version") && i + 1 < argc) version = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--poisson-every") && i + 1 < argc)
      poisson_every = atoi(argv[++i]);
    else {
      fprintf(stderr,
              "usage: %s [--preset N] [--steps K] [--fp64|--fp32] [--det] "
              "[--diag D] [--fused] [--suspend FILE] [--suspend-every S] "
              "[--resume FILE] [--output-every N] [--tecplot] "
              "[--thrust-every N] [--thrust-ave] [--version V] "
              "[--poisson-every P]\n", argv[0]);
      return 2;
    }
  }
  pic_config cfg;
  pic_ctx *ctx = NULL;
  CHECK(ctx, pic_config_preset(preset, &cfg));
  if (precision >= 0) cfg.precision = precision;
  if (det) cfg.deposit = PIC_DEPOSIT_DETERMINISTIC;
  if (poisson_every >= 0) cfg.poisson_every = poisson_every;
  CHECK(ctx, pic_create(&cfg, &ctx));                 /* init_grid :291 */
  int c_sort = 0, c_add = 0;
  int64_t c0 = 0;
  FILE *probe = resume_path ? fopen(resume_path, "rb") : NULL;
  if (probe) {                                        /* continue_job :282 */
    fclose(probe);
    CHECK(ctx, pic_checkpoint_load(ctx, resume_path));
    CHECK(ctx, pic_get_counters(ctx, &c0, &c_add, &c_sort));
    fprintf(stderr, "continue_job from %s at step %lld\n", resume_path,
            (long long)c0);
  } else {
    CHECK(ctx, pic_load_particles(ctx));              /* :412-625 */
    CHECK(ctx, pic_begin(ctx));                       /* :673-674 */
  }
  fprintf(stderr, "grid %dx%dx%d ppc %d form %s deposit %s\n", cfg.nx, cfg.ny,
          cfg.nz, cfg.ppc, cfg.precision == PIC_FP64 ? "fp64" : "fp32",
          cfg.deposit == PIC_DEPOSIT_DETERMINISTIC ? "deterministic" : "atomic");
  const double t0 = now();
  double Fave[27] = {0};
  char thrust_path[64], ave_path[64];
  snprintf(thrust_path, sizeof thrust_path, "thrust%d.txt", version);
  snprintf(ave_path, sizeof ave_path, "thrust%d_ave.txt", version);
  if (fused && thrust_ave && thrust_every > 0) CHECK(ctx, pic_thrust_ave_enable(ctx, 1));
  for (int64_t c = c0; c < steps; ++c) {
    if (fused) {
      CHECK(ctx, pic_step(ctx, 1));
    } else {
      CHECK(ctx, step_routines(ctx, &cfg, c, &c_sort, &c_add));
    }
    if (output_every > 0 && c % output_every == 0) {   /* output :913-915 */
      char name[96];
      snprintf(name, sizeof name, "fdtd%d-0-0-%lld.dat", version, (long long)c);
      CHECK(ctx, pic_output_fields(ctx, name, tecplot ? PIC_OUTPUT_TECPLOT
                                                      : PIC_OUTPUT_COLUMNS, c));
    }
    /* the fused step sums thrust_fave's force itself (no second gather) */
    const int ave_fused = fused && thrust_ave;
    if (thrust_every > 0 && (c % thrust_every == 0 || (thrust_ave && !ave_fused))) {
      pic_diag d;
      double F[27];
      CHECK(ctx, pic_diagnose(ctx, &d));
      forces(&d, F);
      if (c % thrust_every == 0)                        /* thrust_f :917   */
        CHECK(ctx, thrust_line(thrust_path, c, F, d.energy_box, 1));
      if (thrust_ave && !ave_fused) {                   /* thrust_fave     */
        for (int i = 0; i < 27; ++i) Fave[i] += F[i] / thrust_every;
        if (c % thrust_every == 0) {
          CHECK(ctx, thrust_line(ave_path, c, Fave, NULL, 0));
          for (int i = 0; i < 27; ++i) Fave[i] = 0.;
        }
      }
    }
    if (thrust_every > 0 && ave_fused && c % thrust_every == 0) {
      double F[27];
      int64_t nsum = 0;
      CHECK(ctx, pic_thrust_ave_get(ctx, F, &nsum, 1)); /* thrust_fave :5265-5281 */
      for (int i = 0; i < 27; ++i) Fave[i] = F[i] / thrust_every;
      CHECK(ctx, thrust_line(ave_path, c, Fave, NULL, 0));
    }
    const int last = c + 1 == steps;
    if (suspend_path && ((suspend_every > 0 && (c + 1) % suspend_every == 0) || last))
      CHECK(ctx, suspend(ctx, suspend_path, fused, c + 1, c_add, c_sort));
    if (diag > 0 && ((c + 1) % diag == 0 || last)) {
      pic_diag d;
      CHECK(ctx, pic_diagnose(ctx, &d));
      const double W = d.field_energy_e + d.field_energy_b1 + d.kinetic[0] +
                       d.kinetic[1];
      printf("%lld %.3f s  Ee %.6e  Eb1 %.6e  Ki %.6e  Ke %.6e  total %.6e  "
             "Fx %.4e Fy %.4e Fz %.4e  gauss_l2 %.4e  n %lld %lld\n",
             (long long)(c + 1), now() - t0, d.field_energy_e, d.field_energy_b1,
             d.kinetic[0], d.kinetic[1], W, d.force[0], d.force[1],
             d.force[2], d.gauss_l2, (long long)d.count[0],
             (long long)d.count[1]);
      fflush(stdout);
      FILE *fin = fopen("finish", "r");               /* :918-923 */
      if (fin) {
        fclose(fin);
        if (suspend_path && !last)
          CHECK(ctx, suspend(ctx, suspend_path, fused, c + 1, c_add, c_sort));
        fprintf(stderr, "finish file: stopping after step %lld\n",
                (long long)(c + 1));
        break;
      }
    }
  }
  pic_destroy(ctx);
  return 0;
}

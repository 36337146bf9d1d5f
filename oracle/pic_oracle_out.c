/*
 * pic_oracle_out.c - the oracle's field dumps.  TEST INFRASTRUCTURE ONLY.
 *
 * Restates output() (PIC/3d_pic458_domain.c :4834-4883) and
 * output_tecplot() (:4885-4938) for one rank (myrank 0, one process per
 * axis) on the oracle's Grid records; phi is printed as stored.
 */
#include <math.h>
#include <stdio.h>

#include "pic_oracle_internal.h"

int orc_output_fields(orc_ctx *o, const char *path, int format, int64_t step) {
  const pic_config *c = &o->cfg;
  const int Nx = c->nx, Ny = c->ny, Nz = c->nz;
  const double dx = c->dx;
  FILE *fp = fopen(path, "w");
  if (!fp) return PIC_E_IO;
  if (format == PIC_OUTPUT_COLUMNS) {
    for (int k = 2; k < Nx + 2; k++)
      for (int l = 2; l < Ny + 2; l++)
        for (int m = 2; m < Nz + 2; m++) {
          const Grid *g = &G(o, k, l, m);
          const Grid *gk = &G(o, k + 1, l, m), *gl = &G(o, k, l + 1, m),
                     *gm = &G(o, k, l, m + 1);
          const Grid *glm = &G(o, k, l + 1, m + 1), *gkm = &G(o, k + 1, l, m + 1),
                     *gkl = &G(o, k + 1, l + 1, m);
          fprintf(fp,
                  "%.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E "
                  "%.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E\n",
                  (k - 2 - Nx / 2) * dx, (l - 2 - Ny / 2) * dx,
                  (m - (Nz + 4) / 2) * dx, g->ni, g->ne, g->ni - g->ne,
                  (gk->ex + g->ex) / 2., (gl->ey + g->ey) / 2.,
                  (gm->ez + g->ez) / 2.,
                  (g->bx + gl->bx + gm->bx + glm->bx) / 4.,
                  (g->by + gk->by + gm->by + gkm->by) / 4.,
                  (g->bz + gk->bz + gl->bz + gkl->bz) / 4., g->jix0, g->jiy0,
                  g->jiz0, g->jex0, g->jey0, g->jez0, g->phi,
                  sqrt(pow((g->bx + gl->bx + gm->bx + glm->bx) / 4., 2) +
                       pow((g->by + gk->by + gm->by + gkm->by) / 4., 2) +
                       pow((g->bz + gk->bz + gl->bz + gkl->bz) / 4., 2)),
                  (g->bx + gl->bx + gm->bx + glm->bx) / 4. -
                      (g->b0x + gl->b0x + gm->b0x + glm->b0x) / 4.,
                  (g->by + gk->by + gm->by + gkm->by) / 4. -
                      (g->b0y + gk->b0y + gm->b0y + gkm->b0y) / 4.,
                  (g->bz + gk->bz + gl->bz + gkl->bz) / 4. -
                      (g->b0z + gk->b0z + gl->b0z + gkl->b0z) / 4.);
        }
  } else if (format == PIC_OUTPUT_TECPLOT) {
    fprintf(fp,
            "VARIABLES = \"X[m]\" \"Y[m]\" \"Z[m]\" \"ni[/m3]\" \"ne[/m3]\" "
            "\"dens[/m3]\" \"Ex[V/m]\" \"Ey[V/m]\" \"Ez[V/m]\" \"Bx[T]\" "
            "\"By[T]\" \"Bz[T]\" \"jix[A/m2]\" \"jiy[A/m2]\" \"jiz[A/m2]\" "
            "\"jex[A/m2]\" \"jey[A/m2]\" \"jez[A/m2]\" \"j[A/m2]\" \"phi[V]\" "
            "\"B[T]\" \"Bpx[T]\" \"Bpy[T]\" \"Bpz[T]\"\n");
    fprintf(fp,
            "ZONE T=\"STP:%lld\", STRANDID=1, SOLUTIONTIME=%lld, I=%d, J=%d, "
            "K=%d\n",
            (long long)step, (long long)step, Nz, Ny + 1, Nx + 1);
    for (int k = 2; k < Nx + 3; k++) {
      for (int l = 2; l < Ny + 3; l++) {
        for (int m = 2; m < Nz + 2; m++) {
          const Grid *g = &G(o, k, l, m);
          const Grid *gk = &G(o, k + 1, l, m), *gl = &G(o, k, l + 1, m),
                     *gm = &G(o, k, l, m + 1);
          const Grid *glm = &G(o, k, l + 1, m + 1), *gkm = &G(o, k + 1, l, m + 1),
                     *gkl = &G(o, k + 1, l + 1, m);
          fprintf(fp,
                  "%.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E "
                  "%.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E %.3E\n",
                  (k - 2 - Nx / 2) * dx, (l - 2 - Ny / 2) * dx,
                  (m - (Nz + 4) / 2) * dx, g->ni, g->ne, g->ni - g->ne,
                  (gk->ex + g->ex) / 2., (gl->ey + g->ey) / 2.,
                  (gm->ez + g->ez) / 2.,
                  (g->bx + gl->bx + gm->bx + glm->bx) / 4.,
                  (g->by + gk->by + gm->by + gkm->by) / 4.,
                  (g->bz + gk->bz + gl->bz + gkl->bz) / 4.,
                  (gk->jix0 + g->jix0) / 2., (gl->jiy0 + g->jiy0) / 2.,
                  (gm->jiz0 + g->jiz0) / 2., (gk->jex0 + g->jex0) / 2.,
                  (gl->jey0 + g->jey0) / 2., (gm->jez0 + g->jez0) / 2.,
                  sqrt(pow((gk->jix0 + g->jix0) / 2. - (gk->jex0 + g->jex0) / 2., 2) +
                       pow((gl->jiy0 + g->jiy0) / 2. - (gl->jey0 + g->jey0) / 2., 2) +
                       pow((gm->jiz0 + g->jiz0) / 2. - (gm->jez0 + g->jez0) / 2., 2)),
                  g->phi,
                  sqrt(pow((g->bx + gl->bx + gm->bx + glm->bx) / 4., 2) +
                       pow((g->by + gk->by + gm->by + gkm->by) / 4., 2) +
                       pow((g->bz + gk->bz + gl->bz + gkl->bz) / 4., 2)),
                  (g->bx + gl->bx + gm->bx + glm->bx) / 4. -
                      (g->b0x + gl->b0x + gm->b0x + glm->b0x) / 4.,
                  (g->by + gk->by + gm->by + gkm->by) / 4. -
                      (g->b0y + gk->b0y + gm->b0y + gkm->b0y) / 4.,
                  (g->bz + gk->bz + gl->bz + gkl->bz) / 4. -
                      (g->b0z + gk->b0z + gl->b0z + gkl->b0z) / 4.);
        }
        fprintf(fp, "\n");
      }
    }
  } else {
    fclose(fp);
    return PIC_E_ARG;
  }
  return fclose(fp) == 0 ? 0 : PIC_E_IO;
}

/*
 * pic_oracle_phi3.c - the oracle's 3D Poisson correction.  TEST
 * INFRASTRUCTURE ONLY.
 *
 * Restates renew_phi3 (PIC/3d_pic458_domain.c :8315-8665) for one rank:
 * the source (:8336-8353), Numerical Recipes' sinft (:8182-8215) along z,
 * y and x as its defining sum S(k) = sum_{j=1}^{n-1} y_j sin(pi j k / n)
 * (sinft sets y_0 = 0), the division by the symbol (:8500-8512), the
 * inverse transforms 2 S / n (:8086-8180), then sr_phi's x and y guard
 * copies and the E correction (:1389-1573) and sr_e.  The library computes
 * the same transforms with FFTs; the sums here are independent of that.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "pic_oracle_internal.h"

static const double Pi = 3.14159265358979323846;

/* one sinft line: out[k] = scale * sum_{j=1}^{n-1} y[j*st] sin(pi j k / n) */
static void dst_line(double *y, long st, int n, const double *sn, double *tmp,
                     int inverse) {
  for (int k = 0; k < n; k++) {
    double s = 0.;
    for (int j = 1; j < n; j++) s += y[(long)j * st] * sn[(size_t)j * n + k];
    tmp[k] = inverse ? 2. * s / n : s;
  }
  for (int k = 0; k < n; k++) y[(long)k * st] = tmp[k];
}

static double *sin_table(int n) {
  double *t = (double *)malloc(sizeof(double) * (size_t)n * n);
  if (!t) return NULL;
  for (int j = 0; j < n; j++)
    for (int k = 0; k < n; k++) t[(size_t)j * n + k] = sin(Pi * (double)j * k / n);
  return t;
}

/* transform every line of rho[NX][NY][NZ] along one axis */
static void dst_axis(double *rho, int NX, int NY, int NZ, int axis, const double *sn,
                     double *tmp, int inverse) {
  if (axis == 2) {
    for (int k = 0; k < NX; k++)
      for (int l = 0; l < NY; l++)
        dst_line(rho + ((size_t)k * NY + l) * NZ, 1, NZ, sn, tmp, inverse);
  } else if (axis == 1) {
    for (int k = 0; k < NX; k++)
      for (int m = 0; m < NZ; m++)
        dst_line(rho + (size_t)k * NY * NZ + m, NZ, NY, sn, tmp, inverse);
  } else {
    for (int l = 0; l < NY; l++)
      for (int m = 0; m < NZ; m++)
        dst_line(rho + (size_t)l * NZ + m, (long)NY * NZ, NX, sn, tmp, inverse);
  }
}

int orc_poisson(orc_ctx *o) {
  const pic_config *c = &o->cfg;
  const int NX = o->NX, NY = o->NY, NZ = o->NZ, A3 = c->phi_margin;
  const double q = c->q, e0 = c->eps0, dx = c->dx;
  const size_t n = (size_t)NX * NY * NZ;
  double *rho = (double *)calloc(n, sizeof(double));
  double *tmp = (double *)malloc(sizeof(double) * (size_t)(NX + NY + NZ));
  double *sn[3] = {sin_table(NX), sin_table(NY), sin_table(NZ)};
  if (!rho || !tmp || !sn[0] || !sn[1] || !sn[2]) {
    free(rho);
    free(tmp);
    for (int a = 0; a < 3; a++) free(sn[a]);
    return PIC_E_NOMEM;
  }
  /* source (:8336-8353) */
  for (int k = 2; k < NX + 2; k++)
    for (int l = 2; l < NY + 2; l++)
      for (int m = A3 + 2; m < NZ + 2 - A3; m++)
        if (k - 2 > A3 && k - 2 < NX - A3 && l - 2 > A3 && l - 2 < NY - A3)
          rho[((size_t)(k - 2) * NY + (l - 2)) * NZ + (m - 2)] =
              -(-q * (G(o, k, l, m).ni - G(o, k, l, m).ne) / e0 +
                (-G(o, k + 2, l, m).ex + 27 * G(o, k + 1, l, m).ex -
                 27 * G(o, k, l, m).ex + G(o, k - 1, l, m).ex - G(o, k, l + 2, m).ey +
                 27 * G(o, k, l + 1, m).ey - 27 * G(o, k, l, m).ey +
                 G(o, k, l - 1, m).ey - G(o, k, l, m + 2).ez +
                 27 * G(o, k, l, m + 1).ez - 27 * G(o, k, l, m).ez +
                 G(o, k, l, m - 1).ez) /
                    (24. * dx));
  /* fft_z, fft_y, fft_x */
  dst_axis(rho, NX, NY, NZ, 2, sn[2], tmp, 0);
  dst_axis(rho, NX, NY, NZ, 1, sn[1], tmp, 0);
  dst_axis(rho, NX, NY, NZ, 0, sn[0], tmp, 0);
  /* division (:8500-8512): k runs over y, l over z, m over x */
  for (int kx = 0; kx < NX; kx++)
    for (int ly = 0; ly < NY; ly++)
      for (int mz = 0; mz < NZ; mz++) {
        const double A = -(pow(sin(Pi * (double)ly / (double)NY / 2.), 2) +
                           pow(sin(Pi * (double)mz / (double)NZ / 2.), 2) +
                           pow(sin(Pi * (double)kx / (double)NX / 2.), 2)) /
                         pow(dx / 2., 2);
        double *r = &rho[((size_t)kx * NY + ly) * NZ + mz];
        *r = A == 0. ? 0. : *r / A;
      }
  /* inv_fft_x, inv_fft_y, inv_fft_z */
  dst_axis(rho, NX, NY, NZ, 0, sn[0], tmp, 1);
  dst_axis(rho, NX, NY, NZ, 1, sn[1], tmp, 1);
  dst_axis(rho, NX, NY, NZ, 2, sn[2], tmp, 1);
  for (int k = 0; k < NX; k++)
    for (int l = 0; l < NY; l++)
      for (int m = 0; m < NZ; m++)
        G(o, k + 2, l + 2, m + 2).phi = rho[((size_t)k * NY + l) * NZ + m];
  /* sr_phi: x then y guard layers from the periodic images (:1403-1560) */
  for (int l = 0; l < o->SY; l++)
    for (int m = 0; m < o->SZ; m++) {
      G(o, 0, l, m).phi = G(o, NX, l, m).phi;
      G(o, 1, l, m).phi = G(o, NX + 1, l, m).phi;
      G(o, NX + 2, l, m).phi = G(o, 2, l, m).phi;
      G(o, NX + 3, l, m).phi = G(o, 3, l, m).phi;
    }
  for (int k = 0; k < o->SX; k++)
    for (int m = 0; m < o->SZ; m++) {
      G(o, k, 0, m).phi = G(o, k, NY, m).phi;
      G(o, k, 1, m).phi = G(o, k, NY + 1, m).phi;
      G(o, k, NY + 2, m).phi = G(o, k, 2, m).phi;
      G(o, k, NY + 3, m).phi = G(o, k, 3, m).phi;
    }
  /* E correction (:1563-1571) */
  for (int k = 2; k < NX + 2; k++)
    for (int l = 2; l < NY + 2; l++)
      for (int m = 2; m < NZ + 2; m++) {
        Grid *g = &G(o, k, l, m);
        g->ex += (-G(o, k + 1, l, m).phi + 27 * g->phi - 27 * G(o, k - 1, l, m).phi +
                  G(o, k - 2, l, m).phi) /
                 (24. * dx);
        g->ey += (-G(o, k, l + 1, m).phi + 27 * g->phi - 27 * G(o, k, l - 1, m).phi +
                  G(o, k, l - 2, m).phi) /
                 (24. * dx);
        g->ez += (-G(o, k, l, m + 1).phi + 27 * g->phi - 27 * G(o, k, l, m - 1).phi +
                  G(o, k, l, m - 2).phi) /
                 (24. * dx);
      }
  free(rho);
  free(tmp);
  for (int a = 0; a < 3; a++) free(sn[a]);
  return orc_halo_e(o); /* sr_e */
}

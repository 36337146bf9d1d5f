/*
 * pic_oracle_ckpt.c - the oracle's reader and writer of FP64-form
 * checkpoints.  TEST INFRASTRUCTURE ONLY.
 *
 * An independent implementation of the checkpoint format specified in
 * include/plasmasim.h (pic_checkpoint_header and the payload layout), which
 * replaces the reference's suspend_job / continue_job (PIC/3d_pic458_domain.c
 * :7629-7911): the oracle resumes checkpoints written by the library and
 * the library resumes the oracle's, so the two implementations check each
 * other.  The oracle keeps no tile ranges: it writes ntiles = 0 and skips
 * stored tile starts; it drops the library's dead slots on reading.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pic_oracle_internal.h"

static const char kMagic[8] = {'P', 'L', 'A', 'S', 'C', 'K', 'P', 'T'};

typedef struct {
  uint64_t h[4];
  uint64_t nwords;
  unsigned char tail[8];
  int ntail;
} Hash4;

static void h_init(Hash4 *s) {
  for (int j = 0; j < 4; j++) s->h[j] = 0xcbf29ce484222325ull + (uint64_t)j;
  s->nwords = 0;
  s->ntail = 0;
}

static void h_word(Hash4 *s, uint64_t w) {
  uint64_t *x = &s->h[s->nwords & 3];
  *x = (*x ^ w) * 0x100000001b3ull;
  s->nwords++;
}

static void h_bytes(Hash4 *s, const void *p, size_t n) {
  const unsigned char *c = (const unsigned char *)p;
  for (size_t i = 0; i < n; i++) {
    s->tail[s->ntail++] = c[i];
    if (s->ntail == 8) {
      uint64_t w;
      memcpy(&w, s->tail, 8);
      h_word(s, w);
      s->ntail = 0;
    }
  }
}

static uint64_t h_finish(Hash4 *s) {
  if (s->ntail) {
    memset(s->tail + s->ntail, 0, (size_t)(8 - s->ntail));
    uint64_t w;
    memcpy(&w, s->tail, 8);
    h_word(s, w);
    s->ntail = 0;
  }
  return ((s->h[0] * 31 + s->h[1]) * 31 + s->h[2]) * 31 + s->h[3];
}

typedef struct {
  FILE *f;
  Hash4 hash;
  uint64_t bytes;
  int ok;
} Io;

static void io_put(Io *io, const void *p, size_t n) {
  if (!io->ok) return;
  h_bytes(&io->hash, p, n);
  io->ok = fwrite(p, 1, n, io->f) == n;
  io->bytes += n;
}

static void io_get(Io *io, void *p, size_t n) {
  if (!io->ok) return;
  io->ok = fread(p, 1, n, io->f) == n;
  if (io->ok) h_bytes(&io->hash, p, n);
  io->bytes += n;
}

/* the fields two configs must share (capacity, deposit mode and tiles may
 * differ), field by field so padding does not take part */
static int same_physics(const pic_config *a, const pic_config *b) {
  int ok = a->nx == b->nx && a->ny == b->ny && a->nz == b->nz &&
           a->ppc == b->ppc && a->dx == b->dx && a->dt == b->dt &&
           a->c == b->c && a->eps0 == b->eps0 && a->mu0 == b->mu0 &&
           a->q == b->q && a->dipole_m == b->dipole_m &&
           a->dipole_alpha == b->dipole_alpha &&
           a->dipole_soft == b->dipole_soft &&
           a->dipole_rmax == b->dipole_rmax && a->sigma_amp == b->sigma_amp &&
           a->inflow_every == b->inflow_every &&
           a->sort_every == b->sort_every && a->refill_far == b->refill_far &&
           a->precision == b->precision && a->seed == b->seed &&
           a->poisson_every == b->poisson_every && a->phi_margin == b->phi_margin;
  for (int s = 0; s < 2; s++)
    ok = ok && a->mass[s] == b->mass[s] && a->weight[s] == b->weight[s] &&
         a->vth[s] == b->vth[s] && a->periodic[s] == b->periodic[s];
  for (int d = 0; d < 3; d++)
    ok = ok && a->vdrift[d] == b->vdrift[d] && a->imf[d] == b->imf[d] &&
         a->absorb[d] == b->absorb[d] &&
         a->ballistic_lo[d] == b->ballistic_lo[d] &&
         a->ballistic_hi[d] == b->ballistic_hi[d];
  return ok;
}

int orc_checkpoint_save(orc_ctx *o, const char *path) {
  const size_t nc = (size_t)o->SX * o->SY * o->SZ;
  pic_checkpoint_header h;
  memset(&h, 0, sizeof h);
  memcpy(h.magic, kMagic, 8);
  h.version = 1;
  h.header_bytes = (uint32_t)sizeof h;
  h.config_bytes = (uint32_t)sizeof(pic_config);
  h.precision = PIC_FP64;
  h.nx = o->cfg.nx;
  h.ny = o->cfg.ny;
  h.nz = o->cfg.nz;
  h.step = o->step;
  h.c_add = o->c_add;
  h.c_sort = o->c_sort;
  h.inflow_event = o->inflow_event;
  h.ntiles = 0;
  for (int s = 0; s < 2; s++) h.n[s] = o->sp[s].n;

  Io io;
  io.f = fopen(path, "wb");
  if (!io.f) return PIC_E_IO;
  h_init(&io.hash);
  io.bytes = 0;
  io.ok = fwrite(&h, 1, sizeof h, io.f) == sizeof h;
  pic_config cfg = o->cfg;
  cfg.precision = PIC_FP64;
  io_put(&io, &cfg, sizeof cfg);
  double *plane = (double *)malloc(nc * sizeof(double));
  if (!plane) {
    fclose(io.f);
    return PIC_E_NOMEM;
  }
  for (int w = 0; w < PIC_NFIELDS && io.ok; w++) {
    orc_download_field(o, w, plane);
    io_put(&io, plane, nc * sizeof(double));
  }
  free(plane);
  for (int s = 0; s < 2 && io.ok; s++) {
    const PList *L = &o->sp[s];
    for (int a = 0; a < 6; a++)
      for (int64_t i = 0; i < L->n; i++) {
        const Particle *p = &L->p[i];
        const double v = a == 0 ? p->x : a == 1 ? p->y : a == 2 ? p->z
                         : a == 3 ? p->vx : a == 4 ? p->vy : p->vz;
        io_put(&io, &v, sizeof v);
      }
  }
  h.payload_bytes = io.bytes;
  h.checksum = h_finish(&io.hash);
  if (io.ok) io.ok = fseek(io.f, 0, SEEK_SET) == 0 &&
                     fwrite(&h, 1, sizeof h, io.f) == sizeof h;
  if (fclose(io.f) != 0) io.ok = 0;
  return io.ok ? 0 : PIC_E_IO;
}

int orc_checkpoint_load(orc_ctx *o, const char *path) {
  const size_t nc = (size_t)o->SX * o->SY * o->SZ;
  Io io;
  io.f = fopen(path, "rb");
  if (!io.f) return PIC_E_IO;
  h_init(&io.hash);
  io.bytes = 0;
  io.ok = 1;
  pic_checkpoint_header h;
  int rc = 0;
  double *plane = NULL;
  double *arr[6] = {NULL, NULL, NULL, NULL, NULL, NULL};
  if (fread(&h, 1, sizeof h, io.f) != sizeof h || memcmp(h.magic, kMagic, 8) ||
      h.version != 1 || h.header_bytes != sizeof h ||
      h.config_bytes != sizeof(pic_config) || h.precision != PIC_FP64) {
    rc = PIC_E_FORMAT;
    goto done;
  }
  pic_config fc;
  io_get(&io, &fc, sizeof fc);
  pic_config mine = o->cfg;
  mine.precision = PIC_FP64;
  if (!io.ok || !same_physics(&fc, &mine)) {
    rc = PIC_E_FORMAT;
    goto done;
  }
  plane = (double *)malloc(nc * sizeof(double));
  if (!plane) {
    rc = PIC_E_NOMEM;
    goto done;
  }
  for (int w = 0; w < PIC_NFIELDS && io.ok; w++) {
    io_get(&io, plane, nc * sizeof(double));
    if (io.ok) orc_upload_field(o, w, plane);
  }
  for (int s = 0; s < 2 && io.ok; s++) {
    const int64_t n = h.n[s];
    for (int a = 0; a < 6; a++) {
      arr[a] = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
      if (!arr[a]) {
        rc = PIC_E_NOMEM;
        goto done;
      }
      io_get(&io, arr[a], (size_t)n * sizeof(double));
    }
    /* the library keeps dead slots (z = NaN) until its next sort; the
     * oracle deletes in place, keeping the order of the live particles */
    int64_t live = 0;
    for (int64_t i = 0; io.ok && i < n; i++) {
      if (isnan(arr[2][i])) continue;
      for (int a = 0; a < 6; a++) arr[a][live] = arr[a][i];
      live++;
    }
    if (io.ok)
      rc = orc_upload_particles(o, s, live, arr[0], arr[1], arr[2], arr[3],
                                arr[4], arr[5]);
    for (int a = 0; a < 6; a++) {
      free(arr[a]);
      arr[a] = NULL;
    }
    if (rc) goto done;
  }
  if (h.ntiles > 0) {
    uint32_t t;
    for (uint64_t i = 0; i < 2 * ((uint64_t)h.ntiles + 1) && io.ok; i++)
      io_get(&io, &t, sizeof t);
  }
  if (!io.ok) {
    rc = PIC_E_IO;
    goto done;
  }
  if (io.bytes != h.payload_bytes || h_finish(&io.hash) != h.checksum) {
    rc = PIC_E_FORMAT;
    goto done;
  }
  o->step = h.step;
  o->c_add = h.c_add;
  o->c_sort = h.c_sort;
  o->inflow_event = h.inflow_event;
done:
  free(plane);
  for (int a = 0; a < 6; a++) free(arr[a]);
  fclose(io.f);
  return rc;
}

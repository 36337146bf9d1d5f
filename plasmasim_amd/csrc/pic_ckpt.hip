// Checkpoint and resume of a context: pic_checkpoint_save / _load.
//
// Replaces the reference's suspend_job / continue_job (:7629-7911).  The
// state between steps is written in the context's own numeric form and
// streamed through a bounded pinned host buffer, so a checkpoint of a
// multi-GB state needs only that buffer on the host.  The file layout and
// the payload hash are specified in include/plasmasim.h.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>

#include "pic_internal.h"

namespace pic {
namespace {

constexpr char kMagic[8] = {'P', 'L', 'A', 'S', 'C', 'K', 'P', 'T'};
constexpr uint32_t kVersion = 1;
constexpr size_t kStage = (size_t)64 << 20;  // pinned staging bytes

// four interleaved FNV-1a-style lanes over 64-bit little-endian words
struct Hash4 {
  uint64_t h[4];
  uint64_t nwords = 0;
  unsigned char tail[8];
  int ntail = 0;
  Hash4() {
    for (int j = 0; j < 4; ++j) h[j] = 0xcbf29ce484222325ull + (uint64_t)j;
  }
  void word(uint64_t w) {
    uint64_t &x = h[nwords & 3];
    x = (x ^ w) * 0x100000001b3ull;
    ++nwords;
  }
  void bytes(const void *p, size_t n) {
    const unsigned char *c = (const unsigned char *)p;
    while (ntail && n) {
      tail[ntail++] = *c++;
      --n;
      if (ntail == 8) {
        uint64_t w;
        memcpy(&w, tail, 8);
        word(w);
        ntail = 0;
      }
    }
    for (; n >= 8; n -= 8, c += 8) {
      uint64_t w;
      memcpy(&w, c, 8);
      word(w);
    }
    while (n) {
      tail[ntail++] = *c++;
      --n;
    }
  }
  uint64_t finish() {
    if (ntail) {
      memset(tail + ntail, 0, 8 - ntail);
      uint64_t w;
      memcpy(&w, tail, 8);
      word(w);
      ntail = 0;
    }
    return ((h[0] * 31 + h[1]) * 31 + h[2]) * 31 + h[3];
  }
};

// the config fields that must agree between a checkpoint and the context
// that loads it (capacity, deposit mode and tiles may differ); copied field
// by field so that struct padding does not take part in the comparison
pic_config canon(const pic_config &c) {
  pic_config z;
  memset(&z, 0, sizeof z);
  z.nx = c.nx;
  z.ny = c.ny;
  z.nz = c.nz;
  z.ppc = c.ppc;
  z.dx = c.dx;
  z.dt = c.dt;
  z.c = c.c;
  z.eps0 = c.eps0;
  z.mu0 = c.mu0;
  z.q = c.q;
  for (int s = 0; s < 2; ++s) {
    z.mass[s] = c.mass[s];
    z.weight[s] = c.weight[s];
    z.vth[s] = c.vth[s];
    z.periodic[s] = c.periodic[s];
  }
  for (int a = 0; a < 3; ++a) {
    z.vdrift[a] = c.vdrift[a];
    z.imf[a] = c.imf[a];
    z.absorb[a] = c.absorb[a];
    z.ballistic_lo[a] = c.ballistic_lo[a];
    z.ballistic_hi[a] = c.ballistic_hi[a];
  }
  z.dipole_m = c.dipole_m;
  z.dipole_alpha = c.dipole_alpha;
  z.dipole_soft = c.dipole_soft;
  z.dipole_rmax = c.dipole_rmax;
  z.sigma_amp = c.sigma_amp;
  z.inflow_every = c.inflow_every;
  z.sort_every = c.sort_every;
  z.refill_far = c.refill_far;
  z.precision = c.precision;
  z.seed = c.seed;
  z.poisson_every = c.poisson_every;
  z.phi_margin = c.phi_margin;
  return z;
}

struct Io {
  pic_ctx *ctx;
  FILE *f = nullptr;
  char *stage = nullptr;
  Hash4 hash;
  uint64_t bytes = 0;
  bool ok = true;
  ~Io() {
    if (f) fclose(f);
    if (stage) (void)hipHostFree(stage);
  }
  void put(const void *p, size_t n) {
    if (!ok) return;
    hash.bytes(p, n);
    ok = fwrite(p, 1, n, f) == n;
    bytes += n;
  }
  void get(void *p, size_t n) {
    if (!ok) return;
    ok = fread(p, 1, n, f) == n;
    if (ok) hash.bytes(p, n);
    bytes += n;
  }
};

int put_device(Io &io, const void *dev, size_t n) {
  for (size_t off = 0; off < n && io.ok; off += kStage) {
    const size_t c = std::min(kStage, n - off);
    PIC_HIP(io.ctx, hipMemcpy(io.stage, (const char *)dev + off, c,
                              hipMemcpyDeviceToHost));
    io.put(io.stage, c);
  }
  return io.ok ? 0 : set_error(io.ctx, PIC_E_IO, "checkpoint: write failed");
}

int get_device(Io &io, void *dev, size_t n) {
  for (size_t off = 0; off < n && io.ok; off += kStage) {
    const size_t c = std::min(kStage, n - off);
    io.get(io.stage, c);
    if (!io.ok) break;
    PIC_HIP(io.ctx, hipMemcpy((char *)dev + off, io.stage, c,
                              hipMemcpyHostToDevice));
  }
  return io.ok ? 0 : set_error(io.ctx, PIC_E_IO, "checkpoint: file truncated");
}

// device arrays of one species' particle slots, in file order
int particle_arrays(pic_ctx *ctx, int sp, void *ptr[7], size_t esz[7]) {
  if (ctx->cfg.precision == PIC_FP64) {
    const Particles64 P = p64(ctx, sp, ctx->cur[sp]);
    double *a[6] = {P.x, P.y, P.z, P.vx, P.vy, P.vz};
    for (int i = 0; i < 6; ++i) {
      ptr[i] = a[i];
      esz[i] = sizeof(double);
    }
    return 6;
  }
  const Particles32 P = p32(ctx, sp, ctx->cur[sp]);
  ptr[0] = P.cell;
  esz[0] = sizeof(int32_t);
  float *a[6] = {P.ox, P.oy, P.oz, P.vx, P.vy, P.vz};
  for (int i = 0; i < 6; ++i) {
    ptr[i + 1] = a[i];
    esz[i + 1] = sizeof(float);
  }
  return 7;
}

}  // namespace
}  // namespace pic

using namespace pic;

extern "C" int pic_checkpoint_save(pic_ctx *ctx, const char *path) {
  if (!ctx || !path) return PIC_E_ARG;
  const bool f64 = ctx->cfg.precision == PIC_FP64;
  if (!f64) {
    int rc = f32::finish_deposit(ctx);
    if (!rc) rc = f32::settle_sorts(ctx);
    if (rc) return rc;
  }
  PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const Geom &g = ctx->geom;
  const size_t nc = (size_t)g.SX * g.SY * g.SZ;
  const size_t esz = f64 ? sizeof(double) : sizeof(float);
  const TileGeom tg = tile_geom(ctx);
  const uint32_t ntiles = (uint32_t)tg.ntiles();

  pic_checkpoint_header h;
  memset(&h, 0, sizeof h);
  memcpy(h.magic, kMagic, 8);
  h.version = kVersion;
  h.header_bytes = sizeof h;
  h.config_bytes = sizeof(pic_config);
  h.precision = ctx->cfg.precision;
  h.nx = ctx->cfg.nx;
  h.ny = ctx->cfg.ny;
  h.nz = ctx->cfg.nz;
  h.tile[0] = tg.TX;
  h.tile[1] = tg.TY;
  h.tile[2] = tg.TZ;
  h.step = ctx->step;
  h.c_add = ctx->c_add;
  h.c_sort = ctx->c_sort;
  h.inflow_event = ctx->inflow_event;
  h.ntiles = ntiles;
  for (int s = 0; s < 2; ++s) {
    h.n[s] = ctx->n[s];
    h.sorted_end[s] = ctx->sorted_end[s];
    h.drift_halves[s] = f64 ? 0 : ctx->drift_halves[s];
  }

  Io io{ctx};
  io.f = fopen(path, "wb");
  if (!io.f) return set_error(ctx, PIC_E_IO, "checkpoint: cannot create %s", path);
  PIC_HIP(ctx, hipHostMalloc(&io.stage, kStage));
  // header placeholder, rewritten with the size and hash at the end
  if (fwrite(&h, 1, sizeof h, io.f) != sizeof h)
    return set_error(ctx, PIC_E_IO, "checkpoint: write failed");
  io.put(&ctx->cfg, sizeof(pic_config));
  int rc;
  for (int i = 0; i < PIC_NFIELDS; ++i)
    if ((rc = put_device(io, (const char *)ctx->field_mem + (size_t)i * g.plane * esz,
                         nc * esz)))
      return rc;
  if (!f64) io.put(ctx->b0_host, 3 * nc * sizeof(double));
  for (int s = 0; s < 2; ++s) {
    void *ptr[7];
    size_t pe[7];
    const int na = particle_arrays(ctx, s, ptr, pe);
    for (int a = 0; a < na; ++a)
      if ((rc = put_device(io, ptr[a], (size_t)ctx->n[s] * pe[a]))) return rc;
  }
  for (int s = 0; s < 2; ++s)
    if ((rc = put_device(io, ctx->tile_start[s], (size_t)(ntiles + 1) * sizeof(uint32_t))))
      return rc;
  if (!io.ok) return set_error(ctx, PIC_E_IO, "checkpoint: write failed");
  h.payload_bytes = io.bytes;
  h.checksum = io.hash.finish();
  if (fseek(io.f, 0, SEEK_SET) != 0 || fwrite(&h, 1, sizeof h, io.f) != sizeof h ||
      fflush(io.f) != 0)
    return set_error(ctx, PIC_E_IO, "checkpoint: write failed");
  return 0;
}

extern "C" int pic_checkpoint_load(pic_ctx *ctx, const char *path) {
  if (!ctx || !path) return PIC_E_ARG;
  const bool f64 = ctx->cfg.precision == PIC_FP64;
  if (!f64) {
    int rc = f32::settle_sorts(ctx);  // the load replaces the particle state
    if (rc) return rc;
  }
  PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const Geom &g = ctx->geom;
  const size_t nc = (size_t)g.SX * g.SY * g.SZ;
  const size_t esz = f64 ? sizeof(double) : sizeof(float);
  const TileGeom tg = tile_geom(ctx);

  Io io{ctx};
  io.f = fopen(path, "rb");
  if (!io.f) return set_error(ctx, PIC_E_IO, "checkpoint: cannot open %s", path);
  pic_checkpoint_header h;
  if (fread(&h, 1, sizeof h, io.f) != sizeof h)
    return set_error(ctx, PIC_E_FORMAT, "checkpoint: short header");
  if (memcmp(h.magic, kMagic, 8) != 0 || h.version != kVersion ||
      h.header_bytes != sizeof h)
    return set_error(ctx, PIC_E_FORMAT, "checkpoint: not a version-%u checkpoint",
                     kVersion);
  if (h.config_bytes != sizeof(pic_config))
    return set_error(ctx, PIC_E_FORMAT, "checkpoint: config size %u, library %zu",
                     h.config_bytes, sizeof(pic_config));
  if (h.precision != ctx->cfg.precision)
    return set_error(ctx, PIC_E_FORMAT, "checkpoint: numeric form %d, context %d",
                     h.precision, ctx->cfg.precision);
  for (int s = 0; s < 2; ++s)
    if (h.n[s] < 0 || h.n[s] > ctx->cap || h.sorted_end[s] < 0 ||
        h.sorted_end[s] > h.n[s])
      return set_error(ctx, PIC_E_CAPACITY,
                       "checkpoint: %lld slots of species %d, capacity %lld",
                       (long long)h.n[s], s, (long long)ctx->cap);
  PIC_HIP(ctx, hipHostMalloc(&io.stage, kStage));
  pic_config fc;
  io.get(&fc, sizeof fc);
  if (!io.ok) return set_error(ctx, PIC_E_FORMAT, "checkpoint: short config");
  const pic_config a = canon(fc), b = canon(ctx->cfg);
  if (memcmp(&a, &b, sizeof a) != 0)
    return set_error(ctx, PIC_E_FORMAT,
                     "checkpoint: config differs from the context's");
  int rc;
  for (int i = 0; i < PIC_NFIELDS; ++i)
    if ((rc = get_device(io, (char *)ctx->field_mem + (size_t)i * g.plane * esz,
                         nc * esz)))
      return rc;
  if (!f64) {
    io.get(ctx->b0_host, 3 * nc * sizeof(double));
    if (!io.ok) return set_error(ctx, PIC_E_IO, "checkpoint: file truncated");
  }
  for (int s = 0; s < 2; ++s) {
    ctx->n[s] = h.n[s];
    void *ptr[7];
    size_t pe[7];
    const int na = particle_arrays(ctx, s, ptr, pe);
    for (int k = 0; k < na; ++k)
      if ((rc = get_device(io, ptr[k], (size_t)h.n[s] * pe[k]))) return rc;
  }
  const bool same_tiles = h.ntiles == (uint32_t)tg.ntiles() && h.tile[0] == tg.TX &&
                          h.tile[1] == tg.TY && h.tile[2] == tg.TZ;
  if (h.ntiles > 0) {
    const size_t tb = (size_t)(h.ntiles + 1) * sizeof(uint32_t);
    for (int s = 0; s < 2; ++s) {
      if (same_tiles) {
        if ((rc = get_device(io, ctx->tile_start[s], tb))) return rc;
      } else {
        for (size_t off = 0; off < tb && io.ok; off += kStage)
          io.get(io.stage, std::min(kStage, tb - off));
      }
    }
  }
  if (!io.ok) return set_error(ctx, PIC_E_IO, "checkpoint: file truncated");
  if (io.bytes != h.payload_bytes || io.hash.finish() != h.checksum)
    return set_error(ctx, PIC_E_FORMAT,
                     "checkpoint: payload size or checksum mismatch (state "
                     "overwritten)");
  ctx->step = h.step;
  ctx->c_add = h.c_add;
  ctx->c_sort = h.c_sort;
  ctx->inflow_event = h.inflow_event;
  for (int s = 0; s < 2; ++s) {
    ctx->sorted_end[s] = same_tiles && h.ntiles > 0 ? h.sorted_end[s] : 0;
    // the windows of restored tiles follow the drift from where the writer's
    // stood, so a resumed run deposits exactly as an uninterrupted one
    ctx->drift_halves[s] = ctx->sorted_end[s] > 0 ? h.drift_halves[s] : 0;
  }
  if (!f64) return f32::state_restored(ctx);
  return 0;
}

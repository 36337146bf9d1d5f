// Internal state of a pic_ctx and helpers shared by the HIP translation
// units.  Not part of the C-ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/plasmasim.h"

namespace pic {

constexpr int kBlock = 256;  // threads per workgroup for streaming kernels

// Device view of the grid geometry.
struct Geom {
  int NX, NY, NZ;  // interior cells
  int SX, SY, SZ;  // with 2 guards per side
  int64_t plane;   // elements between consecutive field planes
  __host__ __device__ inline int64_t idx(int k, int l, int m) const {
    return ((int64_t)k * SY + l) * SZ + m;
  }
};

// Sort tiles: key = tile-major, lexicographic (k,l,m) inside a tile, over
// interior cell indices.  One tile spanning the interior gives the
// reference's key ((k*NY)+l)*NZ+m.
struct TileGeom {
  int TX, TY, TZ;     // cells per tile
  int NTX, NTY, NTZ;  // tiles per axis
  int shift[3];       // cells by which the tiles' LDS windows are moved
  int LX, LY, LZ;     // log2 of TX, TY, TZ when all three are powers of two, else -1
  __host__ __device__ inline uint32_t key(int k, int l, int m) const {
    if (LX >= 0) {  // shifts and masks instead of integer divisions
      const uint32_t tile =
          ((uint32_t)(k >> LX) * NTY + (uint32_t)(l >> LY)) * NTZ + (uint32_t)(m >> LZ);
      const uint32_t in = ((((uint32_t)k & (TX - 1)) << LY | ((uint32_t)l & (TY - 1)))
                           << LZ) | ((uint32_t)m & (TZ - 1));
      return (tile << (LX + LY + LZ)) | in;
    }
    const uint32_t tile = ((uint32_t)(k / TX) * NTY + l / TY) * NTZ + m / TZ;
    const uint32_t in = ((uint32_t)(k % TX) * TY + l % TY) * TZ + m % TZ;
    return tile * (uint32_t)(TX * TY * TZ) + in;
  }
  // tile and cell-in-tile of a key
  __host__ __device__ inline uint32_t tile_of(uint32_t key) const {
    return LX >= 0 ? key >> (LX + LY + LZ) : key / (uint32_t)(TX * TY * TZ);
  }
  __host__ __device__ inline uint32_t cell_of(uint32_t key) const {
    return LX >= 0 ? key & (uint32_t)(TX * TY * TZ - 1) : key % (uint32_t)(TX * TY * TZ);
  }
  __host__ __device__ inline int ntiles() const { return NTX * NTY * NTZ; }
};

// Per-axis sigma ingredients on the 2x lattice: depth^4 and depth^12.
struct SigmaProfile {
  const double *d4[3];
  const double *d12[3];
  double amp;
};

// Field planes in the order of enum pic_field.  In the FP32 form the B
// planes hold the induced field B - B0 (the curl only needs B - B0, and
// storing it avoids the cancellation of B - B0 in single precision near the
// dipole); downloads add B0 back.
template <typename T>
struct Fields {
  T *f[PIC_NFIELDS];
};

struct Particles64 {
  double *x, *y, *z, *vx, *vy, *vz;
};

// FP32 form: cell index into the guard-inclusive grid, offsets in [0,1).
struct Particles32 {
  int32_t *cell;
  float *ox, *oy, *oz, *vx, *vy, *vz;
};

// Constants of one step shared by the particle kernels.
struct StepConst {
  double dx, dt, c, q;
  double mass;      // signed: +m_i for ions, -m_e for electrons (:784)
  double weight;    // p->n
  double vdrift_z;  // V in cal_track4's IMF terms (:2817-2818)
  double imf_x, imf_y;
  double blo[3], bhi[3];  // ballistic zone bounds
  int periodic[2];
};

}  // namespace pic

namespace pic {
namespace tsort {
struct Bufs;
}
}  // namespace pic

struct pic_ctx {
  pic_config cfg;
  pic::Geom geom;
  hipStream_t stream = nullptr;
  char err[512];

  // fields: PIC_NFIELDS planes of double (FP64) or float (FP32)
  void *field_mem = nullptr;
  size_t field_bytes = 0;
  double *sig_mem = nullptr;  // 6 x 2*Smax doubles
  int sig_stride = 0;
  double *b0_host = nullptr;  // FP32 form: B0 in double, 3 planes

  // particles: two buffers per species, swapped by compaction and sort
  int64_t cap = 0;
  int64_t n[2] = {0, 0};  // stored slots, tombstones included
  void *pbuf[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  int cur[2] = {0, 0};

  // scratch for compaction and sorting (sized to the particle capacity)
  uint32_t *keys = nullptr, *keys_alt = nullptr;
  // idx: sort and selection output; idx_alt: the ions' sort permutation
  // while it waits for their fused pass (FP32 form, inside pic_step)
  uint32_t *idx = nullptr, *idx_alt = nullptr;
  uint8_t *flags = nullptr;
  void *prim_tmp = nullptr;        // radix sort and selection temporaries
  size_t prim_tmp_bytes = 0;
  int64_t *dev_counter = nullptr;  // 8 int64 device scratch slots
  int64_t *host_counter = nullptr; // pinned mirror
  double *diag_partial = nullptr;  // reduction partials
  int64_t diag_partial_cap = 0;
  int device = 0;                   // HIP device current at pic_create
  double *noalloc_scratch = nullptr;  // this device's static scratch array

  // thrust_fave accumulated by pic_step (pic_thrust_ave_enable).  The force
  // of step c is summed during the particle pass of step c + 1, which
  // gathers the same positions, velocities and E; B of the end of step c is
  // kept in thrust_bprev before the first B half-step of step c + 1.
  int thrust_on = 0;
  int thrust_pending = 0;            // the last step's force not yet added
  int thrust_now = 0;                // the running pass adds it
  int64_t thrust_steps = 0;          // steps added to thrust_acc
  double *thrust_acc = nullptr;      // device: per species E part (3), v x B part (3)
  double *thrust_part = nullptr;     // device: 6 partials per pass workgroup
  int64_t thrust_part_slots = 0;
  void *thrust_bprev = nullptr;      // device: 3 FP64 B planes (the FP32 form
                                     // keeps its copy beside its total-B planes)

  // sort tiles: per-species particle ranges of each tile after the last
  // sort; [0, sorted_end) is tile-sorted, later slots are the inflow tail
  uint32_t *tile_start[2] = {nullptr, nullptr};   // ntiles + 1 per species
  int64_t sorted_end[2] = {0, 0};
  int32_t drift_halves[2] = {0, 0};  // half-step drifts since the tile sort
  int64_t sort_tile_count = 0;   // sorts done by the tile sort
  pic::tsort::Bufs *ts64 = nullptr;  // the FP64 form's tile-sort lists (pic_tsort.h)
  int64_t sort_radix_count = 0;  // sorts done by the radix sort

  // state private to the numeric form (FP32: sigma copy, tile ranges,
  // deterministic accumulators)
  void *form_state = nullptr;

  // adjoint: checkpoints, tape and adjoint planes (pic_adj.hip)
  void *adj = nullptr;
  // adjoint sweeps: inflows append to the store, so that each batch of
  // injected particles stays contiguous for the tape
  int adj_mode = 0;

  // 3D Poisson correction (poisson_every > 0): phi (a field plane of the
  // form's type, guard-inclusive), the transform buffer rho (nx ny nz) and
  // the transform tables of the three axes plus the per-line symbol of the
  // x pass (pic_poisson.hip)
  void *phi = nullptr, *rho = nullptr;
  void *dst_tab[3] = {nullptr, nullptr, nullptr};
  double *sym_yz = nullptr;

  // step counters (reference fdtd(): c, c_add, c_thrust)
  int64_t step = 0;
  int32_t c_add = 0, c_sort = 0;
  uint32_t inflow_event = 0;

  // timing: pairs of HIP events around launches, resolved in batches
  static constexpr int kEventPool = 512;
  int timing = 0;  // 0 off, 1 every class, 2 the particle passes only
  double t_ms[PIC_T_COUNT] = {0};
  int64_t t_n[PIC_T_COUNT] = {0};
  hipEvent_t tev[2][kEventPool] = {};
  int tclass[kEventPool] = {};
  int tpending = 0;
};

namespace pic {

int set_error(pic_ctx *ctx, int code, const char *fmt, ...);

#define PIC_HIP(ctx, call)                                                  \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return ::pic::set_error((ctx), PIC_E_HIP, "%s:%d %s: %s", __FILE__,   \
                              __LINE__, #call, hipGetErrorString(e_));      \
  } while (0)

#define PIC_LAUNCH_CHECK(ctx) PIC_HIP(ctx, hipGetLastError())

inline int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }

inline Particles64 p64(pic_ctx *ctx, int sp, int buf) {
  double *b = (double *)ctx->pbuf[sp][buf];
  const int64_t c = ctx->cap;
  return Particles64{b, b + c, b + 2 * c, b + 3 * c, b + 4 * c, b + 5 * c};
}
inline Particles32 p32(pic_ctx *ctx, int sp, int buf) {
  char *b = (char *)ctx->pbuf[sp][buf];
  const int64_t c = ctx->cap;
  float *f = (float *)(b + 4 * c);
  return Particles32{(int32_t *)b, f, f + c, f + 2 * c, f + 3 * c, f + 4 * c,
                     f + 5 * c};
}
template <typename T>
inline Fields<T> fields(pic_ctx *ctx) {
  Fields<T> F;
  T *base = (T *)ctx->field_mem;
  for (int i = 0; i < PIC_NFIELDS; ++i) F.f[i] = base + i * ctx->geom.plane;
  return F;
}
inline SigmaProfile sigma(pic_ctx *ctx) {
  SigmaProfile s;
  for (int a = 0; a < 3; ++a) {
    s.d4[a] = ctx->sig_mem + (2 * a) * ctx->sig_stride;
    s.d12[a] = ctx->sig_mem + (2 * a + 1) * ctx->sig_stride;
  }
  s.amp = ctx->cfg.sigma_amp;
  return s;
}
StepConst step_const(const pic_ctx *ctx, int sp);
inline TileGeom tile_geom(const pic_ctx *ctx) {
  const pic_config &c = ctx->cfg;
  TileGeom t;
  t.TX = c.tile[0] > 0 ? c.tile[0] : c.nx;
  t.TY = c.tile[1] > 0 ? c.tile[1] : c.ny;
  t.TZ = c.tile[2] > 0 ? c.tile[2] : c.nz;
  t.NTX = c.nx / t.TX;
  t.NTY = c.ny / t.TY;
  t.NTZ = c.nz / t.TZ;
  for (int a = 0; a < 3; ++a) t.shift[a] = 0;
  const auto lg = [](int v) {
    int b = 0;
    while ((1 << b) < v) ++b;
    return (1 << b) == v ? b : -1;
  };
  t.LX = lg(t.TX);
  t.LY = lg(t.TY);
  t.LZ = lg(t.TZ);
  if (t.LX < 0 || t.LY < 0 || t.LZ < 0) t.LX = t.LY = t.LZ = -1;
  return t;
}

// timing helpers (HIP events on the context stream)
void timing_begin(pic_ctx *ctx, int which);
void timing_end(pic_ctx *ctx, int which);
int timing_flush(pic_ctx *ctx);

// per-form implementations (fp64 / fp32 translation units)
namespace f64 {
int renew_grid(pic_ctx *);
int renew_b_half(pic_ctx *);
int halo(pic_ctx *, int first);
int fold_current(pic_ctx *);
int renew_e(pic_ctx *);
int load_particles(pic_ctx *);
int push(pic_ctx *, int sp, int do_push_j, int do_drift_n, int do_boundary);
int particle_boundary(pic_ctx *, int sp);
int inflow(pic_ctx *, int sp, uint32_t event);
int sort(pic_ctx *, int sp);
int tsort_setup(pic_ctx *);  // the tile sort's lists, at create
void tsort_free(pic_ctx *);
int diagnose(pic_ctx *, pic_diag *);
int count_live(pic_ctx *, int sp, int64_t *n);
}  // namespace f64
namespace f32 {
int renew_grid(pic_ctx *);
int renew_b_half(pic_ctx *);
int halo(pic_ctx *, int first);
int fold_current(pic_ctx *);
int renew_e(pic_ctx *);
int load_particles(pic_ctx *);
int push(pic_ctx *, int sp, int do_push_j, int do_drift_n, int do_boundary);
int particle_boundary(pic_ctx *, int sp);
int inflow(pic_ctx *, int sp, uint32_t event);
int sort(pic_ctx *, int sp);
// pic_step's tile sort: the permutation is left for the species' next fused
// pass, which reads the particles through it (no copy); settle applies a
// left permutation with a copy, settle_sorts for both species
int sort_for_pass(pic_ctx *, int sp);
int settle(pic_ctx *, int sp);
int settle_sorts(pic_ctx *);
bool pending(const pic_ctx *, int sp);
int diagnose(pic_ctx *, pic_diag *);
int count_live(pic_ctx *, int sp, int64_t *n);
int upload_particles(pic_ctx *, int sp, int64_t n, const double *const *h);
int download_particles(pic_ctx *, int sp, int64_t cap, int64_t *n,
                       double *const *h);
int finish_deposit(pic_ctx *);
// after a checkpoint load: total-B planes rebuilt, no pending fixed-point
// deposit
int state_restored(pic_ctx *);
int init_state(pic_ctx *);
void free_state(pic_ctx *);
int refresh_total_b(pic_ctx *);
int download_raw32(pic_ctx *, int sp, int64_t cap, int64_t *n, int32_t *cell,
                   float *const *f);
}  // namespace f32

// Diagnostics partials, the same layout in both forms: kDiagNBF workgroup
// partials of kDiagNQF field quantities, then per species kDiagNBP
// workgroup partials of kDiagNQP particle quantities.  Summed in workgroup
// order per quantity (quantity 5 of the field partials by max).
constexpr int kDiagNBF = 1024, kDiagNBP = 2048;
constexpr int kDiagNQF = 9, kDiagNQP = 21;
constexpr int64_t kDiagPartials =
    (int64_t)kDiagNBF * kDiagNQF + 2 * (int64_t)kDiagNBP * kDiagNQP;
namespace f64 {
// launch the partial sums into partials[kDiagPartials] and add the live
// particle counts of the two species into counts[0..1] (zeroed by the caller)
int diagnose_partials(pic_ctx *, double *partials, unsigned long long *counts);
}
namespace f32 {
int diagnose_partials(pic_ctx *, double *partials, unsigned long long *counts);
}
// thrust_fave: add one pass's workgroup partials (6 per workgroup, nb
// workgroups) times qn = q_s w dx^3 to the species' running sums
int thrust_add(pic_ctx *ctx, int sp, int nb, double qn);
namespace f32 {
float *total_b(pic_ctx *);  // the 3 total-B planes the gather reads
}
// static device scratch of the allocation-free entry points (pic_noalloc.hip):
// resolved for the current device at pic_create
int noalloc_init(pic_ctx *ctx);

// the fused particle pass of one species (gather, push, J and n deposit,
// particle boundary), as pic_step runs it (pic_api.hip)
int push_fused(pic_ctx *ctx, int sp);
// adjoint state (pic_adj.hip)
void adj_free(pic_ctx *ctx);

// device-wide building blocks, hand-written (pic_prim.hip)
// stable selection: idx_out[0..*count) = indices i < n with flags[i] != 0
int select_flagged(pic_ctx *ctx, const uint8_t *flags, uint32_t *idx_out,
                   int64_t n, int64_t *count);
// stable key sort of (key, index) pairs; indices start as 0..n-1
int sort_keys(pic_ctx *ctx, uint32_t *keys_in, uint32_t *keys_out,
              uint32_t *idx_out, int64_t n, int end_bit);
int ensure_prim_tmp(pic_ctx *ctx, size_t bytes);
// 3D Poisson correction (pic_poisson.hip)
int poisson_init(pic_ctx *ctx);
void poisson_free(pic_ctx *ctx);
int poisson_zero_phi(pic_ctx *ctx);
int poisson(pic_ctx *ctx);  // renew_phi3 + sr_phi + E correction, no sr_e
int poisson_download_phi(pic_ctx *ctx, double *host);

// tile_start[0..nt] from keys sorted ascending (cpt cells per tile; keys
// >= nt * cpt are dropped slots and land in tile nt)
int tile_bounds(pic_ctx *ctx, const uint32_t *skeys, int64_t n, uint32_t cpt,
                int nt, uint32_t *tile_start);

}  // namespace pic

// Internal state of a pic2d_ctx (the 2D3V step) and helpers shared by its
// translation units.  Not part of the C-ABI.
//
// The context embeds a pic_ctx as its runtime: the HIP stream, the error
// text, the compaction/sort scratch, the sort's temporary storage and the
// timing events are the 3D context's, so the device-wide building blocks
// of pic_prim.hip serve both.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/plasmasim2d.h"
#include "pic_internal.h"

namespace pic2d {

using pic::kBlock;
using pic::div_up;

struct Geom2 {
  int NX, NZ;  // interior cells
  int SX, SZ;  // with 2 guards per side
  int64_t plane;
  __host__ __device__ inline int64_t idx(int k, int m) const {
    return (int64_t)k * SZ + m;
  }
};

// sort tiles over interior cells: tile-major, lexicographic (k,m) inside
struct Tile2 {
  int TX, TZ, NTX, NTZ;
  __host__ __device__ inline uint32_t key(int k, int m) const {
    const uint32_t tile = (uint32_t)(k / TX) * NTZ + m / TZ;
    return tile * (uint32_t)(TX * TZ) + (uint32_t)(k % TX) * TZ + m % TZ;
  }
  __host__ __device__ inline int ntiles() const { return NTX * NTZ; }
};

// FP64 particles: x, z, vx, vy, vz (reference units, +2 guard offset); a
// dead slot has x = NaN
struct P64 {
  double *x, *z, *vx, *vy, *vz;
};
// FP32 particles: cell k*SZ+m (guard-inclusive; -1 dead), offsets in [0,1)
struct P32 {
  int32_t *cell;
  float *ox, *oz, *vx, *vy, *vz;
};

// constants of the particle pass for one species
struct Push2 {
  double dx, dt, q, T;   // T = q dt / (2 m), m signed (-m_e for electrons)
  double qw, w;          // q * weight, weight
  double eimf_x, eimf_y; // V*IMF_y and -V*IMF_x (cal_track_c4 :1534-1535)
  double blo[2], bhi[2]; // ballistic bounds (x, z), length units
};

}  // namespace pic2d

namespace pic {
namespace tsort {
struct Bufs;
}
}  // namespace pic

struct pic2d_ctx {
  pic_ctx rt;  // stream, error text, scratch, timing (its cfg is unused)
  pic2d_config cfg;
  pic2d::Geom2 g;
  pic2d::Tile2 tg;
  bool f64 = true;

  void *field_mem = nullptr;  // PIC2D_NFIELDS planes of double or float
  double *sig64 = nullptr;    // 2SX x 2SZ sigma lattice (FP64 form)
  float *sig32 = nullptr;     // the same in fp32 (FP32 form)
  void *rho = nullptr;        // NX x NZ Poisson work plane
  void *tw_x = nullptr;       // FFT twiddles for the x and z DSTs
  void *tw_z = nullptr;

  int64_t cap = 0;
  int64_t n[2] = {0, 0};  // stored slots, dead ones included
  void *pbuf[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  int cur[2] = {0, 0};
  uint32_t *tile_start[2] = {nullptr, nullptr};
  int64_t sorted_end[2] = {0, 0};

  // FP32 form: fixed-point scale of the LDS sums and, in deterministic
  // mode, 64-bit fixed-point planes jx, jy, jz, n of the species in flight
  double fx_scale = 1.0;
  unsigned long long *fx = nullptr;

  int64_t step = 0;
  int32_t c_add = 0, c_sort = 0;
  uint32_t inflow_event = 0;

  // FP32 form: the tile sort's device lists (pic_tsort.h) and how many
  // sorts took the tile sort and the radix sort
  pic::tsort::Bufs *ts = nullptr;
  int64_t sort_tile_count = 0, sort_radix_count = 0;
  // pic2d_step's tile sort leaves its permutation to the species' fused
  // pass of the same step (pending), which reads the particles through it:
  // the ions' waits in perm_ion, the electrons' in rt.idx
  uint32_t *perm_ion = nullptr;
  bool pending[2] = {false, false};
};

namespace pic2d {

#define P2_HIP(ctx, call) PIC_HIP(&(ctx)->rt, call)
#define P2_LAUNCH_CHECK(ctx) PIC_LAUNCH_CHECK(&(ctx)->rt)

inline int err(pic2d_ctx *ctx, int code, const char *msg) {
  return pic::set_error(&ctx->rt, code, "%s", msg);
}

template <typename T>
inline T *plane(pic2d_ctx *ctx, int which) {
  return (T *)ctx->field_mem + (int64_t)which * ctx->g.plane;
}
inline P64 p64(pic2d_ctx *ctx, int sp, int buf) {
  double *b = (double *)ctx->pbuf[sp][buf];
  const int64_t c = ctx->cap;
  return P64{b, b + c, b + 2 * c, b + 3 * c, b + 4 * c};
}
inline P32 p32(pic2d_ctx *ctx, int sp, int buf) {
  char *b = (char *)ctx->pbuf[sp][buf];
  const int64_t c = ctx->cap;
  float *f = (float *)(b + 4 * c);
  return P32{(int32_t *)b, f, f + c, f + 2 * c, f + 3 * c, f + 4 * c};
}
Push2 push_const(const pic2d_ctx *ctx, int sp);

// field routines (pic2d_fields.hip)
int renew_grid(pic2d_ctx *);
int renew_b_half(pic2d_ctx *);
int renew_e(pic2d_ctx *);
int halo(pic2d_ctx *);
int fold(pic2d_ctx *, int first, int count);
int external_current(pic2d_ctx *, int64_t step);
int poisson(pic2d_ctx *);
int init_poisson(pic2d_ctx *);
int diagnose_fields(pic2d_ctx *, pic2d_diag *);

// particle routines (pic2d_particles.hip)
int load_particles(pic2d_ctx *);
int push(pic2d_ctx *, int sp, int push_j, int drift_n, int boundary,
         int density_only);
int particle_boundary(pic2d_ctx *, int sp);
int inflow(pic2d_ctx *, int sp, uint32_t event);
int sort(pic2d_ctx *, int sp);
// the FP32 tile sort as pic2d_step runs it: the permutation is left to the
// species' next fused pass (no copy); settle applies a left permutation
// with a copy, settle_sorts for both species
int sort_for_pass(pic2d_ctx *, int sp);
int settle(pic2d_ctx *, int sp);
int settle_sorts(pic2d_ctx *);
int tsort_setup(pic2d_ctx *);  // the tile sort's lists, at create
void tsort_free(pic2d_ctx *);
int count_live(pic2d_ctx *, int sp, int64_t *n);
int upload_particles(pic2d_ctx *, int sp, int64_t n, const double *const *h);
int download_particles(pic2d_ctx *, int sp, int64_t cap, int64_t *n,
                       double *const *h);
int diagnose_particles(pic2d_ctx *, pic2d_diag *);
int finish_deposit(pic2d_ctx *);

}  // namespace pic2d

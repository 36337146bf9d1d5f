// FP32 chosen form (placeholder until implemented): every entry reports that
// the form is unavailable instead of computing anything.
#include <hip/hip_runtime.h>

#include "pic_internal.h"

namespace pic {
namespace f32 {
#define NA(ctx) set_error((ctx), PIC_E_STATE, "FP32 form not built")
int renew_grid(pic_ctx *c) { return NA(c); }
int renew_b_half(pic_ctx *c) { return NA(c); }
int halo(pic_ctx *c, int) { return NA(c); }
int fold_current(pic_ctx *c) { return NA(c); }
int renew_e(pic_ctx *c) { return NA(c); }
int load_particles(pic_ctx *c) { return NA(c); }
int push(pic_ctx *c, int, int, int, int) { return NA(c); }
int particle_boundary(pic_ctx *c, int) { return NA(c); }
int inflow(pic_ctx *c, int, uint32_t) { return NA(c); }
int sort(pic_ctx *c, int) { return NA(c); }
int diagnose(pic_ctx *c, pic_diag *) { return NA(c); }
int count_live(pic_ctx *c, int, int64_t *) { return NA(c); }
int upload_particles(pic_ctx *c, int, int64_t, const double *const *) { return NA(c); }
int download_particles(pic_ctx *c, int, int64_t, int64_t *, double *const *) { return NA(c); }
int finish_deposit(pic_ctx *) { return 0; }
}  // namespace f32
}  // namespace pic

extern "C" int pic_download_particles_raw32(pic_ctx *ctx, int, int64_t,
                                            int64_t *, int32_t *, float *,
                                            float *, float *, float *, float *,
                                            float *) {
  return pic::set_error(ctx, PIC_E_STATE, "FP32 form not built");
}

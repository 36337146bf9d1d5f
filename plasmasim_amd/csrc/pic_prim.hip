// Device-wide stable selection and key sort on rocPRIM.  They serve the
// amortised steps (inflow compaction every inflow_every steps, the cell sort
// every sort_every steps); the per-step kernels are hand-written.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "pic_internal.h"

namespace pic {

int ensure_prim_tmp(pic_ctx *ctx, size_t bytes) {
  if (bytes <= ctx->prim_tmp_bytes) return 0;
  if (ctx->prim_tmp) PIC_HIP(ctx, hipFree(ctx->prim_tmp));
  ctx->prim_tmp = nullptr;
  ctx->prim_tmp_bytes = 0;
  PIC_HIP(ctx, hipMalloc(&ctx->prim_tmp, bytes));
  ctx->prim_tmp_bytes = bytes;
  return 0;
}

int select_flagged(pic_ctx *ctx, const uint8_t *flags, uint32_t *idx_out,
                   int64_t n, int64_t *count) {
  *count = 0;
  if (n == 0) return 0;
  rocprim::counting_iterator<uint32_t> first(0u);
  uint32_t *sel = (uint32_t *)(ctx->dev_counter + 1);
  size_t bytes = 0;
  PIC_HIP(ctx, rocprim::select(nullptr, bytes, first, flags, idx_out, sel,
                               (size_t)n, ctx->stream));
  int rc = ensure_prim_tmp(ctx, bytes);
  if (rc) return rc;
  PIC_HIP(ctx, rocprim::select(ctx->prim_tmp, bytes, first, flags, idx_out,
                               sel, (size_t)n, ctx->stream));
  PIC_HIP(ctx, hipMemcpyAsync(ctx->host_counter + 1, sel, sizeof(uint32_t),
                              hipMemcpyDeviceToHost, ctx->stream));
  PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *count = (int64_t)(*(uint32_t *)(ctx->host_counter + 1));
  return 0;
}

int sort_keys(pic_ctx *ctx, uint32_t *keys_in, uint32_t *keys_out,
              uint32_t *idx_out, int64_t n, int end_bit) {
  if (n == 0) return 0;
  rocprim::counting_iterator<uint32_t> first(0u);
  size_t bytes = 0;
  PIC_HIP(ctx, rocprim::radix_sort_pairs(nullptr, bytes, keys_in, keys_out,
                                         first, idx_out, (size_t)n, 0u,
                                         (unsigned)end_bit, ctx->stream));
  int rc = ensure_prim_tmp(ctx, bytes);
  if (rc) return rc;
  PIC_HIP(ctx, rocprim::radix_sort_pairs(ctx->prim_tmp, bytes, keys_in,
                                         keys_out, first, idx_out, (size_t)n,
                                         0u, (unsigned)end_bit, ctx->stream));
  return 0;
}

// tile_start[t] = first sorted slot of tile t, from keys sorted ascending.
// Keys of dropped slots are at least nt * cpt and map to tile nt, so
// tile_start[nt] is the number of kept particles.  Each slot writes the
// starts of the tiles that begin at it (empty tiles included), so every
// entry is written exactly once and no counting pass is needed.
__global__ void k_tile_bounds(const uint32_t *__restrict__ skeys, int64_t n,
                              uint32_t cpt, int nt, uint32_t *tile_start) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int tc = (int)min((uint32_t)nt, skeys[i] / cpt);
  const int tp = i == 0 ? -1 : (int)min((uint32_t)nt, skeys[i - 1] / cpt);
  for (int t = tp + 1; t <= tc; ++t) tile_start[t] = (uint32_t)i;
  if (i == n - 1)
    for (int t = tc + 1; t <= nt; ++t) tile_start[t] = (uint32_t)n;
}

int tile_bounds(pic_ctx *ctx, const uint32_t *skeys, int64_t n, uint32_t cpt,
                int nt, uint32_t *tile_start) {
  if (n == 0) return 0;
  k_tile_bounds<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0,
                  ctx->stream>>>(skeys, n, cpt, nt, tile_start);
  PIC_LAUNCH_CHECK(ctx);
  return 0;
}

}  // namespace pic

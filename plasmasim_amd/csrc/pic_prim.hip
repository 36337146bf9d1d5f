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

// exclusive scan of the per-tile counts in one workgroup: start[t], t<=nt
__global__ __launch_bounds__(1024) void k_scan_tiles(const uint32_t *cnt,
                                                     uint32_t *start, int nt) {
  __shared__ uint32_t part[1024];
  const int t = threadIdx.x;
  const int per = (nt + 1023) / 1024;
  const int lo = t * per, hi = min(nt, lo + per);
  uint32_t s = 0;
  for (int i = lo; i < hi; ++i) s += cnt[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = part[t] - s;
  for (int i = lo; i < hi; ++i) {
    start[i] = run;
    run += cnt[i];
  }
  if (t == 1023) start[nt] = part[1023];
}

int scan_tiles(pic_ctx *ctx, int sp) {
  const int nt = tile_geom(ctx).ntiles();
  k_scan_tiles<<<1, 1024, 0, ctx->stream>>>(ctx->tile_count,
                                            ctx->tile_start[sp], nt);
  PIC_LAUNCH_CHECK(ctx);
  return 0;
}

}  // namespace pic

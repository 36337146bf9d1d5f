// Device-wide stable selection and stable key sort, hand-written.  They
// serve the unsorted states (the load, an upload, the compaction of the
// inflow's dead slots) and the downloads; a periodic sort of a store whose
// tiles are current takes the tile sort (pic_tsort.h).
//
// Both rest on one chunked exclusive scan: chunks of kChunk values (256
// threads x 16 consecutive values), a sum per chunk, one workgroup scanning
// the chunk sums, and each chunk scanned again with its offset.
//   select_flagged: the slots whose flag is set, in slot order;
//   sort_keys: least-significant-digit radix sort of (key, slot) pairs,
//     8-bit digits, each pass a digit histogram per tile of kChunk keys
//     (digit-major, so that one scan gives every (digit, tile) offset) and
//     a stable scatter that ranks the keys of a tile in rounds of 256 (by
//     round, wave and lane, as the slots are ordered).
#include <hip/hip_runtime.h>

#include "pic_internal.h"

namespace pic {

namespace {

constexpr int kT = 256;              // threads per workgroup
constexpr int kPer = 16;             // values per thread
constexpr int kChunk = kT * kPer;    // values per chunk or sort tile
constexpr int kTopMax = 1 << 22;     // chunk sums one workgroup scans (4096 per thread)
constexpr int kBins = 256;           // radix digit values

typedef unsigned long long u64;

// exclusive scan of the 256 threads' values over the workgroup; the total
// in *total
__device__ inline uint32_t block_exclusive(uint32_t v, uint32_t *total) {
  __shared__ uint32_t wsum[kT / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(inc, o);
    if (lane >= o) inc += u;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (int i = 0; i < kT / 64; ++i) {
    if (i < w) before += wsum[i];
    all += wsum[i];
  }
  __syncthreads();  // wsum is reused by the next call
  *total = all;
  return before + inc - v;
}

// the chunk sums of a uint32 array (c-th chunk: [c kChunk, (c+1) kChunk))
__global__ __launch_bounds__(kT) void k_chunk_sum(const uint32_t *__restrict__ a, int64_t m,
                                                  uint32_t *__restrict__ sums) {
  const int64_t base = (int64_t)blockIdx.x * kChunk + (int64_t)threadIdx.x * kPer;
  uint32_t s = 0;
  for (int j = 0; j < kPer; ++j)
    if (base + j < m) s += a[base + j];
  uint32_t total;
  (void)block_exclusive(s, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// exclusive scan in place of up to kTopMax chunk sums (one workgroup of
// 1024 threads, each a contiguous run of them); the grand total in sums[nc]
__global__ __launch_bounds__(1024) void k_scan_top(uint32_t *sums, int nc) {
  __shared__ uint32_t part[1024];
  const int per = (nc + 1023) / 1024;
  const int a = threadIdx.x * per, b = min(a + per, nc);
  uint32_t s = 0;
  for (int i = a; i < b; ++i) s += sums[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {  // inclusive Hillis-Steele scan
    const uint32_t v = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = threadIdx.x > 0 ? part[threadIdx.x - 1] : 0u;
  for (int i = a; i < b; ++i) {
    const uint32_t v = sums[i];
    sums[i] = run;
    run += v;
  }
  if (threadIdx.x == 1023) sums[nc] = part[1023];
}

// each chunk scanned in place (exclusive), offset by its chunk's prefix
__global__ __launch_bounds__(kT) void k_chunk_scan(uint32_t *a, int64_t m,
                                                   const uint32_t *__restrict__ sums) {
  const int64_t base = (int64_t)blockIdx.x * kChunk + (int64_t)threadIdx.x * kPer;
  uint32_t v[kPer], s = 0;
  for (int j = 0; j < kPer; ++j) {
    v[j] = base + j < m ? a[base + j] : 0u;
    s += v[j];
  }
  uint32_t total;
  uint32_t run = sums[blockIdx.x] + block_exclusive(s, &total);
  for (int j = 0; j < kPer; ++j) {
    if (base + j < m) a[base + j] = run;
    run += v[j];
  }
}

// exclusive scan in place of a[0, m); m <= kTopMax kChunk
int scan_exclusive(pic_ctx *ctx, uint32_t *a, int64_t m, uint32_t *sums) {
  const int nc = (int)((m + kChunk - 1) / kChunk);
  k_chunk_sum<<<(unsigned)nc, kT, 0, ctx->stream>>>(a, m, sums);
  k_scan_top<<<1, 1024, 0, ctx->stream>>>(sums, nc);
  k_chunk_scan<<<(unsigned)nc, kT, 0, ctx->stream>>>(a, m, sums);
  PIC_LAUNCH_CHECK(ctx);
  return 0;
}

// ---------------------------------------------------------------------------
// selection

__global__ __launch_bounds__(kT) void k_sel_count(const uint8_t *__restrict__ flags,
                                                  int64_t n, uint32_t *__restrict__ sums) {
  const int64_t base = (int64_t)blockIdx.x * kChunk + (int64_t)threadIdx.x * kPer;
  uint32_t s = 0;
  for (int j = 0; j < kPer; ++j)
    if (base + j < n) s += flags[base + j] != 0;
  uint32_t total;
  (void)block_exclusive(s, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kT) void k_sel_write(const uint8_t *__restrict__ flags,
                                                  int64_t n,
                                                  const uint32_t *__restrict__ sums,
                                                  uint32_t *__restrict__ out) {
  const int64_t base = (int64_t)blockIdx.x * kChunk + (int64_t)threadIdx.x * kPer;
  bool f[kPer];
  uint32_t s = 0;
  for (int j = 0; j < kPer; ++j) {
    f[j] = base + j < n && flags[base + j] != 0;
    s += f[j];
  }
  uint32_t total;
  uint32_t run = sums[blockIdx.x] + block_exclusive(s, &total);
  for (int j = 0; j < kPer; ++j)
    if (f[j]) out[run++] = (uint32_t)(base + j);
}

// ---------------------------------------------------------------------------
// radix sort

// rank of this lane among the active lanes (pred) of its wave holding the
// same value v (nbits wide), and the number of such lanes
__device__ inline int lane_rank(bool pred, uint32_t v, int nbits, int &count) {
  const int lane = threadIdx.x & 63;
  u64 same = __builtin_amdgcn_ballot_w64(pred);
  for (int b = 0; b < nbits; ++b) {
    const u64 on = __builtin_amdgcn_ballot_w64(pred && ((v >> b) & 1u));
    same &= ((v >> b) & 1u) ? on : ~on;
  }
  count = __builtin_popcountll(same);
  const u64 lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  return __builtin_popcountll(same & lt);
}

// digit histogram of each tile of kChunk keys, digit-major: hist[d nt + t]
__global__ __launch_bounds__(kT) void k_rs_hist(const uint32_t *__restrict__ keys, int64_t n,
                                                int shift, uint32_t mask, int nt,
                                                uint32_t *__restrict__ hist) {
  __shared__ uint32_t h[kBins];
  for (int d = threadIdx.x; d < kBins; d += kT) h[d] = 0u;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kChunk;
  for (int r = 0; r < kPer; ++r) {
    const int64_t i = base + (int64_t)r * kT + threadIdx.x;
    const bool ok = i < n;
    const uint32_t d = ok ? (keys[i] >> shift) & mask : 0u;
    // one LDS add per digit present in the wave
    int cnt;
    const int rk = lane_rank(ok, d, 8, cnt);
    if (ok && rk == 0) atomicAdd(&h[d], (uint32_t)cnt);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kBins; d += kT) hist[(int64_t)d * nt + blockIdx.x] = h[d];
}

// stable scatter of one tile by digit: slot = offset of (digit, tile) +
// keys of the digit before it in the tile (earlier rounds, earlier waves of
// the round, earlier lanes of the wave).  vals_in null: the values are the
// slots themselves.
__global__ __launch_bounds__(kT) void k_rs_scatter(
    const uint32_t *__restrict__ keys_in, const uint32_t *__restrict__ vals_in, int64_t n,
    int shift, uint32_t mask, int nbits, int nt, const uint32_t *__restrict__ offs,
    uint32_t *__restrict__ keys_out, uint32_t *__restrict__ vals_out) {
  __shared__ uint32_t base_d[kBins], run_d[kBins];
  __shared__ uint32_t wc[kT / 64][kBins];
  const int t = blockIdx.x, w = threadIdx.x >> 6;
  for (int d = threadIdx.x; d < kBins; d += kT) {
    base_d[d] = offs[(int64_t)d * nt + t];
    run_d[d] = 0u;
  }
  const int64_t base = (int64_t)t * kChunk;
  for (int r = 0; r < kPer; ++r) {
    for (int j = threadIdx.x; j < (kT / 64) * kBins; j += kT) wc[j / kBins][j % kBins] = 0u;
    __syncthreads();
    const int64_t i = base + (int64_t)r * kT + threadIdx.x;
    const bool ok = i < n;
    uint32_t key = 0u, val = 0u, d = 0u;
    if (ok) {
      key = keys_in[i];
      val = vals_in ? vals_in[i] : (uint32_t)i;
      d = (key >> shift) & mask;
    }
    int cnt;
    const int rk = lane_rank(ok, d, nbits, cnt);
    if (ok && rk == 0) wc[w][d] = (uint32_t)cnt;
    __syncthreads();
    // per digit: the waves' starts within the round, then the round's total
    for (int dd = threadIdx.x; dd < kBins; dd += kT) {
      uint32_t s = run_d[dd];
      for (int ww = 0; ww < kT / 64; ++ww) {
        const uint32_t c = wc[ww][dd];
        wc[ww][dd] = s;
        s += c;
      }
      run_d[dd] = s;
    }
    __syncthreads();
    if (ok) {
      const uint32_t pos = base_d[d] + wc[w][d] + (uint32_t)rk;
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();  // wc is cleared for the next round
  }
}

}  // namespace

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
  const int nc = (int)((n + kChunk - 1) / kChunk);
  if (nc > kTopMax)
    return set_error(ctx, PIC_E_CAPACITY, "selection: %lld slots", (long long)n);
  int rc = ensure_prim_tmp(ctx, (size_t)(nc + 1) * sizeof(uint32_t));
  if (rc) return rc;
  uint32_t *sums = (uint32_t *)ctx->prim_tmp;
  k_sel_count<<<(unsigned)nc, kT, 0, ctx->stream>>>(flags, n, sums);
  k_scan_top<<<1, 1024, 0, ctx->stream>>>(sums, nc);
  k_sel_write<<<(unsigned)nc, kT, 0, ctx->stream>>>(flags, n, sums, idx_out);
  PIC_LAUNCH_CHECK(ctx);
  PIC_HIP(ctx, hipMemcpyAsync(ctx->host_counter + 1, sums + nc, sizeof(uint32_t),
                              hipMemcpyDeviceToHost, ctx->stream));
  PIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *count = (int64_t)(*(uint32_t *)(ctx->host_counter + 1));
  return 0;
}

// Stable sort of keys_in[0, n) over bits [0, end_bit): keys_out the sorted
// keys, idx_out[p] the slot the p-th came from.  keys_in is left as it was;
// the temporary buffer holds the passes in between.
int sort_keys(pic_ctx *ctx, uint32_t *keys_in, uint32_t *keys_out,
              uint32_t *idx_out, int64_t n, int end_bit) {
  if (n == 0) return 0;
  const int nt = (int)((n + kChunk - 1) / kChunk);
  const int64_t m = (int64_t)kBins * nt;  // histogram entries
  const int nc = (int)((m + kChunk - 1) / kChunk);
  if (nc > kTopMax || n > ((int64_t)1 << 32) - 1)
    return set_error(ctx, PIC_E_CAPACITY, "sort: %lld keys", (long long)n);
  const int passes = end_bit <= 0 ? 1 : (end_bit + 7) / 8;
  // temporary: the other key and slot buffers, the histogram, the chunk sums
  const size_t bytes = (2 * (size_t)n + (size_t)m + (size_t)nc + 1) * sizeof(uint32_t);
  int rc = ensure_prim_tmp(ctx, bytes);
  if (rc) return rc;
  uint32_t *tkeys = (uint32_t *)ctx->prim_tmp;
  uint32_t *tvals = tkeys + n;
  uint32_t *hist = tvals + n;
  uint32_t *sums = hist + m;
  const uint32_t *kin = keys_in, *vin = nullptr;
  for (int p = 0; p < passes; ++p) {
    const int shift = 8 * p;
    const int nbits = p == passes - 1 && end_bit > 0 ? end_bit - shift : 8;
    const uint32_t mask = (1u << nbits) - 1u;
    // the last pass writes the outputs; earlier ones alternate with them
    const bool to_out = (passes - 1 - p) % 2 == 0;
    uint32_t *kout = to_out ? keys_out : tkeys;
    uint32_t *vout = to_out ? idx_out : tvals;
    k_rs_hist<<<(unsigned)nt, kT, 0, ctx->stream>>>(kin, n, shift, mask, nt, hist);
    PIC_LAUNCH_CHECK(ctx);
    rc = scan_exclusive(ctx, hist, m, sums);
    if (rc) return rc;
    k_rs_scatter<<<(unsigned)nt, kT, 0, ctx->stream>>>(kin, vin, n, shift, mask, nbits, nt,
                                                        hist, kout, vout);
    PIC_LAUNCH_CHECK(ctx);
    kin = kout;
    vin = vout;
  }
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

// Tile sort of the FP32 particle store, shared by the 3D and the 2D3V
// forms.  A destination map KM gives each stored cell index its tile, its
// cell in the tile and its tile's home slot in a 64-entry table:
//   struct KM {
//     __host__ __device__ uint32_t cells() const;  // cells per tile
//     __device__ bool dest(int32_t cell, uint32_t &tile, uint32_t &cin,
//                          int &home) const;       // false: dropped
//   };
// with key = tile * cells() + cin the sort key of the form.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pic_internal.h"
#include "pic_tiles.h"

namespace pic {
namespace tsort {

// The tile sort (sorting :7475-7590 with the tile-major key), written for the
// state a periodic sort meets: the particles are still tile-sorted from the
// last sort except for a small appended tail, and most of them are still in
// their tile.  The result is the stable sort by key, the same permutation
// as a radix sort of (key, slot) pairs:
//   1. each source block (a tile range of the last sort, or a tail chunk)
//      counts the particles still in their tile and writes the others, in
//      slot order, as one block per destination tile into a mover pool
//      shared by all tiles (the block reserved with one atomic), recording
//      for the destination (block's first slot, pool offset, count);
//   2. a scan of (stayers + movers) per tile gives the new tile starts;
//   3. one workgroup per tile orders its mover blocks by first slot, so the
//      sequence (blocks from lower slots, its own stayers, blocks from
//      higher slots) is in slot order, and places that sequence by a stable
//      counting sort over the tile's cells.
// Waves take contiguous quarters of a sequence: each counts, one scan sets
// every (wave, value) start, and each wave then places its quarter in
// order, ranking the lanes that share a value by ballots.  The pool holds
// as many movers as the store holds particles, so it cannot overflow; a
// block with more than kDestMax destination tiles, a tile receiving mover
// blocks from more than kDescCap source blocks, or a tile of more than
// kTileCells cells takes the radix sort instead.
constexpr int kDescCap = 64;
constexpr int kDestMax = 64;
constexpr int kTileCells = 1024;
constexpr uint32_t kEmpty = 0xffffffffu;
// tail particles per source block: the inflow overflow is appended tile by
// tile, a few to some tens per slab tile, so a short block meets few tiles
constexpr int kSortTail = 256;

namespace {

typedef unsigned long long u64;
constexpr int kBlk = 256;

// rank of this lane among the active lanes (pred) of its wave holding the
// same value v (nbits wide), and the number of such lanes
__device__ inline int wave_rank(bool pred, uint32_t v, int nbits, int &count) {
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

__global__ void k_tsort_clear(uint32_t *a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0u;
}

// pass 1: one workgroup per tile range of the last sort or tail chunk
template <class KM>
__global__ __launch_bounds__(256) void k_tsort_split(KM km, TileRun tr,
                                                     const int32_t *__restrict__ cell,
                                                     uint32_t *stay, uint32_t *mcount,
                                                     uint32_t *ndesc, uint4 *desc,
                                                     uint2 *movers, uint32_t pool_cap,
                                                     uint32_t *flag) {
  __shared__ uint32_t tkey[kDestMax];
  __shared__ uint32_t wc[kBlk / 64][kDestMax];
  __shared__ unsigned int kept_w[kBlk / 64];
  const int bid = blockIdx.x, tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const bool own_tile = bid < tr.ntiles;
  // slots fit 32 bits: the library keeps at most 2^30 particles per species
  uint32_t s, e;
  if (own_tile) {
    s = tr.tile_start[bid];
    e = tr.tile_start[bid + 1];
  } else {
    s = (uint32_t)tr.sorted_end + (uint32_t)(bid - tr.ntiles) * kSortTail;
    e = min(s + (uint32_t)kSortTail, (uint32_t)tr.n);
  }
  for (int h = tid; h < kDestMax; h += kBlk) {
    tkey[h] = kEmpty;
    for (int w = 0; w < kBlk / 64; ++w) wc[w][h] = 0u;
  }
  __syncthreads();
  // this wave's contiguous quarter of the block
  const uint32_t per = (e - s + 3) / 4;
  const uint32_t q0 = min(s + wv * per, e), q1 = min(q0 + per, e);
  // the destination's table slot by linear probing from its home slot h,
  // -1 when the table is full (insert) or the tile is absent (find)
  const auto slot_of = [&](uint32_t t, int h, bool insert) -> int {
    for (int i = 0; i < kDestMax; ++i) {
      const uint32_t cur = tkey[h];
      if (cur == t) return h;
      if (cur == kEmpty) {
        if (!insert) return -1;
        const uint32_t old = atomicCAS(&tkey[h], kEmpty, t);
        if (old == kEmpty || old == t) return h;
      }
      h = (h + 1) & (kDestMax - 1);
    }
    return -1;
  };
  // a particle's table slot when it moves to another tile, else -1
  const auto classify = [&](int32_t c, bool insert, uint32_t &t, uint32_t &cin,
                            bool &kept) -> int {
    int h0;
    kept = false;
    if (!km.dest(c, t, cin, h0)) return -1;
    if (own_tile && t == (uint32_t)bid) {
      kept = true;
      return -1;
    }
    const int h = slot_of(t, h0, insert);
    if (h < 0 && insert) atomicOr(flag, 1u);
    return h;
  };
  unsigned int kept = 0;
  int32_t cnext = q0 + lane < q1 ? cell[q0 + lane] : -1;
  for (uint32_t i0 = q0; i0 < q1; i0 += 64) {
    const int32_t c = cnext;
    cnext = i0 + 64 + lane < q1 ? cell[i0 + 64 + lane] : -1;
    uint32_t t = 0, cin = 0;
    bool kp;
    const int h = classify(c, true, t, cin, kp);
    kept += kp;
    // one LDS add per destination present in the wave
    int cnt;
    const int r = wave_rank(h >= 0, (uint32_t)(h >= 0 ? h : 0), 6, cnt);
    if (h >= 0 && r == 0) atomicAdd(&wc[wv][h], (uint32_t)cnt);
  }
  for (int o = 32; o > 0; o >>= 1) kept += __shfl_xor(kept, o);
  if (lane == 0) kept_w[wv] = kept;
  __syncthreads();
  // one reservation per destination: its total, its offset, one descriptor
  if (tid < kDestMax && tkey[tid] != kEmpty) {
    const uint32_t t = tkey[tid];
    uint32_t tot = 0;
    for (int w = 0; w < kBlk / 64; ++w) tot += wc[w][tid];
    (void)atomicAdd(&mcount[t], tot);
    const uint32_t off = atomicAdd(flag + 1, tot);  // the pool's fill
    const uint32_t k = atomicAdd(&ndesc[t], 1u);
    if (k < (uint32_t)kDescCap)
      desc[(int64_t)t * kDescCap + k] = make_uint4(s, off, tot, 0u);
    if (k >= (uint32_t)kDescCap || off + tot > pool_cap) atomicOr(flag, 1u);
    uint32_t run = off;
    for (int w = 0; w < kBlk / 64; ++w) {
      const uint32_t c = wc[w][tid];
      wc[w][tid] = run;
      run += c;
    }
  }
  if (tid == 0 && own_tile) {
    unsigned int k = 0;
    for (int w = 0; w < kBlk / 64; ++w) k += kept_w[w];
    stay[bid] = k;
  }
  __syncthreads();
  // each wave writes its movers in slot order
  cnext = q0 + lane < q1 ? cell[q0 + lane] : -1;
  for (uint32_t i0 = q0; i0 < q1; i0 += 64) {
    const int32_t c = cnext;
    cnext = i0 + 64 + lane < q1 ? cell[i0 + 64 + lane] : -1;
    uint32_t t = 0, cin = 0;
    bool kp;
    const int h = classify(c, false, t, cin, kp);
    const bool mv = h >= 0;
    int cnt;
    const int r = wave_rank(mv, (uint32_t)(mv ? h : 0), 6, cnt);
    if (mv) {
      const uint32_t pos = wc[wv][h] + (uint32_t)r;
      if (pos < pool_cap) movers[pos] = make_uint2(i0 + (uint32_t)lane, cin);
    }
    // the lanes' reads above precede this update in the wave's LDS order
    if (mv && r == 0) wc[wv][h] += (uint32_t)cnt;
  }
}

// pass 2: start[t] = exclusive scan of stay[t] + mcount[t] over the tiles,
// start[nt] = the kept total (one workgroup of 1024 threads)
__global__ __launch_bounds__(1024) void k_tsort_scan(const uint32_t *stay,
                                                     const uint32_t *mcount, int nt,
                                                     uint32_t *start) {
  __shared__ uint32_t part[1024];
  const int per = (nt + 1023) / 1024;
  const int a = threadIdx.x * per, b = min(a + per, nt);
  uint32_t sum = 0;
  for (int i = a; i < b; ++i) sum += stay[i] + mcount[i];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {  // inclusive Hillis-Steele scan
    const uint32_t v = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = threadIdx.x > 0 ? part[threadIdx.x - 1] : 0u;
  for (int i = a; i < b; ++i) {
    start[i] = run;
    run += stay[i] + mcount[i];
  }
  if (threadIdx.x == 1023) start[nt] = part[1023];
}

// pass 3: one workgroup per tile t; writes idx[start[t] + position] = slot.
// Dynamic LDS: per-wave cell counters (4 x cells per tile) and the cells'
// totals (cells per tile), 20 bytes per cell.
template <class KM>
__global__ __launch_bounds__(256) void k_tsort_place(KM km,
                                                     const uint32_t *__restrict__ old_start,
                                                     const int32_t *__restrict__ cell,
                                                     const uint32_t *__restrict__ ndesc,
                                                     const uint4 *__restrict__ desc,
                                                     const uint2 *__restrict__ movers,
                                                     uint32_t pool_cap,
                                                     const uint32_t *__restrict__ start,
                                                     uint32_t *idx) {
  // the sequence as segments: first sequence index and source offset; the
  // segment s_own is the tile's own range from slot s0, the others are
  // mover blocks at that offset of the tile's list
  __shared__ uint32_t seg_q[kDescCap + 2], seg_src[kDescCap + 2];
  __shared__ int s_own;
  __shared__ uint4 d[kDescCap];
  __shared__ uint32_t part[kBlk];
  extern __shared__ uint32_t place_lds[];
  const int t = blockIdx.x, tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const uint32_t cpt = km.cells();
  uint32_t *cnt = place_lds + wv * cpt;  // this wave's counters
  uint32_t *tot = place_lds + 4 * cpt;
  const uint32_t s0 = old_start[t], len = old_start[t + 1] - s0;
  const int nd = (int)min(ndesc[t], (uint32_t)kDescCap);
  if (wv == 0) {
    // the mover blocks in order of first slot: the first slots differ, so
    // each block's rank is the number of blocks with a smaller one
    uint4 x = make_uint4(kEmpty, 0u, 0u, 0u);
    if (lane < nd) x = desc[(int64_t)t * kDescCap + lane];
    int rank = 0;
    for (int j = 0; j < nd; ++j) rank += (uint32_t)__builtin_amdgcn_readlane((int)x.x, j) < x.x;
    if (lane < nd) d[rank] = x;
  }
  for (uint32_t j = tid; j < 4 * cpt; j += kBlk) place_lds[j] = 0u;
  __syncthreads();
  if (wv == 0) {
    // the sequence: blocks from lower slots, the own range, blocks from
    // higher slots.  A block past the list's end overflowed: the result is
    // then discarded, but its clamped length keeps every read in the list.
    const uint4 y = lane < nd ? d[lane] : make_uint4(kEmpty, 0u, 0u, 0u);
    const uint32_t off = min(y.y, pool_cap);
    const uint32_t ln = lane < nd ? min(y.z, pool_cap - off) : 0u;
    const int own = __builtin_popcountll(__builtin_amdgcn_ballot_w64(lane < nd && y.x < s0));
    uint32_t inc = ln;  // inclusive scan over the wave
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t v = __shfl_up(inc, o);
      if (lane >= o) inc += v;
    }
    const uint32_t below = own > 0 ? (uint32_t)__builtin_amdgcn_readlane((int)inc, own - 1) : 0u;
    const uint32_t sum = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
    if (lane < nd) {
      const int si = lane < own ? lane : lane + 1;
      seg_q[si] = inc - ln + (lane < own ? 0u : len);
      seg_src[si] = y.y;
    }
    if (lane == 0) {
      seg_q[own] = below;
      seg_src[own] = s0;
      seg_q[nd + 1] = sum + len;  // end of the sequence
      s_own = own;
    }
  }
  __syncthreads();
  const int ns = nd + 1, own = s_own;
  const uint32_t total = seg_q[ns];
  const uint32_t per = (total + 3) / 4;
  const uint32_t q0 = min(wv * per, total), q1 = min(q0 + per, total);
  // this lane's segment: the last one starting at or before q, found once
  // by bisection, then followed as q grows
  const auto seg_find = [&](uint32_t q) -> int {
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (seg_q[mid] <= q) lo = mid; else hi = mid - 1;
    }
    return lo;
  };
  const int sg0 = seg_find(q0 + lane);
  int sg = sg0;
  // element q of the sequence: (slot, cell in the tile), false for a range
  // slot whose particle left the tile or dropped out
  const auto elem = [&](uint32_t q, uint32_t &slot, uint32_t &c) -> bool {
    while (sg + 1 < ns && seg_q[sg + 1] <= q) ++sg;
    const uint32_t r = q - seg_q[sg];
    if (sg != own) {
      const uint2 m = movers[seg_src[sg] + r];
      slot = m.x;
      c = m.y;
      return true;
    }
    slot = s0 + r;
    uint32_t tt;
    int h;
    return km.dest(cell[slot], tt, c, h) && tt == (uint32_t)t;
  };
  for (uint32_t qb = q0; qb < q1; qb += 64) {
    const uint32_t q = qb + lane;
    uint32_t slot, c = 0;
    const bool ok = q < q1 && elem(q, slot, c);
    // lanes in runs of one cell: the run's first lane adds the run's length
    const uint32_t cv = ok ? c : kEmpty;
    const uint32_t cp = __shfl_up(cv, 1);
    const bool first = lane == 0 || cv != cp;
    const u64 heads = __builtin_amdgcn_ballot_w64(first);
    if (ok && first) {
      const u64 above = (heads >> lane) >> 1;
      const uint32_t run = above ? (uint32_t)__builtin_ctzll(above) + 1u : 64u - lane;
      atomicAdd(&cnt[c], run);
    }
  }
  __syncthreads();
  // cell starts (exclusive scan over the cells' totals), then each wave's
  // start per cell
  for (uint32_t c = tid; c < cpt; c += kBlk) {
    uint32_t v = 0;
    for (int w = 0; w < kBlk / 64; ++w) v += place_lds[w * cpt + c];
    tot[c] = v;
  }
  __syncthreads();
  const uint32_t cper = (cpt + kBlk - 1) / kBlk;
  const uint32_t ca = min(tid * cper, cpt), cb = min(ca + cper, cpt);
  uint32_t sum = 0;
  for (uint32_t c = ca; c < cb; ++c) sum += tot[c];
  part[tid] = sum;
  __syncthreads();
  for (int o = 1; o < kBlk; o <<= 1) {  // inclusive Hillis-Steele scan
    const uint32_t v = tid >= o ? part[tid - o] : 0u;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  uint32_t run = start[t] + (tid > 0 ? part[tid - 1] : 0u);
  for (uint32_t c = ca; c < cb; ++c) {
    for (int w = 0; w < kBlk / 64; ++w) {
      const uint32_t v = place_lds[w * cpt + c];
      place_lds[w * cpt + c] = run;
      run += v;
    }
  }
  __syncthreads();
  int nbits = 1;
  while ((1u << nbits) < cpt) ++nbits;
  sg = sg0;
  for (uint32_t qb = q0; qb < q1; qb += 64) {
    const uint32_t q = qb + lane;
    uint32_t slot = 0, c = 0;
    const bool ok = q < q1 && elem(q, slot, c);
    int n;
    const int r = wave_rank(ok, c, nbits, n);
    if (ok) idx[cnt[c] + (uint32_t)r] = slot;
    if (ok && r == 0) cnt[c] += (uint32_t)n;
  }
}

}  // namespace

// the sort's device lists for nt tiles and a mover pool of `pool` entries
struct Bufs {
  uint32_t *stay = nullptr, *mcount = nullptr, *ndesc = nullptr;
  uint4 *desc = nullptr;
  uint2 *movers = nullptr;
  uint32_t *start = nullptr;
  uint32_t *flag = nullptr;  // the overflow flag, then the mover pool's fill
  int ntiles = 0;
  int64_t pool = 0;
};

inline void release(Bufs &B) {
  for (void *p : {(void *)B.stay, (void *)B.mcount, (void *)B.desc, (void *)B.movers,
                  (void *)B.start, (void *)B.flag})
    if (p) (void)hipFree(p);
  B = Bufs();
}

inline int reserve(pic_ctx *rt, Bufs &B, int nt, int64_t pool) {
  if (B.ntiles >= nt && B.pool >= pool) return 0;
  release(B);
  PIC_HIP(rt, hipMalloc(&B.stay, (size_t)nt * sizeof(uint32_t)));
  PIC_HIP(rt, hipMalloc(&B.mcount, (size_t)2 * nt * sizeof(uint32_t)));
  B.ndesc = B.mcount + nt;
  PIC_HIP(rt, hipMalloc(&B.desc, (size_t)nt * kDescCap * sizeof(uint4)));
  PIC_HIP(rt, hipMalloc(&B.movers, (size_t)pool * sizeof(uint2)));
  PIC_HIP(rt, hipMalloc(&B.start, (size_t)(nt + 1) * sizeof(uint32_t)));
  PIC_HIP(rt, hipMalloc(&B.flag, 2 * sizeof(uint32_t)));
  B.ntiles = nt;
  B.pool = pool;
  return 0;
}

// The tile sort of one species when the tiles of the last sort are current
// (old_start: nt + 1 tile starts, [0, sorted_end) tile-sorted, the tail
// after it appended; n stored slots, at most 2^30).  *done is false when a
// source block met more than kDestMax destination tiles or a tile more than
// kDescCap source blocks (the caller then takes the radix sort).  Otherwise
// idx[0, *valid) holds the permutation (new slot p takes old slot idx[p]),
// old_start the new tile starts, *valid the kept count.
template <class KM>
int run(pic_ctx *rt, const KM &km, int nt, uint32_t *old_start, int64_t sorted_end,
        int64_t n, const int32_t *cell, Bufs &B, uint32_t *idx, bool *done,
        int64_t *valid) {
  *done = false;
  if (B.ntiles < nt || B.pool < n)
    return set_error(rt, PIC_E_STATE, "tile sort: lists for %d tiles / %lld movers, need %d / %lld",
                     B.ntiles, (long long)B.pool, nt, (long long)n);
  hipStream_t st = rt->stream;
  TileRun tr;
  tr.tile_start = old_start;
  tr.sorted_end = sorted_end;
  tr.ntiles = nt;
  tr.n = n;
  tr.ntail = (int)div_up(n - sorted_end, kSortTail);
  // mover counts and descriptor counts (adjacent), the flag and the fill
  k_tsort_clear<<<(unsigned)div_up(2 * nt, kBlk), kBlk, 0, st>>>(B.mcount, 2 * nt);
  k_tsort_clear<<<1, 64, 0, st>>>(B.flag, 2);
  k_tsort_split<KM><<<(unsigned)(nt + tr.ntail), kBlk, 0, st>>>(
      km, tr, cell, B.stay, B.mcount, B.ndesc, B.desc, B.movers, (uint32_t)B.pool, B.flag);
  k_tsort_scan<<<1, 1024, 0, st>>>(B.stay, B.mcount, nt, B.start);
  const size_t place_lds = (size_t)5 * km.cells() * sizeof(uint32_t);
  k_tsort_place<KM><<<(unsigned)nt, kBlk, place_lds, st>>>(
      km, old_start, cell, B.ndesc, B.desc, B.movers, (uint32_t)B.pool, B.start, idx);
  PIC_LAUNCH_CHECK(rt);
  uint32_t *hc = (uint32_t *)rt->host_counter;
  PIC_HIP(rt, hipMemcpyAsync(hc, B.start + nt, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  PIC_HIP(rt, hipMemcpyAsync(hc + 1, B.flag, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  PIC_HIP(rt, hipStreamSynchronize(st));
  if (hc[1] != 0u) return 0;  // a table overflowed: radix sort
  PIC_HIP(rt, hipMemcpyAsync(old_start, B.start, (size_t)(nt + 1) * sizeof(uint32_t),
                             hipMemcpyDeviceToDevice, st));
  *valid = (int64_t)hc[0];
  *done = true;
  return 0;
}

}  // namespace tsort
}  // namespace pic

// Tile-window machinery shared by the FP64 and FP32 particle passes.
//
// One workgroup takes one sort tile (its particles are contiguous after the
// sort) or one chunk of the unsorted inflow tail.  It accumulates current
// and density in an LDS window around an anchor box: the tile's cells, or
// the bounding box of a few sampled particles of a tail chunk.  The window
// spans the anchor +- kMargin cells; a particle whose cell at the start of
// the step lies within the anchor +- (kMargin - 2) has all its gather and
// deposit nodes inside the window (it moves less than one cell per step and
// its stencils reach one node below and two above its cell).
#pragma once
#include <hip/hip_runtime.h>

#include "pic_internal.h"

namespace pic {

constexpr int kMargin = 3;
constexpr int kTailChunk = 2048;
constexpr int kStray = 8;

struct TileRun {
  const uint32_t *tile_start;  // ntiles + 1 entries
  int ntiles;
  int64_t sorted_end;          // [0, sorted_end) is tile-sorted
  int64_t n;                   // stored slots
  int ntail;                   // tail chunks of kTailChunk
};

// LDS window: origin (cells) and extents (nodes), and the cell range whose
// stencils fit inside it
struct Win {
  int k0, l0, m0, WX, WY, WZ;
  int vk0, vk1, vl0, vl1, vm0, vm1;
  bool on;
};

// Between sorts the particles move with the bulk drift (vdrift) and leave
// their tile on its downstream side; those more than one cell out deposit
// through global atomics.  The tiles' windows follow the drift: after
// `halves` half-step drifts since the sort they are moved by the bulk
// displacement rounded to whole cells.
inline void add_drift_shift(TileGeom &t, const pic_config &c, int halves) {
  for (int a = 0; a < 3; ++a)
    t.shift[a] = (int)lround(c.vdrift[a] * 0.5 * c.dt * halves / c.dx);
}

__device__ inline int xcd_remap(int b, int nt, int nb) {
  // Workgroups are dealt round-robin over the 8 XCDs in launch order.  Each
  // XCD gets a contiguous range of the nt tiles (neighbouring tiles share its
  // L2) and then an equal share of the nb - nt tail chunks, so the light
  // chunk workgroups do not displace tiles onto the other XCDs.
  const int nc = nb - nt;
  const int qt = nt / 8, qc = nc / 8, q = qt + qc;
  if (b >= 8 * q) {  // remainders of both kinds, in order
    const int r = b - 8 * q, rt = nt - 8 * qt;
    return r < rt ? 8 * qt + r : nt + 8 * qc + (r - rt);
  }
  const int x = b % 8, s = b / 8;
  return s < qt ? x * qt + s : nt + x * qc + (s - qt);
}

// Range and window of workgroup bid.  cell_of(i, k, l, m) returns false for
// a dead slot.  Runs on one thread.
template <typename CellOf>
__device__ inline void tile_setup(int bid, const TileRun &tr, const TileGeom &tg,
                                  int win_max, CellOf cell_of, int64_t &s,
                                  int64_t &e, Win &w) {
  int ak0, ak1, al0, al1, am0, am1;  // anchor box (cells)
  if (bid < tr.ntiles) {
    s = tr.tile_start[bid];
    e = tr.tile_start[bid + 1];
    const int tz = bid % tg.NTZ, ty = (bid / tg.NTZ) % tg.NTY,
              tx = bid / (tg.NTZ * tg.NTY);
    ak0 = 2 + tx * tg.TX + tg.shift[0];
    ak1 = ak0 + tg.TX - 1;
    al0 = 2 + ty * tg.TY + tg.shift[1];
    al1 = al0 + tg.TY - 1;
    am0 = 2 + tz * tg.TZ + tg.shift[2];
    am1 = am0 + tg.TZ - 1;
  } else {
    const int c = bid - tr.ntiles;
    s = tr.sorted_end + (int64_t)c * kTailChunk;
    e = s + kTailChunk < tr.n ? s + kTailChunk : tr.n;
    ak0 = al0 = am0 = 1 << 20;
    ak1 = al1 = am1 = -1;
    // 8 samples; those more than kStray cells from a middle sample on some
    // axis (a particle that wrapped across a periodic face, a straggler) are
    // left out of the box and deposit globally
    int ks[8], ls[8], ms[8];
    bool ok[8];
    for (int j = 0; j < 8; ++j) {
      const int64_t i = s + (e - s - 1) * j / 7;
      ok[j] = i >= s && i < e && cell_of(i, ks[j], ls[j], ms[j]);
    }
    int r = -1;
    for (int q = 0; q < 8 && r < 0; ++q) {
      const int j = (q & 1) ? 3 - (q >> 1) : 4 + (q >> 1);  // 4 3 5 2 6 1 7 0
      if (ok[j]) r = j;
    }
    for (int j = 0; j < 8 && r >= 0; ++j) {
      if (!ok[j] || abs(ks[j] - ks[r]) > kStray || abs(ls[j] - ls[r]) > kStray ||
          abs(ms[j] - ms[r]) > kStray)
        continue;
      ak0 = min(ak0, ks[j]); ak1 = max(ak1, ks[j]);
      al0 = min(al0, ls[j]); al1 = max(al1, ls[j]);
      am0 = min(am0, ms[j]); am1 = max(am1, ms[j]);
    }
  }
  w.on = ak1 >= ak0;
  if (w.on) {
    w.k0 = ak0 - kMargin;
    w.l0 = al0 - kMargin;
    w.m0 = am0 - kMargin;
    w.WX = ak1 - ak0 + 1 + 2 * kMargin;
    w.WY = al1 - al0 + 1 + 2 * kMargin;
    w.WZ = am1 - am0 + 1 + 2 * kMargin;
    w.vk0 = ak0 - (kMargin - 2);
    w.vk1 = ak1 + (kMargin - 2);
    w.vl0 = al0 - (kMargin - 2);
    w.vl1 = al1 + (kMargin - 2);
    w.vm0 = am0 - (kMargin - 2);
    w.vm1 = am1 + (kMargin - 2);
    if ((int64_t)w.WX * w.WY * w.WZ > win_max) w.on = false;
  }
}

__device__ inline bool in_win(const Win &w, int k, int l, int m) {
  return w.on && k >= w.vk0 && k <= w.vk1 && l >= w.vl0 && l <= w.vl1 &&
         m >= w.vm0 && m <= w.vm1;
}

}  // namespace pic

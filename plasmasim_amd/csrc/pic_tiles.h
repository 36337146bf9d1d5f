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

__device__ inline int xcd_remap(int b, int nb) {
  // consecutive workgroups are dealt round-robin over the 8 XCDs; give each
  // XCD a contiguous range of tiles so neighbouring tiles share its L2
  const int q = nb / 8;
  if (b >= 8 * q) return b;
  return (b % 8) * q + b / 8;
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
    ak0 = 2 + tx * tg.TX;
    ak1 = ak0 + tg.TX - 1;
    al0 = 2 + ty * tg.TY;
    al1 = al0 + tg.TY - 1;
    am0 = 2 + tz * tg.TZ;
    am1 = am0 + tg.TZ - 1;
  } else {
    const int c = bid - tr.ntiles;
    s = tr.sorted_end + (int64_t)c * kTailChunk;
    e = s + kTailChunk < tr.n ? s + kTailChunk : tr.n;
    ak0 = al0 = am0 = 1 << 20;
    ak1 = al1 = am1 = -1;
    for (int j = 0; j < 8; ++j) {
      const int64_t i = s + (e - s - 1) * j / 7;
      if (i < s || i >= e) continue;
      int k, l, m;
      if (!cell_of(i, k, l, m)) continue;
      ak0 = min(ak0, k); ak1 = max(ak1, k);
      al0 = min(al0, l); al1 = max(al1, l);
      am0 = min(am0, m); am1 = max(am1, m);
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

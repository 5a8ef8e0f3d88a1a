// rcg_tiled.h -- device-side state of a tiled FP32 gridding plan.
#pragma once

#include "rcg_internal.h"

namespace rcg {

// Tiles per grid side^2 handled by the LDS histograms of the binning pass
// (16384 ints = 64 KiB of LDS).
constexpr int MAX_BIN_TILES = 16384;
// Error word bit: a pair did not fit the plan's pair buffer (a chunk
// support above the plan's max_support).
constexpr int ERR_CAPACITY = 2;

struct TiledDev {
  int T, lt, ntx, n_tiles, chunk, pitch, grid_size;
  long long pair_cap;
  int *tile_count;   // pairs per tile
  int *tile_start;   // first pair of each tile
  int *tile_cursor;  // fill cursors
  int *item_start;   // first work item of each tile
  int *totals;       // [0] pairs, [1] items
  int4 *pairs;       // (point, tile) records sorted by tile
  float2 *partial;   // degridding: per (point, tile slot) partial sums
  int mpp;           // tile slots per point (max pairs per point)
};

}  // namespace rcg

struct rcg_tiled_plan {
  rcg_geom geom;
  long long n;
  int max_support;
  int max_pairs_pt;
  rcg_pg *pg = nullptr;
  unsigned long long *taps = nullptr;
  int *err = nullptr;
  rcg::TiledDev dev{};
};

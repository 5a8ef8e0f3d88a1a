// rcg_tiled.h -- device-side state of a tiled FP32 gridding plan.
#pragma once

#include <vector>

#include "rcg_internal.h"

namespace rcg {

// Tiles per instance handled by the LDS histograms of the binning pass
// (16384 ints = 64 KiB of LDS).
constexpr int MAX_BIN_TILES = 16384;
// Tiles of a whole plan (instances x tiles per instance) handled by the
// three-kernel scan: 1024 segments of 1024 tiles.
constexpr int MAX_PLAN_TILES = 1 << 20;
// Error word bit: a pair did not fit the plan's pair buffer (a chunk
// support above the plan's max_support).
constexpr int ERR_CAPACITY = 2;
// Work-queue counters: [0, 8) gridding, [8, 16) degridding, one per XCD;
// [16] the halo item sort.
constexpr int N_QUEUE = 17;

// Records per work item of a halo-layout plan (its kernels sort an item's
// records in LDS).
constexpr int MAX_HALO_ITEM = 1024;

// Record layouts of a plan (rcg_tiled_plan_create_ex).
constexpr int LAYOUT_PAIRS = 0;  // one record per (point, tile) the window overlaps
constexpr int LAYOUT_HALO = 1;   // one record per sub-window, in its origin's tile

struct TiledDev {
  int layout;        // LAYOUT_PAIRS or LAYOUT_HALO
  int wmax, lw;      // halo: largest sub-window side (16 or 32) and its log2
  int gbins;         // binning counts and ranks with global atomics (tiles
                     // per instance above the LDS histogram)
  int T, lt, ntx;
  int tpi;           // tiles per instance (ntx * ntx)
  int n_inst;
  int n_tiles;       // n_inst * tpi
  int chunk, pitch, grid_size;
  int mpp;           // tile slots per point (max pairs per point)
  int ppt;           // points per binning thread
  long long gstride; // grid elements per instance (grid_size * pitch)
  long long pair_cap;
  long long npts;    // points of the plan (all instances)
  double inv_pitch;  // 1 / pitch
  int *tile_count;   // pairs per tile
  int *tile_start;   // first pair of each tile
  int *tile_cursor;  // fill cursors
  int *item_start;   // first work item of each tile
  int *totals;       // [0] pairs, [1] items
  int *seg;          // scan scratch: 2 ints per segment of 1024 tiles
  int *queue;        // N_QUEUE work counters
  const int *pt_off;   // n_inst + 1: first point of each instance
  const int *blk_off;  // n_inst + 1: first binning workgroup of each instance
  int4 *pairs;       // (point, tile) records sorted by tile
  float2 *partial;   // degridding: partial sums, [tile slot][point]
  int adj;           // gridding with conjugated taps (set per launch)
  float2 *lo;        // compensated flush: rounding errors of the atomic adds
  float2 *ibuf;      // halo gridding: each work item's region, items below
  int ibuf_cap;      // ibuf_cap (later items flush with atomics)
  int item_cap;      // halo: bound on the number of work items
  uint16_t *runs;    // halo: per work item, its records per run key (64)
  float2 *vsort;     // halo gridding: the visibility of every record, in record order
};

// Event pairs kept per direction for rcg_tiled_kernel_times.
constexpr int N_TIMED = 64;

}  // namespace rcg

struct rcg_tiled_plan;

namespace rcg {

// Bins the records the caller wrote into the plan's record buffer
// (rcg_tiled_plan::pg, one per point; cell -1 for a point left out) in place
// of the pregrid that rcg_tiled_bin runs.
rcg_status tiled_bin_records(rcg_tiled_plan *pl, hipStream_t st);

// Halo-layout kernels (rcg_halo.hip): persistent workgroups per CU x CUs
// for the grid and degrid kernels of tile T and sub-windows of at most
// wmax cells, and their launches.
rcg_status halo_blocks(int T, int wmax, int *grid_blocks, int *degrid_blocks);
rcg_status halo_grid(const TiledDev &d, int blocks, const float2 *vis, const float2 *taps,
                     long long n_taps, float2 *grid, hipStream_t st);
// Cells of a tile's region (T + wmax - 1)^2: the item-buffer stride.
inline long long halo_region_cells(int T, int wmax)
{
  return (long long)(T + wmax - 1) * (T + wmax - 1);
}
rcg_status halo_degrid(const TiledDev &d, int blocks, const float2 *grid, const float2 *taps,
                       long long n_taps, float2 *vis, hipStream_t st);
// Sorts every work item's records by run key in place and writes the
// items' run tables (the last pass of binning a halo-layout plan).
rcg_status halo_sort_items(const TiledDev &d, hipStream_t st);

// FP64 gridding over the plan's pairs, added to the grid's contents.
rcg_status tiled_grid_f64_add(rcg_tiled_plan *pl, const double2 *taps, long long n_taps,
                              const double2 *vis, double2 *grid, hipStream_t st);

}  // namespace rcg

struct rcg_tiled_plan {
  std::vector<rcg_geom> geoms;
  rcg_geom *geoms_dev = nullptr;
  long long n = 0;
  int max_support = 0;
  int max_pairs_pt = 0;
  int grid_waves = 0;     // wavefronts (private tile copies) per gridding workgroup
  int degrid_waves = 4;   // wavefronts per degridding workgroup (sharing the staged tile)
  int grid_blocks = 0;    // persistent workgroups of the grid kernel
  int degrid_blocks = 0;  // ... of the degrid kernel
  int bin_blocks = 0;
  rcg_pg *pg = nullptr;
  unsigned long long *taps = nullptr;
  int *err = nullptr;
  int *offs = nullptr;    // device pt_off | blk_off
  float2 *lo = nullptr;   // compensated flush grid (rcg_tiled_compensated)
  bool compensated = false;
  rcg::TiledDev dev{};
  // kernel timing (rcg_tiled_timing)
  bool timing = false;
  hipEvent_t ev[2][rcg::N_TIMED][2] = {};
  int n_ev[2] = {0, 0};
};

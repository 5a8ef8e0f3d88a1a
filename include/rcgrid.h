/*
 * rcgrid.h -- C ABI of the MI355X (gfx950) w-projection gridder/degridder.
 *
 * Two layers:
 *
 *  1. The reference's MS4 kernel entry points, same names, argument order and
 *     meaning, all buffers in HOST memory, so that the Haskell DNA layer
 *     (MS4/dna-programs/Kernel/CPU/ScatterGrid.hs:48-51,87) and the C++
 *     driver (MS4/misc/cppcycle.cpp:210-292) link against this library
 *     unchanged.  Each call copies its inputs to the GPU, runs the HIP
 *     kernels and copies the result back; nothing persists between calls.
 *     Types: rcg_c128 and rcg_double3 have the layout of the reference's
 *     complexd (std::complex<double>) and Double3 (MS4/kernel/common/
 *     common.h:37-39,72-77); every one of them is passed by pointer.
 *
 *  2. A device-resident API (rcg_*) on caller-owned device buffers and a
 *     caller-chosen HIP stream, used by the Python binding, the benchmark
 *     and the batched/FP32 paths.
 *
 * Errors: the MS4 functions have no error path in the reference.  Here a
 * failing call returns 0 taps (or does nothing) and leaves the output zeroed
 * where it can; rcg_last_error() / rcg_last_error_string() report the cause
 * (thread-local).  All rcg_* functions return an rcg_status.
 */
#ifndef RCGRID_H
#define RCGRID_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef unsigned long long ull;

typedef struct rcg_c128 { double re, im; } rcg_c128;
typedef struct rcg_c64 { float re, im; } rcg_c64;
typedef struct rcg_double3 { double u, v, w; } rcg_double3;

typedef enum rcg_status {
  RCG_OK = 0,
  RCG_ERR_ARG = 1,        /* invalid argument (sizes, pointers, supports) */
  RCG_ERR_HIP = 2,        /* a HIP runtime call failed                     */
  RCG_ERR_NOMEM = 3,      /* device or host allocation failed              */
  RCG_ERR_RANGE = 4,      /* a point referenced a w-plane or GCF layer
                             outside the table, or a GCF window that does
                             not fit its layer                              */
  RCG_ERR_UNSUPPORTED = 5,
  RCG_ERR_FFT = 6         /* rocFFT plan or execution failed               */
} rcg_status;

/* ---------------------------------------------------------------------
 * 1. Reference MS4 entry points (host memory)
 * ------------------------------------------------------------------- */

/* Replaces gridKernelCPUFullGCF / gridKernelCPUHalfGCF,
 * MS4/kernel/common/scatter_gridder_w_dependent_dyn_1p.h:16-34 (bodies
 * MS4/kernel/cpu/scatter_grid/scatter_gridder_w_dependent_dyn_1p.cpp:
 * 245-266).  Zeroes grid (grid_size rows x grid_pitch complex), grids
 * baselines*ts_ch points; uvw[bl], vis[bl] point at ts_ch consecutive
 * points of chunk bl whose support is bl_supps[bl].  gcf and gcf_supps point
 * at the w_plane = 0 centre of the layer-pointer and support tables
 * (negative indices allowed).  Returns the number of in-bound taps.
 * Accumulation is FP64 (result within 1e-12 relative L2 of the reference). */
ull gridKernelCPUFullGCF(double scale, double wstep, int baselines,
                         const int bl_supps[], rcg_c128 grid[],
                         const rcg_c128 *gcf[], const rcg_double3 *uvw[],
                         const rcg_c128 *vis[], int ts_ch, int grid_pitch,
                         int grid_size, int gcf_supps[]);
ull gridKernelCPUHalfGCF(double scale, double wstep, int baselines,
                         const int bl_supps[], rcg_c128 grid[],
                         const rcg_c128 *gcf[], const rcg_double3 *uvw[],
                         const rcg_c128 *vis[], int ts_ch, int grid_pitch,
                         int grid_size, int gcf_supps[]);

/* Replaces deGridKernelCPUFullGCF / deGridKernelCPUHalfGCF,
 * scatter_gridder_w_dependent_dyn_1p.h:73-91 (bodies .cpp:368-431).
 * Overwrites every vis[bl][i] with the transpose-degridded value
 * (unconjugated taps; out-of-bound points become 0).  Returns taps. */
ull deGridKernelCPUFullGCF(double scale, double wstep, int baselines,
                           const int bl_supps[], const rcg_c128 grid[],
                           const rcg_c128 *gcf[], const rcg_double3 *uvw[],
                           rcg_c128 *vis[], int ts_ch, int grid_pitch,
                           int grid_size, int gcf_supps[]);
ull deGridKernelCPUHalfGCF(double scale, double wstep, int baselines,
                           const int bl_supps[], const rcg_c128 grid[],
                           const rcg_c128 *gcf[], const rcg_double3 *uvw[],
                           rcg_c128 *vis[], int ts_ch, int grid_pitch,
                           int grid_size, int gcf_supps[]);

/* Replaces grid0, scatter_gridder_w_dependent_dyn_1p.h:36-45 (.cpp:268-291):
 * nearest-cell gridding into a grid pointer the caller has centred. */
void grid0(const rcg_double3 uvw[], const rcg_c128 vis[], rcg_c128 grid[],
           double scale, int baselines_ts_ch, int grid_pitch, int grid_size);

/* Replaces normalizeCPU, .h:47-52 (.cpp:293-306): grid *= 1/(size*pitch). */
void normalizeCPU(rcg_c128 src[], int grid_pitch, int grid_size);

/* Replaces reweight, .h:54-61 (.cpp:311-353): uniform weighting. */
void reweight(const rcg_double3 uvw[], rcg_c128 vis[], double scale,
              int baselines_ts_ch, int grid_size);

/* Replaces rotateCPU, .h:63-69 (.cpp:355-363): vis *= exp(2 pi i w scale),
 * evaluated in FP64. */
void rotateCPU(const rcg_double3 uvw[], rcg_c128 vis[], int baselines_ts_ch,
               double scale);

/* Replaces herm_padded_inplace, MS4/kernel/common/herm_padded.h:9
 * (MS4/kernel/cpu/herm/herm_padded.cpp:6-27). */
void herm_padded_inplace(rcg_c128 *data, int size, int pitch);

/* Replaces fft_inplace_even / fftInitThreading,
 * MS4/kernel/common/fft_dyn_padded.h:10-11 (MS4/kernel/cpu/fft/
 * fft_dyn_padded.cpp:77-127).  sign: -1 forward C2C, +1 backward C2C,
 * 0 R2C, 2 C2R, with the reference's fftshift before and after and its
 * in-place strides.  Returns an opaque plan handle (the argument plan may be
 * NULL); release it with rcg_fft_destroy_plan (the reference's Haskell
 * binding calls fftw_destroy_plan, MS4/dna-programs/Kernel/CPU/FFT.hs:23). */
void *fft_inplace_even(void *plan, int sign, void *data, int size, int pitch);
void fftInitThreading(void);
void rcg_fft_destroy_plan(void *plan);

/* Replaces deconvolve, MS4/kernel/common/hogbom.h:7-16
 * (MS4/kernel/cpu/hogbom/hogbom.cpp:94-121) with textbook Hogbom CLEAN:
 * signed peak, model += gain*peak/psf_peak, stop when |peak| < threshold. */
void deconvolve(double *mod_p, double *res_p, double *psf_p, int siz,
                int pitch, unsigned int niters, double gain, double threshold);

/* Replaces mkGCFLayer (MS4/kernel/cpu/gcf/GCF.cpp:123-146, FFI import
 * MS4/dna-programs/Kernel/CPU/GCF.hs:15-25): the 64 oversampled layers
 * [over_u][over_v][support][support] of one w-plane for phase w (scaled)
 * and theta/2 = t2 into host dst, table[i] = dst + i*support^2.  p is the
 * handle returned by the previous call (NULL the first time), reused while
 * max_support and src_pad stay the same; `arena` is not used (the work
 * buffer is on the device).  Release with rcg_gcf_gen_destroy.  The
 * normalisation sums are pad-correct unless rcg_gcf_pad_quirk(1) selects
 * the reference's contiguous walk (thread-local switch). */
void *mkGCFLayer(void *p, rcg_c128 dst[], rcg_c128 *table[], rcg_c128 arena[],
                 int support, int max_support, int src_pad, double t2,
                 double w);
void rcg_gcf_pad_quirk(int enable);
/* Replaces calcAccums (GCF.cpp:148-219): per w-plane sum of w and point
 * count over numDataPoints points, plane = round(w / wstep) + numOfWPlanes/2.
 * Points outside the planes are skipped and reported (RCG_ERR_RANGE); the
 * reference writes outside its arrays for them. */
void calcAccums(const rcg_double3 uvw[], double sums[], int npts[],
                double wstep, int numDataPoints, int numOfWPlanes);

/* Replace bin / binm, MS4/kernel/common/binner.h:8-9 (MS4/kernel/cpu/binner/
 * binner.cpp:100-189; called from MS4/dna-programs/Kernel/GPU/
 * GatherGrid.hs:20-44).  Bins num_points points into 32 x 32 uv bins of
 * grid_size/32 cells: bin (p.u / ds, p.v / ds) of the untranslated pregrid
 * cell (C truncating division), plus copies in the neighbour bins that the
 * margin supp/2 reaches, supp = gcfsupps[w_plane] (centred table).  Each copy
 * carries the 8-byte Pregridded record (u, v translated by
 * grid_size/2 - supp/2, layer index, supp) and the amplitude times
 * exp(2 pi i w scale); binm negates the layer index for w < 0.  Each bin
 * keeps its copies in input order.  The result is one pinned host block laid
 * out as the reference's: an rcg_bin_table (RCG_BIN_TABLE_BYTES, padded),
 * then every bin's amplitudes, then every bin's records; offsets count bytes
 * from the end of the table.  Returns the address past the table, as the
 * reference does (its table is at ret - RCG_BIN_TABLE_BYTES).  Free with
 * rcg_bin_free(ret).  Returns NULL with rcg_last_error() set where the
 * reference fails its assert (a home bin outside [0, 32)^2), where a field
 * leaves int16 or the block exceeds 2^31 - 1 bytes.  Empty bins get
 * n_items = 0 and the offset where they would start (unset in the
 * reference). */
#define RCG_BIN_DIVS 32
#define RCG_BIN_TABLE_BYTES 12320
typedef struct rcg_bin_table {
  int data_size;                               /* whole block, bytes     */
  int data_off;                                /* RCG_BIN_TABLE_BYTES    */
  int vis_off[RCG_BIN_DIVS][RCG_BIN_DIVS];
  int pre_off[RCG_BIN_DIVS][RCG_BIN_DIVS];
  int n_items[RCG_BIN_DIVS][RCG_BIN_DIVS];
} rcg_bin_table;
void *bin(int num_points, double scale, double wstep, rcg_c128 *amps,
          rcg_double3 *uvws, int gcfsupps[], int grid_size);
void *binm(int num_points, double scale, double wstep, rcg_c128 *amps,
           rcg_double3 *uvws, int gcfsupps[], int grid_size);
/* Releases a block returned by bin / binm (the reference has no release
 * function; its block comes from cudaHostAlloc). */
void rcg_bin_free(void *data);

/* Thread-local error of the last call made on this thread (0 = none). */
int rcg_last_error(void);
const char *rcg_last_error_string(void);

/* Which device kernel the calling thread's last gridKernelCPU*GCF or
 * deGridKernelCPU*GCF call ran: RCG_PATH_TILED (binned into uv tiles),
 * RCG_PATH_PLAIN (one wavefront per point, FP64 atomics), RCG_PATH_NONE
 * (no kernel: an argument error or no in-bound point).  Diagnostics only;
 * the reference has no counterpart. */
#define RCG_PATH_NONE 0
#define RCG_PATH_TILED 1
#define RCG_PATH_PLAIN 2
int rcg_legacy_path(void);

/* ---------------------------------------------------------------------
 * 2. Device-resident API
 * ------------------------------------------------------------------- */

/* A GCF table in device memory.  Layer slot s (0 <= s < n_slots) starts at
 * element layer_off[s] of data and holds layer_len[s] complex taps, row
 * stride = its plane's support.  The slot of layer index L (the reference's
 * signed (w_plane*8+over_u)*8+over_v) is L + center for a full table and
 * |L| for a half table (center must be 0).  supps[p + max_w_plane] is the
 * support of w-plane p for -max_w_plane <= p <= max_w_plane.
 * data is rcg_c128 for the FP64 entry points and rcg_c64 for FP32 ones;
 * the FP32 entry points require n_taps * 8 < 2^31. */
typedef struct rcg_gcf {
  const void *data;
  const int64_t *layer_off;
  const int32_t *layer_len;
  int32_t n_slots;
  int32_t center;
  const int32_t *supps;
  int32_t max_w_plane;
  int32_t half;
  int64_t n_taps;     /* complex elements in data (bounds of tap reads) */
} rcg_gcf;

/* Visibility geometry in device memory: n_chunks chunks of ts_ch points,
 * uvw[n_chunks*ts_ch] (metres or wavelengths; scaled by `scale`), the
 * per-chunk support chunk_supps[n_chunks]. */
typedef struct rcg_geom {
  const rcg_double3 *uvw;
  const int32_t *chunk_supps;
  int32_t n_chunks;
  int32_t ts_ch;
  double scale;
  double wstep;
  int32_t grid_size;
  int32_t grid_pitch;
} rcg_geom;

/* Per-point record written by rcg_pregrid (16 bytes). */
typedef struct rcg_pg {
  int32_t cell;     /* u0*grid_pitch + v0 of the window's top-left cell,
                       or -1 when the point is not in-bound            */
  int32_t tap0;     /* element offset of the recentred window in data  */
  int16_t stride;   /* layer row stride (plane support)                */
  int16_t supp;     /* window support (chunk support)                  */
  int32_t flags;    /* bit 0: conjugate taps; bit 1: range error       */
} rcg_pg;

/* FP64 exact pregrid of every point (reference pregridPoint + scatter_gridder_w_dependent_dyn_1p.cpp:128-144).
 * pg must hold n_chunks*ts_ch records.  *taps_dev (device, optional)
 * receives the in-bound tap count.  Status RCG_ERR_RANGE is reported by
 * rcg_plain_check() after the stream is synchronised. */
rcg_status rcg_pregrid(const rcg_geom *geom, const rcg_gcf *gcf, rcg_pg *pg,
                       unsigned long long *taps_dev, int *err_dev,
                       void *stream);

/* FP64 "plain" path: one launch over all points, one wavefront per point,
 * FP64 global atomics.  grid is zeroed first.  vis/grid rcg_c128. */
rcg_status rcg_grid_plain_f64(const rcg_geom *geom, const rcg_gcf *gcf,
                              const rcg_pg *pg, const rcg_c128 *vis,
                              rcg_c128 *grid, void *stream);
rcg_status rcg_degrid_plain_f64(const rcg_geom *geom, const rcg_gcf *gcf,
                                const rcg_pg *pg, const rcg_c128 *grid,
                                rcg_c128 *vis, void *stream);

/* Tiled FP32 path.  A plan owns the device scratch for one or more
 * instances (snapshots, facets) of one grid shape: pregrid records, per-tile
 * bins of (point, tile) pairs, work-item tables and queues.  tile = 16, 32
 * or 64 grid cells; chunk = maximum pairs per work item; max_support bounds
 * every chunk support.
 *
 * rcg_tiled_plan_create_batch takes n_inst geometries (copied into the
 * plan: their device buffers must stay valid while the plan is used) with
 * equal grid_size and grid_pitch; their points are concatenated in
 * instance order in vis and vis_out, and grid holds n_inst grids of
 * grid_size x grid_pitch one after the other.  All instances share the GCF
 * table.  rcg_tiled_plan_create is the one-instance case.
 *
 * rcg_tiled_bin pregrids and bins every point (device only, no host sync);
 * grid/degrid then run on complex64 visibilities, taps (gcf->data) and grid,
 * accumulating in FP32, each as one launch over all instances with a fixed
 * number of workgroups drawing work items from per-XCD queues.  Every call
 * is asynchronous on `stream`; calls that share a plan must be ordered on
 * one stream.  None of them allocates memory, synchronises or copies from
 * the host, so a sequence of them can be captured in a graph.  These entry
 * points have no counterpart in the reference, whose gridder runs on host
 * memory (MS4/kernel/common/scatter_gridder_w_dependent_dyn_1p.h:16-91). */
typedef struct rcg_tiled_plan rcg_tiled_plan;
rcg_status rcg_tiled_plan_create(const rcg_geom *geom, int tile, int chunk,
                                 int max_support, rcg_tiled_plan **plan);
rcg_status rcg_tiled_plan_create_batch(const rcg_geom *geoms, int n_inst,
                                       int tile, int chunk, int max_support,
                                       rcg_tiled_plan **plan);
/* The same with a choice of record layout.  RCG_LAYOUT_PAIRS (the layout of
 * the two functions above): one record per (point, tile) that the point's
 * window overlaps, tile 16, 32 or 64; the only layout rcg_grid_tiled_f64
 * and rcg_degrid_tiled_f64 take.  RCG_LAYOUT_HALO: one record per point whose support is at most 32
 * (16 when max_support <= 16), binned in the tile of its window's first
 * cell; larger windows are cut into ceil(S/32)^2 near-equal sub-windows,
 * each binned likewise; tile 16 or 32.  The kernels then hold a tile plus a
 * 31-cell (15-cell) halo in LDS, so every record is a whole window with its
 * lane geometry fixed at compile time.  chunk = maximum records per work
 * item.  Grids of more tiles per instance than the binning's LDS histogram
 * holds (16384) are binned with global atomics; a plan holds at most 2^20
 * tiles over all instances. */
#define RCG_LAYOUT_PAIRS 0
#define RCG_LAYOUT_HALO 1
#define RCG_LAYOUT_HALO16 2 /* halo with sub-windows of at most 16 cells */
rcg_status rcg_tiled_plan_create_ex(const rcg_geom *geoms, int n_inst,
                                    int tile, int chunk, int max_support,
                                    int layout, rcg_tiled_plan **plan);
rcg_status rcg_tiled_plan_destroy(rcg_tiled_plan *plan);
/* Wavefronts per workgroup of a pairs-layout plan's FP32 gridder, 1..4 (0:
 * the default, 3 at tile 32, 4 otherwise).  Each holds a private FP32 copy
 * of the tile in LDS, so the choice trades copies per workgroup against
 * workgroups per CU; the autotuner (rc_amd/autotune.py) searches it with
 * tile and chunk.  Not to be called while the plan's work is in flight. */
rcg_status rcg_tiled_set_grid_waves(rcg_tiled_plan *plan, int waves);
/* Wavefronts per workgroup of a pairs-layout plan's FP32 degridder: 2, 4 or 8
 * (0: the default 4).  They share one staged copy of the tile.  Not to be
 * called while the plan's work is in flight. */
rcg_status rcg_tiled_set_degrid_waves(rcg_tiled_plan *plan, int waves);
rcg_status rcg_tiled_bin(rcg_tiled_plan *plan, const rcg_gcf *gcf,
                         void *stream);
rcg_status rcg_grid_tiled_f32(rcg_tiled_plan *plan, const rcg_gcf *gcf,
                              const rcg_c64 *vis, rcg_c64 *grid, void *stream);
rcg_status rcg_degrid_tiled_f32(rcg_tiled_plan *plan, const rcg_gcf *gcf,
                                const rcg_c64 *grid, rcg_c64 *vis,
                                void *stream);
/* FP64 gridding over a plan's pairs (after rcg_tiled_bin with an FP64
 * tap table, the same GCF layout): each tap adds the reference's complex
 * product in FP64 (scatter_gridder_w_dependent_dyn_1p.cpp:169, the body of
 * gridKernel_scatter, :54-186, behind gridKernelCPUFullGCF / HalfGCF); only the summation order
 * differs from the sequential one.  grid is zeroed first.  vis/grid
 * rcg_c128.  The gridder of the MS4 host-memory entry points. */
rcg_status rcg_grid_tiled_f64(rcg_tiled_plan *plan, const rcg_gcf *gcf,
                              const rcg_c128 *vis, rcg_c128 *grid, void *stream);
/* FP64 degridding over a plan's pairs (binned as for rcg_grid_tiled_f64):
 * each visibility is the sum of the reference's products grid * tap in FP64
 * (scatter_gridder_w_dependent_dyn_1p.cpp:368-431, behind
 * deGridKernelCPUFullGCF / HalfGCF) in another order; points left out get
 * 0.  The degridder of the MS4 host-memory entry points. */
rcg_status rcg_degrid_tiled_f64(rcg_tiled_plan *plan, const rcg_gcf *gcf,
                                const rcg_c128 *grid, rcg_c128 *vis,
                                void *stream);
/* The Hermitian adjoint of rcg_degrid_tiled_f32: gridding with conjugated
 * taps (the reference's degridder is the plain transpose of its gridder,
 * scatter_gridder_w_dependent_dyn_1p.cpp:368-431, so the adjoint conjugates).  Used as the backward pass of a
 * measurement operator. */
rcg_status rcg_grid_tiled_adjoint_f32(rcg_tiled_plan *plan, const rcg_gcf *gcf,
                                      const rcg_c64 *vis, rcg_c64 *grid,
                                      void *stream);
/* Compensated flush (off by default): gridding adds each tile's sum to the
 * grid with atomics whose rounding errors (exact, by TwoSum on the value
 * the atomic returns) are accumulated in a second grid and added at the
 * end, so the error of the flushes drops to second order.  Enabling it
 * allocates that grid once (not capturable); the switch itself may change
 * between calls. */
rcg_status rcg_tiled_compensated(rcg_tiled_plan *plan, int enable);
/* Synchronises `stream`; returns the in-bound tap count, pair and work-item
 * counts and the device error word of the last rcg_tiled_bin (bit 0: a point
 * referenced a layer outside the table; bit 1: pair buffer overflow). */
rcg_status rcg_tiled_stats(rcg_tiled_plan *plan, unsigned long long *taps,
                           long long *n_pairs, long long *n_items, int *err,
                           void *stream);
/* Kernel timing: with timing on, every grid (which = 0) or degrid
 * (which = 1) call records HIP events around its main kernel on its
 * stream; rcg_tiled_kernel_times waits for and returns the durations (ms)
 * of the last min(count, 64, max) such launches, oldest first, and clears
 * the record.  Switching timing on or off clears it too. */
rcg_status rcg_tiled_timing(rcg_tiled_plan *plan, int enable);
rcg_status rcg_tiled_kernel_times(rcg_tiled_plan *plan, int which, float *ms,
                                  int max, int *count);

/* Graphs: rcg_tiled_graph_create captures the chosen stages (bitwise OR of
 * RCG_STAGE_*) of one plan on fixed buffers into an executable HIP graph;
 * rcg_graph_launch replays it on `stream` (the buffers' current contents
 * are read on every replay).  The graph holds no memset nodes. */
#define RCG_STAGE_BIN 1
#define RCG_STAGE_GRID 2
#define RCG_STAGE_DEGRID 4
typedef struct rcg_graph rcg_graph;
rcg_status rcg_tiled_graph_create(rcg_tiled_plan *plan, const rcg_gcf *gcf,
                                  const rcg_c64 *vis, rcg_c64 *grid,
                                  rcg_c64 *vis_out, int stages,
                                  rcg_graph **graph);
rcg_status rcg_graph_launch(rcg_graph *graph, void *stream);
rcg_status rcg_graph_destroy(rcg_graph *graph);

/* Small kernels on device memory (f64 = 1: complex128/double data, 0:
 * complex64/float).  Semantics as the host entry points above. */
rcg_status rcg_normalize(void *grid, int f64, int grid_pitch, int grid_size,
                         void *stream);
/* vis *= exp(2 pi i w scale) with the phase and its sine/cosine in FP64. */
rcg_status rcg_rotate(const rcg_double3 *uvw, void *vis, int f64, int n,
                      double scale, void *stream);
rcg_status rcg_grid0(const rcg_double3 *uvw, const rcg_c128 *vis,
                     rcg_c128 *grid_centred, double scale, int n,
                     int grid_pitch, int grid_size, void *stream);
/* count_scratch: grid_size^2 ints of device memory. */
rcg_status rcg_reweight(const rcg_double3 *uvw, rcg_c128 *vis, double scale,
                        int n, int grid_size, int *count_scratch, void *stream);

/* rcg_reweight without the count buffer argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops): no
 * counterpart in the reference.  The counts go to a buffer of the library
 * that each device holds for the life of the process (64 MiB), so the call
 * allocates nothing; grid_size above RCG_REWEIGHT_STATIC_MAX_SIZE is refused
 * with RCG_ERR_ARG.  Calls that overlap on one device share the buffer and
 * race: serialise them (one stream, or events). */
#define RCG_REWEIGHT_STATIC_MAX_SIZE 4096
rcg_status rcg_reweight_static(const rcg_double3 *uvw, rcg_c128 *vis, double scale,
                               int n, int grid_size, void *stream);
rcg_status rcg_herm_padded(rcg_c128 *data, int size, int pitch, void *stream);
/* herm_padded_inplace on complex128 (f64) or complex64 data. */
rcg_status rcg_herm(void *data, int f64, int size, int pitch, void *stream);
/* Major-cycle helpers: out -= in over n complex values (the reference's
 * subtractVis, MS4/kernel/gpu/nvidia_degrid/nvidia_degrid.cu:11-18), and
 * vis = re + i im (constVis, MS4/dna-programs/Data.hs:242-247). */
rcg_status rcg_subtract_vis(void *out, const void *in, int f64, long long n,
                            void *stream);
rcg_status rcg_const_vis(void *vis, int f64, long long n, double re, double im,
                         void *stream);
/* Ingest of the MS6 visibility layout, five doubles (u, v, w, re, im) per
 * point (MS6/programs/Kernel/IO.hs:59-71): split into uvw and complex
 * visibilities (complex128 if f64, else complex64), with rotate != 0
 * multiplying each visibility by exp(2 pi i w scale) in FP64 (rotw,
 * MS4/kernel/common/common.h:90-94, as MS4/kernel/cpu/binner/binner.cpp:135
 * does while binning); and the way back. */
rcg_status rcg_unpack_vis5(const double *rec, long long n, int rotate, double scale,
                           rcg_double3 *uvw, void *vis, int f64, void *stream);
rcg_status rcg_pack_vis5(const rcg_double3 *uvw, const void *vis, int f64,
                         long long n, double *rec, void *stream);
rcg_status rcg_fftshift_even(void *data, int elem_bytes, int size, int pitch,
                             void *stream);

/* rocFFT-backed transform with the semantics of fft_inplace_even (sign -1,
 * 0, 1, 2; f64 selects double or single precision) on device memory laid
 * out as the reference's in-place buffer: size rows of pitch complex. */
typedef struct rcg_fft_plan rcg_fft_plan;
rcg_status rcg_fft_plan_create(int sign, int f64, int size, int pitch,
                               rcg_fft_plan **plan);
rcg_status rcg_fft_execute(rcg_fft_plan *plan, void *data, void *stream);
rcg_status rcg_fft_plan_destroy(rcg_fft_plan *plan);

/* Releases process-wide library state before the process exits: the rocFFT
 * library state that the first FFT plan set up (rocfft_setup) is cleaned up
 * (rocfft_cleanup).  Refused with RCG_ERR_ARG while an FFT plan (of
 * rcg_fft_plan_create, fft_inplace_even or a GCF generator) is alive.  A
 * later FFT plan sets the state up again.  No counterpart in the reference,
 * which never releases FFTW's state. */
rcg_status rcg_shutdown(void);

/* GCF generation on device memory (the reference's mkGCFLayer and
 * calcAccums, GCF.cpp:16-219).  A generator owns the (8 max_support)^2
 * FP64 arena with row pitch 8 max_support + src_pad and its rocFFT plan;
 * rcg_gcf_gen_layer writes one plane's 64 layers [8][8][support][support]
 * to dst (complex128 if f64, else complex64).  rcg_calc_accums: device uvw,
 * sums (double) and npts (int) of n_planes entries, err_dev one device int
 * (bit 0: points outside the planes, skipped). */
typedef struct rcg_gcf_gen rcg_gcf_gen;
rcg_status rcg_gcf_gen_create(int max_support, int src_pad, rcg_gcf_gen **gen);
rcg_status rcg_gcf_gen_layer(rcg_gcf_gen *gen, void *dst, int f64, int support,
                             double t2, double w, int pad_quirk, void *stream);
rcg_status rcg_gcf_gen_destroy(rcg_gcf_gen *gen);
rcg_status rcg_calc_accums(const rcg_double3 *uvw, long long n, double wstep,
                           int n_planes, double *sums, int *npts, int *err_dev,
                           void *stream);

/* Hogbom CLEAN on device images (size siz x siz, row pitch `pitch`
 * elements, double if f64 else float).  scratch: rcg_hogbom_scratch_bytes()
 * of device memory; rcg_hogbom_iterations reads back how many peaks were
 * accepted (it synchronises the stream). */
size_t rcg_hogbom_scratch_bytes(void);
rcg_status rcg_hogbom(void *model, void *res, const void *psf, int f64, int siz,
                      int pitch, unsigned niters, double gain, double threshold,
                      void *scratch, void *stream);
rcg_status rcg_hogbom_iterations(const void *scratch, unsigned *iters,
                                 void *stream);

#ifdef __cplusplus
}
#endif

#endif /* RCGRID_H */

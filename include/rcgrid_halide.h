/*
 * rcgrid_halide.h -- the MS6 dna-flow kernel entry points (Halide buffer_t
 * ABI) over librcgrid.so's HIP kernels.
 *
 * buffer_t has the layout of Halide's legacy buffer descriptor, which the
 * MS6 kernels and their Haskell marshalling use
 * (MS6/kernel/common/halide_buf.h:17-27,
 * MS6/dna/flow/Flow/Halide/BufferT.hs:26-99).  Every entry point below runs
 * on the GPU and leaves its result in host memory.
 *
 * Semantics (MS6/kernel/cpu/gridding/scatter.cpp:17-145 and
 * degrid.cpp:17-124):
 *   vis     doubles, extent {5, n}: (u, v, w, re, im) per point, point index
 *           from min[1];
 *   gcf     doubles, extent {2, S, S, 64}: one oversampled GCF of support S
 *           (8, 16, 32 or 64), [over][y][x][re|im] with over = ou + 8 ov;
 *   uvg     doubles, extent {2, X, Y} with min[1], min[2] the grid position
 *           of the buffer's first cell: [y][x][re|im];
 *   pixel   us = u*scale, ou = clamp(round(8 (us - floor us)), 0, 7),
 *           x0 = round(us) + grid_size/2 - S/2 (round: halves to even, as
 *           Halide's round), likewise v;
 *   bounds  m = max(0, (margin_size - S)/2); a point is used only if
 *           min[1] + m <= x0 <= min[1] + extent[1] - S - 1 - m (same in y);
 *   scatter uvg(x0 + sx, y0 + sy) += vis * gcf[over][sy][sx] over every
 *           point (the buffer is added to, not cleared);
 *   degrid  vis_out = vis, then vis_out.re/im -= sum over the window of
 *           uvg(x0 + sx, y0 + sy) * gcf[over][sy][sx].
 * Return 0, or the reference's codes for an unsupported GCF size (-444
 * scatter, -555 degrid, -222 / -333 for the _gpu1 variants); other
 * failures return -1 with rcg_last_error() set.
 */
#ifndef RCGRID_HALIDE_H
#define RCGRID_HALIDE_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef BUFFER_T_DEFINED
#define BUFFER_T_DEFINED
typedef struct buffer_t {
  uint64_t dev;
  uint8_t *host;
  int32_t extent[4];
  int32_t stride[4];
  int32_t min[4];
  int32_t elem_size;
  __attribute__((aligned(1))) bool host_dirty;
  __attribute__((aligned(1))) bool dev_dirty;
  __attribute__((aligned(1))) uint8_t _padding[10 - sizeof(void *)];
} buffer_t;
#endif

/* Replaces kern_scatter (MS6/kernel/cpu/gridding/scatter1.cpp:9-19). */
int kern_scatter(const double scale, const int32_t grid_size,
                 const int32_t margin_size, buffer_t *vis, buffer_t *gcf,
                 buffer_t *uvg);
/* Replaces kern_degrid (MS6/kernel/cpu/gridding/degrid1.cpp:9-18). */
int kern_degrid(const double scale, const int32_t grid_size,
                const int32_t margin_size, buffer_t *gcf, buffer_t *uvg,
                buffer_t *vis, buffer_t *vis_out);
/* Replace kern_scatter_gpu1 / kern_degrid_gpu1
 * (MS6/kernel/gpu/gridding/kern_scatter_gpu1.cpp:22-40,
 * kern_degrid_gpu1.cpp:21-41): a buffer whose `dev` is non-zero and whose
 * host copy is not dirty is read from device memory (dev is a HIP device
 * pointer); the result is copied to `host`, and `dev` of every buffer is
 * cleared as the reference's halide_device_free does.  Unlike the
 * reference's GPU scatter, every point is gridded (it drops the last
 * n mod 200). */
int kern_scatter_gpu1(const double scale, const int32_t grid_size,
                      const int32_t margin_size, buffer_t *vis, buffer_t *gcf,
                      buffer_t *uvg);
int kern_degrid_gpu1(const double scale, const int32_t grid_size,
                     const int32_t margin_size, buffer_t *gcf, buffer_t *uvg,
                     buffer_t *vis, buffer_t *vis_out);

/* Replaces nvGridder (MS6/kernel/nvidia/gridder/halide_api.cu:4-29,
 * grid_gpu.cu:37-243, binsort.cpp:6-32), with its own conventions:
 *   vis  host array of extent[1] points, 5 contiguous doubles each
 *        (x, y, z, re, im); sorted IN PLACE by 8-pixel block of the unscaled
 *        coordinates (block key (gy*2055)/8 + gx with gx = floor(x)/8 in C
 *        integer arithmetic), then by 8(x - floor x) + (y - floor y); equal
 *        keys keep their order;
 *   gcf  host array of 64 16x16 GCFs, [8 sub_y + sub_x][b][a] complex
 *        doubles;
 *   uvg  host array of grid_size x grid_size complex doubles, x fast,
 *        overwritten with the gridded sum;
 *   each point adds vis * gcf[over][b][a] at pixel (floor(x scale) + a - 7,
 *   floor(y scale) + b - 7) where that lies in the grid; the oversampling
 *   offset is floor(8 (xs - floor xs)), likewise y.
 * `dev` and the dirty flags are cleared as the reference does.  Returns 0,
 * or -1 with rcg_last_error() set (the reference has no error path). */
int nvGridder(const double scale, const int32_t grid_size, buffer_t *vis,
              buffer_t *gcf, buffer_t *uvg);

/* The other exports of the same gridder (grid_gpu.cuh:27-43; binsort.cpp:
 * 33-34), on raw host pointers (the reference's double3 / combined /
 * CmplxType records are 3, 5 and 2 doubles, passed here as double *):
 *   bin_sort3 / bin_sort5  nvGridder's in-place sort of npts records of 3
 *                          (x, y, z) or 5 (x, y, z, re, im) doubles;
 *   gridGPUc  grids npts 5-double records (no sort) into out, img_dim^2
 *             complex doubles, overwritten; the reference's gather assumes
 *             sorted input, this grids every point;
 *   gridGPUs  the same with positions (3 doubles per point) and
 *             visibilities (2 doubles per point) in separate arrays.
 * Errors are reported through rcg_last_error() (the reference has none). */
void bin_sort3(double *in, int npts);
void bin_sort5(double *in, int npts);
void gridGPUc(double scale, double *out, double *in_c, size_t npts, size_t img_dim,
              double *gcf);
void gridGPUs(double scale, double *out, double *in, double *in_vals, size_t npts,
              size_t img_dim, double *gcf);

/* The MS6 imaging-glue kernels (Halide generators in
 * MS6/kernel/cpu/gridding; Haskell bindings MS6/programs/Kernel/
 * {Gridder,Facet,Cleaning}.hs).  Each is computed over its output buffer's
 * region (dimension i covers [min[i], min[i] + extent[i])) on host buffers;
 * the computation runs on the GPU.  Return 0, or -1 with rcg_last_error()
 * set where Halide would report a constraint or bounds error.
 *   kern_init        (init.cpp)        grid(c, x, y) = 0;
 *   kern_image_init  (image_init.cpp)  image(x, y) = 0;
 *   kern_psf_vis     (psf_vis.cpp:17-20) out = vis with re = 1, im = 0
 *                    (5 doubles per point, dimension 0 min 0 extent 5);
 *   kern_rotate      (rotate.cpp:28-68, utils.h:137-158) facet (l, m) =
 *                    the output's min[2], min[3]; centre (out_lon0 + l
 *                    out_lon_incr, out_lat0 + m out_lat_incr); R =
 *                    xyz2uvw(out) uvw2xyz(in); uvw' = uvw, or R^T (inv2x2(R)^T
 *                    R) uvw when uvproj; vis' = vis exp(2 pi i (w - R w).uvw);
 *   kern_detile      (detile.cpp:19-30) out(c, x, y) += in(c, x, y) over the
 *                    overlap of both buffers' x and y ranges;
 *   kern_defacet     (defacet.cpp:12-20) image(clamp(x + l), clamp(y + m)) =
 *                    facet(x, y, l, m), x fastest, the last write winning;
 *   kern_image_sum   (image_sum.cpp:12-20) sum(clamp(x), clamp(y)) +=
 *                    image(x, y);
 * where clamp() limits to the output's dimension-0 range [left, right], for
 * y as well (the reference clamps y with the x bounds). */
/* The MS6 visibility binning kernels (MS6/programs/Kernel/Binning.hs), which
 * are Haskell loops over host vectors in the reference (no C counterpart);
 * these entry points run the per-point work on the GPU for a Haskell kernel
 * to call.  Regions, boxes and bins are in uv units (gridXY2UV applied,
 * MS6/programs/Kernel/Config.hs:211-212); vis holds n rows of `width`
 * doubles (u, v, w, re, im, ...).
 *   rcg_flow_bin_sizes  binSizer (:34-97): low, high = min and max of 0 and
 *     the w of the points inside the regions' bounding box, high +=
 *     (high - low) 1e-4; for region r (4 doubles: u_low, u_high, v_low,
 *     v_high) and bin b, out[(r bins + b) 3 + {0, 1, 2}] = start, end (as
 *     doubles) and count (as int64) with start = low + b (high - low) / bins;
 *     a point is counted in the region with the greatest u_low <= u, then
 *     the greatest v_low <= v (no upper test, as Data.Map.lookupLE), bin
 *     floor((w - low) / (high - low) bins) clamped to [0, bins); low_high
 *     (optional) receives low and high.  A NaN w inside the box is an error
 *     (the reference's fold result would depend on its position).
 *   rcg_flow_binner     binner (:99-181), width 5: box b (4 doubles) owns the
 *     bins [bin_off[b], bin_off[b + 1]) with ranges bin_range[2 k], [2 k + 1]
 *     and sizes bin_size[k]; out[b] holds their summed sizes in rows of 5
 *     doubles.  Every point goes, in input order, to the box found by the
 *     greatest w_low <= w, then v_low <= v, then u_low <= u (the first box
 *     and bin with a given triple of lows wins) if it lies in that box's
 *     half-open ranges; then rows [filled, size) are zeroed for every bin of
 *     the box.  filled (optional) receives the rows written per box.  A box
 *     that would receive more rows than it holds is an error (the reference
 *     writes past its vector).
 * Return 0, or -1 with rcg_last_error() set. */
int rcg_flow_bin_sizes(const double *vis, int n, int width, const double *regions,
                       int n_regions, int bins, void *out, double *low_high);
int rcg_flow_binner(const double *vis, int n, int width, const double *boxes,
                    int n_boxes, const int *bin_off, const double *bin_range,
                    const long long *bin_size, double *const *out, long long *filled);

int kern_init(buffer_t *grid);
int kern_image_init(buffer_t *image);
int kern_psf_vis(buffer_t *vis, buffer_t *out);
int kern_rotate(const double in_lon, const double in_lat, const double out_lon0,
                const double out_lat0, const double out_lon_incr,
                const double out_lat_incr, const int32_t uvproj, buffer_t *vis,
                buffer_t *rot);
int kern_detile(buffer_t *in_grid, buffer_t *out_grid);
int kern_defacet(buffer_t *facet, buffer_t *image);
int kern_image_sum(buffer_t *image, buffer_t *sum);

#ifdef __cplusplus
}
#endif

#endif /* RCGRID_HALIDE_H */

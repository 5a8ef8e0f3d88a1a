/*
 * rc_oracle.c -- CPU restatement of the reference's MS4 w-projection
 * gridder, degridder and the small kernels next to them.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * never by the product library (rc_amd/librcgrid.so), which must run its own
 * HIP kernels or fail.
 *
 * Every function cites the reference file:line it restates.  Reference paths
 * are relative to the RC repository root.  Shorthand:
 *   CH  = MS4/kernel/common/common.h
 *   SG  = MS4/kernel/cpu/scatter_grid/scatter_gridder_w_dependent_dyn_1p.cpp
 *   SGH = MS4/kernel/common/scatter_gridder_w_dependent_dyn_1p.h
 *
 * Pinning (what checks each function beyond its own formula):
 *   orc_pregrid          golden vectors from the compiled reference (delta-GCF
 *                        probe, tests/golden), integer bit-exact
 *   orc_grid/orc_degrid  golden grids / visibilities from the compiled
 *                        reference at one thread (bit-exact), delta KATs,
 *                        sum invariant, bilinear transpose identity, brute
 *                        force on tiny inputs
 *   orc_normalize/rotate/grid0/reweight
 *                        golden vectors from the compiled reference + closed
 *                        forms (scale, exp(2 pi i w scale), nearest cell,
 *                        1/count)
 *   orc_herm_padded, orc_fftshift_even
 *                        closed forms (index mirror, quadrant swap is an
 *                        involution equal to numpy.fft.fftshift for even N)
 *   orc_bin              the MS4 binner: its pregrid part is orc_pregrid
 *                        (pinned above) minus the translation; bins, margin
 *                        copies and order are pinned by a geometric check
 *                        independent of the if / else-if form (a copy goes
 *                        to each neighbour bin that the window
 *                        [p.u - m, p.u + m) x [p.v - m, p.v + m) reaches,
 *                        m <= size/64 so at most one side per axis), by
 *                        hand-made KATs of the remainder edge cases
 *                        (negative coordinates in bin 0, m > size/64 where
 *                        the if / else-if order picks the side) and by the
 *                        input order inside every bin.  binner.cpp needs the
 *                        CUDA runtime (cudaHostAlloc) and cannot be built
 *                        here.
 *   orc_hogbom           textbook Hogbom CLEAN; the reference's deconvolve
 *                        (MS4/kernel/cpu/hogbom/hogbom.cpp:32-121) reads
 *                        out of bounds in findPeak and cannot serve as a
 *                        vector source, so this function is pinned by brute
 *                        force on tiny images (single point source recovered
 *                        exactly, residual = dirty - model * psf).
 *
 * Strength of the pinning: tests/test_oracle_mutations.py plants ten
 * plausible mistakes here (orc_set_mutation: transposed tap operand,
 * dropped conjugation, rounded oversampling offset, dropped recentring,
 * conjugating degridder, non-strict in-bound test, scaled w-plane, wrong
 * normalisation, flipped rotation sign, signed Hogbom peak) and requires
 * each to fail at least one of the checks above.
 *
 * Build: plain C11, no fast-math, no FP contraction, so that the operation
 * order below is the one executed (oracle/Makefile).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OVER 8 /* MS4/kernel/common/metrix.h:5 (OVER) */

/* Planted mistakes, for tests that show the pinning checks catch them
 * (tests/test_oracle_mutations.py).  0 = the restatement as written.
 *   1 transposed tap operand (layer[sv][su])
 *   2 HalfGCF conjugation dropped
 *   3 oversampling offset rounded instead of floored
 *   4 window recentring dropped
 *   5 degridder conjugates its taps (adjoint instead of transpose)
 *   6 in-bound test not strict (u0 <= N - S accepted)
 *   7 w-plane from the scaled w
 *   8 normalisation by grid_size^2 instead of grid_size*grid_pitch
 *   9 phase rotation with the opposite sign
 *  10 Hogbom peak by signed value instead of magnitude
 *  11 binner: right / top margin test not strict (rem >= size - margin)
 *  12 binner: bins by floor division instead of C's truncating division
 *  13 binner: first and last point swapped in the bin order */
static int orc_mutation = 0;
void orc_set_mutation(int m) { orc_mutation = m; }

typedef struct { double re, im; } ocd;       /* complexd, CH:37-39 */
typedef struct { double u, v, w; } od3;      /* Double3,  CH:72-77 */

/* Pregridded point, CH:79-88, widened to int (see orc_pregrid). */
typedef struct {
  int u, v;         /* top-left grid cell of the support window          */
  int over_u, over_v;
  int w_plane;      /* signed plane                                       */
  int layer_index;  /* signed GCF layer index (negated for mirrored w<0)  */
  int valid;        /* 0 if a coordinate leaves the reference's int16 range */
} orc_pg;

/* CH:115-170, pregridPoint<over=8, do_mirror> with __DYN_GRID_SIZE.
 * The reference stores u, v, over_*, w_plane and the layer index in int16;
 * conversions of out-of-range doubles to int16 are undefined behaviour
 * there.  Here a point whose floor(u*scale), floor(v*scale), round(w/wstep)
 * or layer index leaves the int16 range is marked invalid and is treated as
 * not in-bound by the gridder (skipped, not counted).  Inside the range the
 * arithmetic is the reference's, operation for operation. */
void orc_pregrid(double scale, double wstep, const od3 *p, int max_supp,
                 int grid_size, int mirror, orc_pg *r)
{
  double us = p->u * scale;
  double vs = p->v * scale;
  double wr = round(orc_mutation == 7 ? p->w * scale / wstep : p->w / wstep);
  double fu = floor(us), fv = floor(vs);
  r->valid = 1;
  if (!(fu >= -32768.0 && fu <= 32767.0) || !(fv >= -32768.0 && fv <= 32767.0)
      || !(wr >= -32768.0 && wr <= 32767.0)) {
    r->valid = 0;
    r->u = r->v = r->over_u = r->over_v = r->w_plane = r->layer_index = 0;
    return;
  }
  int u = (int)fu, v = (int)fv, w_plane = (int)wr;
  /* CH:130-131: over * (uvw.u - u) with u the int16 floor */
  int over_u = (int)floor(ORC_OVER * (us - (double)u));
  int over_v = (int)floor(ORC_OVER * (vs - (double)v));
  if (orc_mutation == 3) {
    over_u = (int)round(ORC_OVER * (us - (double)u)) % ORC_OVER;
    over_v = (int)round(ORC_OVER * (vs - (double)v)) % ORC_OVER;
  }
  /* CH:136-137 */
  u += grid_size / 2 - max_supp / 2;
  v += grid_size / 2 - max_supp / 2;
  int idx;
  /* CH:158-166 */
  if (mirror && w_plane < 0)
    idx = -((-w_plane * ORC_OVER + over_u) * ORC_OVER + over_v);
  else
    idx = (w_plane * ORC_OVER + over_u) * ORC_OVER + over_v;
  if (idx < -32768 || idx > 32767) r->valid = 0;
  r->u = u; r->v = v; r->over_u = over_u; r->over_v = over_v;
  r->w_plane = w_plane; r->layer_index = idx;
}

/* Exported probe for tests: pregrid n points of one chunk. */
void orc_pregrid_many(double scale, double wstep, const od3 *uvw, int n,
                      int max_supp, int grid_size, int mirror, int *out6)
{
  for (int i = 0; i < n; i++) {
    orc_pg r;
    orc_pregrid(scale, wstep, uvw + i, max_supp, grid_size, mirror, &r);
    out6[6 * i + 0] = r.u;
    out6[6 * i + 1] = r.v;
    out6[6 * i + 2] = r.over_u;
    out6[6 * i + 3] = r.over_v;
    out6[6 * i + 4] = r.w_plane;
    out6[6 * i + 5] = r.valid ? r.layer_index : INT32_MIN;
  }
}

/* Per-point tap source, SG:128-144: layer pointer recentred to the
 * chunk's support window, its row stride, the conjugation flag (HalfGCF,
 * negative index) and the in-bound flag (strict upper bound). */
typedef struct {
  const ocd *layer;
  int stride;
  int conj;
  int inb;
  int u, v;
} orc_src;

static void orc_resolve(const orc_pg *pg, int half, const ocd *const *gcf,
                        const int *gcf_supps, int max_supp, int grid_size,
                        orc_src *s)
{
  int idx = pg->layer_index;
  s->u = pg->u;
  s->v = pg->v;
  s->inb = pg->valid && !(pg->u < 0 || pg->u >= grid_size - max_supp ||
                          pg->v < 0 || pg->v >= grid_size - max_supp);
  if (orc_mutation == 6)
    s->inb = pg->valid && !(pg->u < 0 || pg->u > grid_size - max_supp ||
                            pg->v < 0 || pg->v > grid_size - max_supp);
  if (!pg->valid) { s->layer = NULL; s->stride = 0; s->conj = 0; return; }
  /* SG:130-137 */
  if (half) {
    s->conj = orc_mutation == 2 ? 0 : idx < 0;
    s->layer = gcf[idx < 0 ? -idx : idx];
  } else {
    s->conj = 0;
    s->layer = gcf[idx];
  }
  /* SG:139-140: gcf_supps indexed by the signed plane in both variants */
  int gs = gcf_supps[pg->w_plane];
  s->stride = gs;
  if (orc_mutation != 4) s->layer += (gs - max_supp) / 2 * (gs + 1);
}

static inline ocd orc_tap(const orc_src *s, int su, int sv)
{
  ocd t = orc_mutation == 1 ? s->layer[sv * s->stride + su]
                            : s->layer[su * s->stride + sv];     /* SG:158 */
  if (s->conj) t.im = -t.im;                 /* SG:160-161 */
  return t;
}

/* SG:49-185 (gridding branch) + SG:187-243 at one thread: the grid is
 * zeroed, then for every chunk in order, for su, for point, for sv:
 * grid[u+su][v+sv] += vis * tap.  Returns the number of in-bound taps
 * (SG:173, 241).
 *
 * nthreads > 1 splits the grid rows into bands; every thread walks all
 * chunks in the same order and only applies taps whose row falls in its
 * band, so each cell receives its contributions in exactly the one-thread
 * order and the result is bit-identical to nthreads == 1. */
unsigned long long orc_grid(int half, double scale, double wstep,
                            int baselines, const int *bl_supps, ocd *grid,
                            const ocd *const *gcf, const od3 *const *uvw,
                            const ocd *const *vis, int ts_ch, int grid_pitch,
                            int grid_size, const int *gcf_supps, int nthreads)
{
  size_t siz = (size_t)grid_size * (size_t)grid_pitch;
  memset(grid, 0, sizeof(ocd) * siz);
  if (nthreads < 1) nthreads = 1;
  unsigned long long taps_total = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads) reduction(+ : taps_total)
#endif
  {
#ifdef _OPENMP
    int tno = omp_get_thread_num();
    int nth = omp_get_num_threads();
#else
    int tno = 0, nth = 1;
#endif
    int row_lo = (int)((long long)grid_size * tno / nth);
    int row_hi = (int)((long long)grid_size * (tno + 1) / nth);
    orc_pg *pa = (orc_pg *)malloc(sizeof(orc_pg) * (size_t)(ts_ch > 0 ? ts_ch : 1));
    orc_src *sa = (orc_src *)malloc(sizeof(orc_src) * (size_t)(ts_ch > 0 ? ts_ch : 1));
    unsigned long long taps = 0;
    for (int bl = 0; bl < baselines; bl++) {
      int ms = bl_supps[bl];                                /* SG:88-89 */
      for (int n = 0; n < ts_ch; n++) {                     /* SG:118-145 */
        orc_pregrid(scale, wstep, uvw[bl] + n, ms, grid_size, half, pa + n);
        orc_resolve(pa + n, half, gcf, gcf_supps, ms, grid_size, sa + n);
      }
      for (int su = 0; su < ms; su++) {                     /* SG:146 */
        for (int i = 0; i < ts_ch; i++) {                   /* SG:147 */
          if (!sa[i].inb) continue;                         /* SG:155 */
          int gsu = sa[i].u + su;
          int mine = gsu >= row_lo && gsu < row_hi;
          ocd vv = vis[bl][i];                              /* SG:123 */
          for (int sv = 0; sv < ms; sv++) {                 /* SG:150 */
            if (mine) {
              ocd t = orc_tap(sa + i, su, sv);
              ocd *g = grid + (size_t)gsu * grid_pitch + (sa[i].v + sv);
              /* SG:169: std::complex product then += */
              double pr = vv.re * t.re - vv.im * t.im;
              double pi = vv.re * t.im + vv.im * t.re;
              g->re += pr;
              g->im += pi;
            }
          }
          if (mine) taps += (unsigned long long)ms;         /* SG:173 */
        }
      }
    }
    taps_total += taps;
    free(pa);
    free(sa);
  }
  return taps_total;
}

/* SG:49-185 (__DEGRID branch) + SG:368-431: every point of every chunk is
 * zeroed (SG:124-126), then vis += grid[u+su][v+sv] * tap in the order su,
 * point, sv, which for one point is row-major over its window.  The
 * transpose of orc_grid: taps are not conjugated beyond the HalfGCF rule.
 * Chunks are independent, so nthreads > 1 is bit-identical to one thread. */
unsigned long long orc_degrid(int half, double scale, double wstep,
                              int baselines, const int *bl_supps,
                              const ocd *grid, const ocd *const *gcf,
                              const od3 *const *uvw, ocd *const *vis,
                              int ts_ch, int grid_pitch, int grid_size,
                              const int *gcf_supps, int nthreads)
{
  if (nthreads < 1) nthreads = 1;
  unsigned long long taps_total = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads) reduction(+ : taps_total)
#endif
  {
    orc_pg *pa = (orc_pg *)malloc(sizeof(orc_pg) * (size_t)(ts_ch > 0 ? ts_ch : 1));
    orc_src *sa = (orc_src *)malloc(sizeof(orc_src) * (size_t)(ts_ch > 0 ? ts_ch : 1));
    unsigned long long taps = 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 23)
#endif
    for (int bl = 0; bl < baselines; bl++) {
      int ms = bl_supps[bl];
      ocd *vp = vis[bl];
      for (int n = 0; n < ts_ch; n++) {
        vp[n].re = 0.0;
        vp[n].im = 0.0;
        orc_pregrid(scale, wstep, uvw[bl] + n, ms, grid_size, half, pa + n);
        orc_resolve(pa + n, half, gcf, gcf_supps, ms, grid_size, sa + n);
      }
      for (int su = 0; su < ms; su++) {
        for (int i = 0; i < ts_ch; i++) {
          if (!sa[i].inb) continue;
          const ocd *grow = grid + (size_t)(sa[i].u + su) * grid_pitch + sa[i].v;
          for (int sv = 0; sv < ms; sv++) {
            ocd t = orc_tap(sa + i, su, sv);
            if (orc_mutation == 5) t.im = -t.im;
            ocd g = grow[sv];
            /* SG:171: vis[i] += grid * supportPixel */
            double pr = g.re * t.re - g.im * t.im;
            double pi = g.re * t.im + g.im * t.re;
            vp[i].re += pr;
            vp[i].im += pi;
          }
          taps += (unsigned long long)ms;
        }
      }
    }
    taps_total += taps;
    free(pa);
    free(sa);
  }
  return taps_total;
}

/* SG:293-306 */
void orc_normalize(ocd *src, int grid_pitch, int grid_size)
{
  size_t siz = (size_t)grid_size * (size_t)grid_pitch;
  double norm = 1.0 / (double)(grid_size * (orc_mutation == 8 ? grid_size : grid_pitch));
  for (size_t i = 0; i < siz; i++) {
    src[i].re *= norm;
    src[i].im *= norm;
  }
}

/* SG:355-363 with rotw, CH:90-94: vis *= exp(+2 pi i * w * scale) */
void orc_rotate(const od3 *uvw, ocd *vis, int n, double scale)
{
  for (int i = 0; i < n; i++) {
    double ph = (orc_mutation == 9 ? -2.0 : 2.0) * M_PI * (uvw[i].w * scale);
    double s = sin(ph), c = cos(ph);
    double re = vis[i].re * c - vis[i].im * s;
    double im = vis[i].re * s + vis[i].im * c;
    vis[i].re = re;
    vis[i].im = im;
  }
}

/* SG:268-291: nearest-cell gridding into a grid pointer that the caller
 * has centred (cppcycle.cpp:93).  The bound test is on the uncentred
 * offset n + trans in [0, last). */
void orc_grid0(const od3 *uvw, const ocd *vis, ocd *grid, double scale, int n,
               int grid_pitch, int grid_size)
{
  long long last = (long long)grid_size * grid_pitch;
  long long trans = (long long)(grid_size / 2) * (grid_pitch + 1);
  for (int i = 0; i < n; i++) {
    int u = (int)round(uvw[i].u * scale);
    int v = (int)round(uvw[i].v * scale);
    long long k = (long long)u * grid_pitch + v;
    long long kg = k + trans;
    if (kg >= 0 && kg < last) {
      grid[k].re += vis[i].re;
      grid[k].im += vis[i].im;
    }
  }
}

/* SG:311-353: uniform weighting; every visibility whose rounded cell lies
 * in the grid is divided by the number of visibilities in that cell. */
void orc_reweight(const od3 *uvw, ocd *vis, double scale, int n, int grid_size)
{
  long long last = (long long)grid_size * grid_size;
  long long trans = (long long)(grid_size / 2) * (grid_size + 1);
  int *cnt = (int *)calloc((size_t)last, sizeof(int));
  long long *cell = (long long *)malloc(sizeof(long long) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++) {
    int u = (int)round(uvw[i].u * scale);
    int v = (int)round(uvw[i].v * scale);
    long long kg = (long long)u * grid_size + v + trans;
    cell[i] = kg;
    if (kg >= 0 && kg < last) cnt[kg]++;
  }
  for (int i = 0; i < n; i++) {
    long long kg = cell[i];
    if (kg >= 0 && kg < last) {
      double c = (double)cnt[kg];
      /* std::complex<double> /= double divides both parts */
      vis[i].re /= c;
      vis[i].im /= c;
    }
  }
  free(cnt);
  free(cell);
}

/* MS4/kernel/cpu/herm/herm_padded.cpp:6-27 (the default, non-MOVE_TO_TOP
 * branch): for every row i and column j < size/2,
 *   A[i][j] += conj(A[size-1-i][size-1-j]).
 * The source cells (columns >= size/2) are never written, so the loop order
 * does not matter. */
void orc_herm_padded(ocd *data, int size, int pitch)
{
  for (int i = 0; i < size; i++)
    for (int j = 0; j < size / 2; j++) {
      const ocd *s = data + (size_t)(size - 1 - i) * pitch + (size - 1 - j);
      ocd *d = data + (size_t)i * pitch + j;
      d->re += s->re;
      d->im -= s->im;
    }
}

/* MS4/kernel/cpu/fft/fft_dyn_padded.cpp:26-48: swap quadrants 0<->3 and
 * 1<->2 of an even size x size array of elem_bytes-sized elements with row
 * pitch `pitch` elements. */
void orc_fftshift_even(void *data, int elem_bytes, int size, int pitch)
{
  char *d = (char *)data;
  int h = size / 2;
  char tmp[32];
  for (int i = 0; i < h; i++)
    for (int j = 0; j < size; j++) {
      int jj = (j + h) % size;
      char *a = d + ((size_t)i * pitch + j) * elem_bytes;
      char *b = d + ((size_t)(i + h) * pitch + jj) * elem_bytes;
      memcpy(tmp, a, (size_t)elem_bytes);
      memcpy(a, b, (size_t)elem_bytes);
      memcpy(b, tmp, (size_t)elem_bytes);
    }
}

/* Textbook Hogbom CLEAN over the size x size window of pitch-strided real
 * images (the interface of MS4/kernel/common/hogbom.h:7-16):
 *   model = 0
 *   psf peak p = argmax |psf| (first index wins ties), pv = psf[p]
 *   repeat niters: r = argmax |res|, v = res[r]; stop if |v| < threshold;
 *     res[x] -= gain * v / pv * psf[x - r + p] wherever x - r + p lies in
 *     the window; model[r] += gain * v / pv.
 * Pad columns are neither read nor written.  Returns the iterations done. */
unsigned orc_hogbom(double *mod, double *res, const double *psf, int siz,
                    int pitch, unsigned niters, double gain, double threshold)
{
  for (int i = 0; i < siz; i++)
    for (int j = 0; j < siz; j++) mod[(size_t)i * pitch + j] = 0.0;
  int px = 0, py = 0;
  double pbest = -1.0;
  for (int i = 0; i < siz; i++)
    for (int j = 0; j < siz; j++) {
      double a = fabs(psf[(size_t)i * pitch + j]);
      if (a > pbest) { pbest = a; px = i; py = j; }
    }
  double pv = psf[(size_t)px * pitch + py];
  unsigned it;
  for (it = 0; it < niters; it++) {
    int rx = 0, ry = 0;
    double best = -1.0;
    for (int i = 0; i < siz; i++)
      for (int j = 0; j < siz; j++) {
        double a = orc_mutation == 10 ? res[(size_t)i * pitch + j]
                                      : fabs(res[(size_t)i * pitch + j]);
        if (a > best) { best = a; rx = i; ry = j; }
      }
    double v = res[(size_t)rx * pitch + ry];
    if (fabs(v) < threshold) break;
    double a = gain * v / pv;
    int dx = px - rx, dy = py - ry;
    for (int i = 0; i < siz; i++) {
      int si = i + dx;
      if (si < 0 || si >= siz) continue;
      for (int j = 0; j < siz; j++) {
        int sj = j + dy;
        if (sj < 0 || sj >= siz) continue;
        res[(size_t)i * pitch + j] -= a * psf[(size_t)si * pitch + sj];
      }
    }
    mod[(size_t)rx * pitch + ry] += a;
  }
  return it;
}

/* MS4/kernel/cpu/binner/binner.cpp:100-189, doit<OVER, DIVIDERS = 32,
 * mirror> (bin / binm), with pregridPoint<8, mirror> compiled under __BINNER
 * (CH:96-171: no translation inside the pregrid).  For every point in input
 * order:
 *   - p = the untranslated pregrid record; (qu, ru) = div(p.u, grid_size/32)
 *     in C's truncating division, likewise v (binner.cpp:135-137);
 *   - supp = gcfsupps[w_plane] (signed plane, centred table); the record is
 *     translated by grid_size/2 - supp/2 and carries supp in place of the
 *     plane (binner.cpp:143-151); the amplitude is rotated by
 *     exp(2 pi i w scale) (binner.cpp:157-158, CH:90-94);
 *   - it goes to bin (qu, qv), then, if margin = supp/2 > 0, to at most one
 *     horizontal neighbour, one vertical neighbour and the diagonal between
 *     them (binner.cpp:161-196; the if / else-if order decides the side).
 * The reference appends to one stream per bin, so each bin holds its copies
 * in input order (binner.cpp:87-90).
 *
 * Output: counts[32*32] (bin qu*32 + qv); if pre_out is not NULL, for each
 * copy in bin order its record (pre_out, 4 int16: u, v, layer index,
 * support), rotated amplitude (vis_out) and source point (src_out), up to
 * cap copies.  Returns the number of copies, or -1 where the reference fails
 * its assert (a home bin outside [0, 32)^2, binner.cpp:153-155), where a
 * coordinate leaves int16 (undefined there) or grid_size < 32 (a division by
 * zero there). */
#define ORC_DIVS 32
typedef struct { int bin[4]; short pre[4]; ocd vis; } orc_binpt;

static int orc_bin_point(int mirror, double scale, double wstep, const ocd *amp,
                         const od3 *uvw, const int *gcfsupps, int grid_size,
                         orc_binpt *o)
{
  orc_pg p;
  orc_pregrid(scale, wstep, uvw, 0, 0, mirror, &p);  /* __BINNER: untranslated */
  if (!p.valid) return -1;
  const int ds = grid_size / ORC_DIVS;
  /* div(): quotient truncated toward zero, remainder of the dividend's sign */
  const int qu = p.u / ds, ru = p.u % ds;
  const int qv = p.v / ds, rv = p.v % ds;
  if (orc_mutation == 12) {  /* planted: floor division */
    const int fqu = (int)floor((double)p.u / ds), fqv = (int)floor((double)p.v / ds);
    if (fqu != qu || fqv != qv) return -1;
  }
  if (qu < 0 || qv < 0 || qu >= ORC_DIVS || qv >= ORC_DIVS) return -1;
  const int supp = gcfsupps[p.w_plane];
  o->pre[0] = (short)(p.u + (short)(grid_size / 2 - supp / 2));
  o->pre[1] = (short)(p.v + (short)(grid_size / 2 - supp / 2));
  o->pre[2] = (short)p.layer_index;
  o->pre[3] = (short)supp;
  const double ph = 2.0 * M_PI * (uvw->w * scale);
  const double s = sin(ph), c = cos(ph);
  o->vis.re = amp->re * c - amp->im * s;
  o->vis.im = amp->re * s + amp->im * c;
  o->bin[0] = qu * ORC_DIVS + qv;
  o->bin[1] = o->bin[2] = o->bin[3] = -1;
  const int margin = supp / 2;
  if (margin > 0) {
    int du = 0, dv = 0;
    const int right = orc_mutation == 11 ? ru >= ds - margin : ru > ds - margin;
    if (ru < margin && qu >= 1) du = -1;
    else if (right && qu < ORC_DIVS - 1) du = 1;
    if (rv < margin && qv >= 1) dv = -1;
    else if (rv > ds - margin && qv < ORC_DIVS - 1) dv = 1;
    if (du) o->bin[1] = (qu + du) * ORC_DIVS + qv;
    if (dv) o->bin[2] = qu * ORC_DIVS + qv + dv;
    if (du && dv) o->bin[3] = (qu + du) * ORC_DIVS + qv + dv;
  }
  return 0;
}

long long orc_bin(int mirror, int n, double scale, double wstep, const ocd *amps,
                  const od3 *uvws, const int *gcfsupps, int grid_size, int *counts,
                  short *pre_out, ocd *vis_out, int *src_out, long long cap)
{
  const int nb = ORC_DIVS * ORC_DIVS;
  for (int b = 0; b < nb; b++) counts[b] = 0;
  if (grid_size < ORC_DIVS) return -1;
  orc_binpt *pts = (orc_binpt *)malloc(sizeof(orc_binpt) * (size_t)(n > 0 ? n : 1));
  if (!pts) return -1;
  long long total = 0;
  for (int i = 0; i < n; i++) {
    if (orc_bin_point(mirror, scale, wstep, amps + i, uvws + i, gcfsupps, grid_size,
                      pts + i) != 0) {
      free(pts);
      return -1;
    }
    for (int k = 0; k < 4; k++)
      if (pts[i].bin[k] >= 0) { counts[pts[i].bin[k]]++; total++; }
  }
  if (pre_out) {
    long long *cur = (long long *)malloc(sizeof(long long) * nb);
    long long run = 0;
    for (int b = 0; b < nb; b++) { cur[b] = run; run += counts[b]; }
    for (int i = 0; i < n; i++) {
      /* planted 13: the last point of each bin first */
      const int ii = orc_mutation == 13 && i == 0 ? n - 1
                     : orc_mutation == 13 && i == n - 1 ? 0 : i;
      for (int k = 0; k < 4; k++) {
        const int b = pts[ii].bin[k];
        if (b < 0) continue;
        const long long at = cur[b]++;
        if (at >= cap) continue;
        memcpy(pre_out + 4 * at, pts[ii].pre, sizeof(pts[ii].pre));
        vis_out[at] = pts[ii].vis;
        src_out[at] = ii;
      }
    }
    free(cur);
  }
  free(pts);
  return total;
}

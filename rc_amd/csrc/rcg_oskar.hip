// rcg_oskar.hip -- reader of OSKAR 2.6 visibility files (binary format
// version 2) into the gridder's [baseline][time][channel] layout, with the
// reshuffle and the coordinate extremes on the GPU (include/rcgrid_oskar.h).
//
// Reference behaviour: MS4/dep/oskar/OskarBinReader.cpp:155-413 (mkFromFile,
// readAndReshuffle) over the vendored oskar_binary library.  This file
// reads the format itself:
//  * a 64-byte header starting "OSKARBIN\0" and the format version;
//  * then chunks, each a 20-byte tag ('T', 0x40 + version, 'G', element
//    size, flags, data type, group id, tag id, 4-byte user index, 8-byte
//    block size = bytes to the next tag), for an extended tag (flag bit 7)
//    the group and tag names, the payload, and with flag bit 6 a 4-byte
//    CRC-32C (Castagnoli, reflected polynomial 0x82F63B78) of everything
//    from the tag to the end of the payload;
//  * a value is found by (data type, group, tag, user index), searching from
//    a start index (oskar_binary_query.c semantics), and its CRC checked
//    when it is read.
// The visibility header (group 11) and block (group 12) tags and their
// meaning are those of OskarBinReader.cpp:59-90.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "rcg_internal.h"
#include "rcgrid_oskar.h"

namespace {

constexpr int OSKAR_INT = 0x02, OSKAR_DOUBLE = 0x08, OSKAR_COMPLEX = 0x20, OSKAR_MATRIX = 0x40;
constexpr int OSKAR_DOUBLE_COMPLEX_MATRIX = OSKAR_DOUBLE | OSKAR_COMPLEX | OSKAR_MATRIX;
constexpr int GROUP_VIS_HEADER = 11, GROUP_VIS_BLOCK = 12;
// header tags
constexpr int NUM_TAGS_PER_BLOCK = 2, AMP_TYPE = 4, MAX_TIMES_PER_BLOCK = 5,
              NUM_TIMES_TOTAL = 6, NUM_CHANNELS = 7, NUM_STATIONS = 8, FREQ_START_HZ = 9,
              FREQ_INC_HZ = 10, CHANNEL_BANDWIDTH_HZ = 11, TIME_START_MJD_UTC = 12,
              TIME_INC_SEC = 13, TIME_AVERAGE_SEC = 14, PHASE_CENTRE = 15,
              TELESCOPE_CENTRE = 16;
// block tags
constexpr int DIM_START_AND_SIZE = 1, FREQ_REF_INC_HZ = 2, TIME_REF_INC_MJD_UTC = 3,
              CROSS_CORRELATIONS = 5, BASELINE_UU = 6, BASELINE_VV = 7, BASELINE_WW = 8;

// CRC-32C, table-driven (Sarwate), reflected polynomial 0x82F63B78.
struct Crc32c {
  unsigned t[256];
  Crc32c()
  {
    for (unsigned i = 0; i < 256; i++) {
      unsigned c = i;
      for (int k = 0; k < 8; k++) c = (c >> 1) ^ ((c & 1u) ? 0x82F63B78u : 0u);
      t[i] = c;
    }
  }
  // state: the running register (start 0xFFFFFFFF; the CRC is ~state)
  unsigned update(unsigned state, const void *data, size_t n) const
  {
    const unsigned char *p = static_cast<const unsigned char *>(data);
    for (size_t i = 0; i < n; i++) state = t[(state ^ p[i]) & 0xFFu] ^ (state >> 8);
    return state;
  }
};

const Crc32c &crc_table()
{
  static const Crc32c c;
  return c;
}

unsigned le32(const unsigned char *p)
{
  return (unsigned)p[0] | (unsigned)p[1] << 8 | (unsigned)p[2] << 16 | (unsigned)p[3] << 24;
}

unsigned long long le64(const unsigned char *p)
{
  return (unsigned long long)le32(p) | (unsigned long long)le32(p + 4) << 32;
}

struct Chunk {
  bool extended;
  int data_type, group, tag, index;
  long long payload_off;
  unsigned long long payload_bytes;
  bool has_crc;
  unsigned crc_state;  // CRC register after the tag (and names)
  unsigned crc;        // stored CRC
};

}  // namespace

struct rcg_oskar_file {
  FILE *f = nullptr;
  std::vector<Chunk> chunks;
  ~rcg_oskar_file()
  {
    if (f) fclose(f);
  }

  // Index of the chunk (data type, group, tag, index) from `start` on, or -1.
  int find(int data_type, int group, int tag, int index, int start) const
  {
    for (int i = std::max(start, 0); i < (int)chunks.size(); i++) {
      const Chunk &c = chunks[i];
      if (!c.extended && c.data_type == data_type && c.group == group && c.tag == tag &&
          c.index == index)
        return i;
    }
    return -1;
  }

  // Reads the payload of (data type, group, tag, index) into `out` of
  // `capacity` bytes and checks its CRC; returns 0 or an OSKAR status.
  int read(int data_type, int group, int tag, int index, int start, void *out,
           size_t capacity, size_t *got = nullptr) const
  {
    const int i = find(data_type, group, tag, index, start);
    if (i < 0) return RCG_OSKAR_ERR_TAG_NOT_FOUND;
    const Chunk &c = chunks[i];
    if (got) *got = (size_t)c.payload_bytes;
    if (c.payload_bytes == 0) return 0;
    if (!out || capacity < c.payload_bytes) return RCG_OSKAR_ERR_MEMORY_NOT_ALLOCATED;
    if (fseeko(f, (off_t)c.payload_off, SEEK_SET) != 0) return RCG_OSKAR_ERR_SEEK_FAIL;
    if (fread(out, 1, (size_t)c.payload_bytes, f) != c.payload_bytes)
      return RCG_OSKAR_ERR_READ_FAIL;
    if (c.has_crc &&
        ~crc_table().update(c.crc_state, out, (size_t)c.payload_bytes) != c.crc)
      return RCG_OSKAR_ERR_CRC_FAIL;
    return 0;
  }
};

namespace {

// Opens and indexes a file; 0 or an OSKAR status.
int open_index(const char *filename, rcg_oskar_file **out)
{
  *out = nullptr;
  FILE *f = filename ? fopen(filename, "rb") : nullptr;
  if (!f) return RCG_OSKAR_ERR_OPEN_FAIL;
  auto *h = new (std::nothrow) rcg_oskar_file();
  if (!h) {
    fclose(f);
    return RCG_OSKAR_ERR_MEMORY_NOT_ALLOCATED;
  }
  h->f = f;
  unsigned char hdr[64];
  if (fread(hdr, 1, 64, f) != 64 || memcmp(hdr, "OSKARBIN", 9) != 0) {
    delete h;
    return RCG_OSKAR_ERR_FILE_INVALID;
  }
  if (hdr[9] < 1 || hdr[9] > 2) {
    delete h;
    return RCG_OSKAR_ERR_VERSION_UNKNOWN;
  }
  long long pos = 64;
  for (;;) {
    unsigned char tg[20];
    if (fread(tg, 1, 20, f) != 20) break;  // end of file
    const int flags = tg[4];
    const int version = tg[1] - 0x40;
    if (tg[0] != 'T' || tg[2] != 'G' || (flags & 0x1F) != 0) {
      delete h;
      return RCG_OSKAR_ERR_FILE_INVALID;
    }
    if (version < 1 || version > 2) {
      delete h;
      return RCG_OSKAR_ERR_VERSION_UNKNOWN;
    }
    Chunk c{};
    c.extended = (flags & 0x80) != 0;
    c.data_type = tg[5];
    c.group = tg[6];
    c.tag = tg[7];
    c.index = (int)le32(tg + 8);
    const unsigned long long block = le64(tg + 12);
    c.has_crc = (flags & 0x40) != 0;
    if (c.has_crc && block < 4) {
      delete h;
      return RCG_OSKAR_ERR_FILE_INVALID;
    }
    unsigned long long payload = block - (c.has_crc ? 4 : 0);
    unsigned state = crc_table().update(0xFFFFFFFFu, tg, 20);
    long long off = pos + 20;
    if (c.extended) {
      // group and tag names (their lengths in the id bytes) follow the tag
      const int nb = tg[6] + tg[7];
      std::string names((size_t)nb, '\0');
      if (payload < (unsigned long long)nb || fread(&names[0], 1, (size_t)nb, f) != (size_t)nb) {
        delete h;
        return RCG_OSKAR_ERR_FILE_INVALID;
      }
      state = crc_table().update(state, names.data(), (size_t)nb);
      payload -= (unsigned long long)nb;
      off += nb;
    }
    c.payload_off = off;
    c.payload_bytes = payload;
    c.crc_state = state;
    if (fseeko(f, (off_t)(off + (long long)payload), SEEK_SET) != 0) {
      delete h;
      return RCG_OSKAR_ERR_FILE_INVALID;
    }
    if (c.has_crc) {
      unsigned char cb[4];
      if (fread(cb, 1, 4, f) != 4) {
        delete h;
        return RCG_OSKAR_ERR_FILE_INVALID;
      }
      c.crc = le32(cb);
    }
    h->chunks.push_back(c);
    pos = off + (long long)payload + (c.has_crc ? 4 : 0);
  }
  *out = h;
  return 0;
}

// ---------------------------------------------------------------------------
// Device side: one block's points reshuffled into the output layout.
// Extremes are kept as ordered 64-bit keys (the bit pattern of a double with
// the sign bit flipped for positive values and all bits flipped for
// negative ones, which orders as the values do) so that unsigned atomic
// max/min give the exact extremes.

__device__ __forceinline__ unsigned long long okey(double x)
{
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}

unsigned long long okey_host(double x)
{
  unsigned long long b;
  memcpy(&b, &x, 8);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}

double unkey_host(unsigned long long k)
{
  const unsigned long long b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFULL) : ~k;
  double x;
  memcpy(&x, &b, 8);
  return x;
}

// ext[0..6): max u, v, w, then min u, v, w (keys); bl[2 b], bl[2 b + 1]:
// max and min w of baseline b.  Thread e = (t * nch + c) * nb + b of the
// block's points (OskarBinReader.cpp:342-389): uvw scaled by the channel's
// 1 / wavelength, amplitudes moved as 8 doubles.
__global__ void k_oskar_reshuffle(const double *__restrict__ amp_blk,
                                  const double *__restrict__ u_blk,
                                  const double *__restrict__ v_blk,
                                  const double *__restrict__ w_blk,
                                  const double *__restrict__ inv_lambda, int nt_blk, int t0,
                                  int c0, int nch, int nb, int n_times, long long n_points,
                                  double *__restrict__ amps, double *__restrict__ uvws,
                                  unsigned long long *__restrict__ ext,
                                  unsigned long long *__restrict__ bl)
{
#pragma clang fp contract(off)
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)nt_blk * nch * nb;
  if (e >= n) return;
  const int b = (int)(e % nb);
  const int c = (int)((e / nb) % nch);
  const int t = (int)(e / ((long long)nb * nch));
  const long long j = b + (long long)nb * t;
  const double il = inv_lambda[c];
  const double u0 = u_blk[j] * il, v0 = v_blk[j] * il, w0 = w_blk[j] * il;
  const long long k = ((long long)n_times * b + t0 + t) * nch + c0 + c;
  if (k < 0 || k >= n_points) return;  // a block outside the header's dimensions
  uvws[3 * k] = u0;
  uvws[3 * k + 1] = v0;
  uvws[3 * k + 2] = w0;
  const double *src = amp_blk + 8 * e;
  double *dst = amps + 8 * k;
#pragma unroll
  for (int q = 0; q < 8; q++) dst[q] = src[q];
  atomicMax(&ext[0], okey(u0));
  atomicMax(&ext[1], okey(v0));
  atomicMax(&ext[2], okey(w0));
  atomicMin(&ext[3], okey(u0));
  atomicMin(&ext[4], okey(v0));
  atomicMin(&ext[5], okey(w0));
  atomicMax(&bl[2 * b], okey(w0));
  atomicMin(&bl[2 * b + 1], okey(w0));
}

struct DevMem {
  void *p = nullptr;
  ~DevMem()
  {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 16); }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

#define OSK_HIP(expr)                                                           \
  do {                                                                          \
    hipError_t e__ = (expr);                                                    \
    if (e__ != hipSuccess) {                                                    \
      rcg::set_error(RCG_ERR_HIP, "readAndReshuffle: %s failed: %s", #expr,     \
                     hipGetErrorString(e__));                                   \
      return RCG_OSKAR_ERR_DEVICE;                                              \
    }                                                                           \
  } while (0)

int reshuffle(const VisData *vdp, double *amps, double *uvws, Metrix *mp, WMaxMin *bl_ws)
{
  const rcg_oskar_file *h = vdp->h;
  const int nb = vdp->num_baselines, nch = vdp->num_channels;
  int ntags = 0, max_t = 0;
  int s = h->read(OSKAR_INT, GROUP_VIS_HEADER, NUM_TAGS_PER_BLOCK, 0, 0, &ntags, sizeof(int));
  if (!s) s = h->read(OSKAR_INT, GROUP_VIS_HEADER, MAX_TIMES_PER_BLOCK, 0, 0, &max_t, sizeof(int));
  if (s) return s;
  if (max_t <= 0 || nb < 0 || nch <= 0) return RCG_OSKAR_ERR_FILE_INVALID;
  const int n_blocks = (vdp->num_times + max_t - 1) / max_t;
  const long long per_blk = (long long)max_t * nb;
  const long long n_points = vdp->num_points;

  // host staging of one block, device copies of it and the full outputs
  std::vector<double> amp((size_t)(per_blk * nch * 8)), u((size_t)per_blk), v((size_t)per_blk),
      w((size_t)per_blk), il((size_t)nch);
  hipStream_t st = nullptr;
  OSK_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard()
    {
      if (s) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
      }
    }
  } sg{st};
  DevMem d_amp, d_u, d_v, d_w, d_il, d_amps, d_uvws, d_ext, d_bl;
  OSK_HIP(d_amp.alloc(amp.size() * 8));
  OSK_HIP(d_u.alloc(u.size() * 8));
  OSK_HIP(d_v.alloc(v.size() * 8));
  OSK_HIP(d_w.alloc(w.size() * 8));
  OSK_HIP(d_il.alloc(il.size() * 8));
  OSK_HIP(d_amps.alloc((size_t)n_points * 64));
  OSK_HIP(d_uvws.alloc((size_t)n_points * 24));
  OSK_HIP(d_ext.alloc(6 * 8));
  OSK_HIP(d_bl.alloc((size_t)std::max(nb, 1) * 16));
  // extremes start at -1e12 (maxima) and 1e12 (minima), OskarBinReader.cpp:268-279
  std::vector<unsigned long long> ext0(6), bl0((size_t)std::max(nb, 1) * 2);
  for (int q = 0; q < 3; q++) {
    ext0[q] = okey_host(-1e12);
    ext0[q + 3] = okey_host(1e12);
  }
  for (int b = 0; b < nb; b++) {
    bl0[2 * b] = okey_host(-1e12);
    bl0[2 * b + 1] = okey_host(1e12);
  }
  OSK_HIP(hipMemcpyAsync(d_ext.p, ext0.data(), 6 * 8, hipMemcpyHostToDevice, st));
  OSK_HIP(hipMemcpyAsync(d_bl.p, bl0.data(), bl0.size() * 8, hipMemcpyHostToDevice, st));

  for (int blk = 0; blk < n_blocks; blk++) {
    const int start = blk * ntags;
    if (start >= (int)h->chunks.size()) return RCG_OSKAR_ERR_TAG_OUT_OF_RANGE;
    int dim[6] = {0, 0, 0, 0, 0, 0};
    double fri[2] = {0.0, 0.0}, tri[2] = {0.0, 0.0};
    s = h->read(OSKAR_INT, GROUP_VIS_BLOCK, DIM_START_AND_SIZE, blk, start, dim, sizeof(dim));
    if (!s) s = h->read(OSKAR_DOUBLE, GROUP_VIS_BLOCK, FREQ_REF_INC_HZ, blk, start, fri, 16);
    if (!s) s = h->read(OSKAR_DOUBLE, GROUP_VIS_BLOCK, TIME_REF_INC_MJD_UTC, blk, start, tri, 16);
    if (s) return s;
    // 1 / wavelength per channel, the frequency summed channel by channel
    double cf = fri[0];
    for (int c = 0; c < nch; c++, cf += fri[1]) il[(size_t)c] = cf / SPEED_OF_LIGHT;
    const int t0 = dim[0], c0 = dim[1], nt = dim[2];
    s = h->read(OSKAR_DOUBLE_COMPLEX_MATRIX, GROUP_VIS_BLOCK, CROSS_CORRELATIONS, blk, start,
                amp.data(), amp.size() * 8);
    if (!s) s = h->read(OSKAR_DOUBLE, GROUP_VIS_BLOCK, BASELINE_UU, blk, start, u.data(), u.size() * 8);
    if (!s) s = h->read(OSKAR_DOUBLE, GROUP_VIS_BLOCK, BASELINE_VV, blk, start, v.data(), v.size() * 8);
    if (!s) s = h->read(OSKAR_DOUBLE, GROUP_VIS_BLOCK, BASELINE_WW, blk, start, w.data(), w.size() * 8);
    if (s) return s;
    if (nt < 0 || nt > max_t) return RCG_OSKAR_ERR_FILE_INVALID;
    const long long n_blk = (long long)nt * nch * nb;
    if (n_blk == 0) continue;
    OSK_HIP(hipMemcpyAsync(d_amp.p, amp.data(), (size_t)n_blk * 64, hipMemcpyHostToDevice, st));
    OSK_HIP(hipMemcpyAsync(d_u.p, u.data(), (size_t)nt * nb * 8, hipMemcpyHostToDevice, st));
    OSK_HIP(hipMemcpyAsync(d_v.p, v.data(), (size_t)nt * nb * 8, hipMemcpyHostToDevice, st));
    OSK_HIP(hipMemcpyAsync(d_w.p, w.data(), (size_t)nt * nb * 8, hipMemcpyHostToDevice, st));
    OSK_HIP(hipMemcpyAsync(d_il.p, il.data(), il.size() * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_oskar_reshuffle, dim3((unsigned)((n_blk + 255) / 256)), dim3(256), 0,
                       st, d_amp.as<double>(), d_u.as<double>(), d_v.as<double>(),
                       d_w.as<double>(), d_il.as<double>(), nt, t0, c0, nch, nb,
                       vdp->num_times, n_points, d_amps.as<double>(), d_uvws.as<double>(),
                       d_ext.as<unsigned long long>(), d_bl.as<unsigned long long>());
    OSK_HIP(hipGetLastError());
    // the staging buffers are refilled for the next block
    OSK_HIP(hipStreamSynchronize(st));
  }
  std::vector<unsigned long long> ext(6), blk_ext(bl0.size());
  OSK_HIP(hipMemcpyAsync(amps, d_amps.p, (size_t)n_points * 64, hipMemcpyDeviceToHost, st));
  OSK_HIP(hipMemcpyAsync(uvws, d_uvws.p, (size_t)n_points * 24, hipMemcpyDeviceToHost, st));
  OSK_HIP(hipMemcpyAsync(ext.data(), d_ext.p, 6 * 8, hipMemcpyDeviceToHost, st));
  OSK_HIP(hipMemcpyAsync(blk_ext.data(), d_bl.p, blk_ext.size() * 8, hipMemcpyDeviceToHost, st));
  OSK_HIP(hipStreamSynchronize(st));
  mp->maxu = unkey_host(ext[0]);
  mp->maxv = unkey_host(ext[1]);
  mp->maxw = unkey_host(ext[2]);
  mp->minu = unkey_host(ext[3]);
  mp->minv = unkey_host(ext[4]);
  mp->minw = unkey_host(ext[5]);
  for (int b = 0; b < nb; b++) {
    bl_ws[b].maxw = unkey_host(blk_ext[2 * (size_t)b]);
    bl_ws[b].minw = unkey_host(blk_ext[2 * (size_t)b + 1]);
  }
  return 0;
}

}  // namespace

extern "C" int mkFromFile(VisData *vdp, const char *filename)
{
  rcg::clear_error();
  if (!vdp) return RCG_OSKAR_ERR_MEMORY_NOT_ALLOCATED;
  rcg_oskar_file *h = nullptr;
  int s = open_index(filename, &h);
  if (s) return s;
  int amp_type = 0, n_times = 0, n_ch = 0, n_st = 0;
  s = h->read(OSKAR_INT, GROUP_VIS_HEADER, AMP_TYPE, 0, 0, &amp_type, sizeof(int));
  if (!s) s = h->read(OSKAR_INT, GROUP_VIS_HEADER, NUM_TIMES_TOTAL, 0, 0, &n_times, sizeof(int));
  if (!s) s = h->read(OSKAR_INT, GROUP_VIS_HEADER, NUM_CHANNELS, 0, 0, &n_ch, sizeof(int));
  if (!s) s = h->read(OSKAR_INT, GROUP_VIS_HEADER, NUM_STATIONS, 0, 0, &n_st, sizeof(int));
  if (s) {
    delete h;
    return s;
  }
  if ((amp_type & 0x0F) != OSKAR_DOUBLE) {  // as the reference: double amplitudes only
    delete h;
    return (int)0xdeadbeef;
  }
  VisData r{};
  r.h = h;
  const int d_tags[] = {FREQ_START_HZ, FREQ_INC_HZ, CHANNEL_BANDWIDTH_HZ, TIME_START_MJD_UTC,
                        TIME_INC_SEC, TIME_AVERAGE_SEC, PHASE_CENTRE, TELESCOPE_CENTRE};
  double *d_dst[] = {&r.freq_start_inc[0], &r.freq_start_inc[1], &r.channel_bandwidth_hz,
                     &r.time_start_inc[0], &r.time_start_inc[1], &r.time_average_sec,
                     r.phase_centre, r.telescope_centre};
  const int d_len[] = {1, 1, 1, 1, 1, 1, 2, 3};
  for (int q = 0; q < 8 && !s; q++)
    s = h->read(OSKAR_DOUBLE, GROUP_VIS_HEADER, d_tags[q], 0, 0, d_dst[q],
                sizeof(double) * (size_t)d_len[q]);
  if (s) {
    delete h;
    return s;
  }
  r.num_stations = n_st;
  r.num_channels = n_ch;
  r.num_times = n_times;
  r.num_baselines = n_st * (n_st - 1) / 2;
  r.num_times_baselines = n_times * r.num_baselines;
  r.num_points = r.num_times_baselines * n_ch;
  *vdp = r;
  return 0;
}

extern "C" void freeBinHandler(VisData *vdp)
{
  if (!vdp) return;
  delete vdp->h;
  vdp->h = nullptr;
}

extern "C" int readAndReshuffle(const VisData *vdp, double *amps, double *uvws, Metrix *mp,
                                WMaxMin *bl_ws)
{
  rcg::clear_error();
  if (!vdp || !vdp->h || !amps || !uvws || !mp || (!bl_ws && vdp->num_baselines > 0))
    return RCG_OSKAR_ERR_MEMORY_NOT_ALLOCATED;
  return reshuffle(vdp, amps, uvws, mp, bl_ws);
}

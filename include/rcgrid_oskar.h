/* rcgrid_oskar.h -- OSKAR 2.6 visibility files (binary format version 2)
 * read into the gridder's layout, the reshuffle on the GPU.
 *
 * Drop-in for the reference's reader MS4/dep/oskar/OskarBinReader.h:17-84
 * (called from MS4/misc/cppcycle.cpp:46-77 and bound in
 * MS4/dep/oskar/OskarBinReaderFFI.hsc): the same structures, function names,
 * arguments and results, over this library's own reader of the OSKAR binary
 * format (no oskar_binary library).
 *
 * mkFromFile opens `filename`, indexes its tags, checks the CRC-32C of every
 * header value it reads and fills *vdp; 0 on success, else an OSKAR status
 * (RCG_OSKAR_ERR_*), or (int)0xdeadbeef when the amplitudes are not double
 * precision, as the reference.  vdp->h is this library's file handle.
 *
 * readAndReshuffle reads every block (CRC-checked) and writes, for point
 * (baseline b, time t, channel c) at k = (num_times * b + t) * num_channels + c,
 *   amps[8k .. 8k+8)  the 2x2 complex (double) cross-correlation matrix,
 *   uvws[3k .. 3k+3)  u, v, w in wavelengths (metres x f_c / c, f_c summed
 *                     channel by channel from the block's start frequency);
 * *mp the extreme u, v, w (starting from -1e12 / 1e12) and bl_ws[b] the
 * extreme w of each baseline.  The reshuffle and the extremes run on the
 * current HIP device; amps/uvws/mp/bl_ws are host memory as in the
 * reference.  0 on success, else an OSKAR status, or RCG_OSKAR_ERR_DEVICE
 * when a HIP call fails (rcg_last_error_string() says which); on failure the
 * outputs are unspecified.
 *
 * freeBinHandler closes the file. */
#ifndef RCGRID_OSKAR_H
#define RCGRID_OSKAR_H

#ifdef __cplusplus
extern "C" {
#endif

#define SPEED_OF_LIGHT 299792458.0

/* OSKAR binary status codes (oskar_binary.h) this reader returns. */
#define RCG_OSKAR_ERR_OPEN_FAIL (-100)
#define RCG_OSKAR_ERR_SEEK_FAIL (-101)
#define RCG_OSKAR_ERR_READ_FAIL (-102)
#define RCG_OSKAR_ERR_FILE_INVALID (-106)
#define RCG_OSKAR_ERR_VERSION_UNKNOWN (-109)
#define RCG_OSKAR_ERR_TYPE_UNKNOWN (-110)
#define RCG_OSKAR_ERR_MEMORY_NOT_ALLOCATED (-114)
#define RCG_OSKAR_ERR_TAG_NOT_FOUND (-115)
#define RCG_OSKAR_ERR_TAG_OUT_OF_RANGE (-117)
#define RCG_OSKAR_ERR_CRC_FAIL (-118)
/* Not an OSKAR code: a HIP call of the reshuffle failed. */
#define RCG_OSKAR_ERR_DEVICE (-200)

typedef struct rcg_oskar_file rcg_oskar_file;

typedef struct VisData_tag {
  rcg_oskar_file *h;
  int num_baselines, num_channels, num_stations, num_times, num_times_baselines, num_points;
  double phase_centre[2], telescope_centre[3], freq_start_inc[2], time_start_inc[2],
      channel_bandwidth_hz, time_average_sec;
} VisData;

typedef struct Metrix_tag {
  double maxu, maxv, maxw, minu, minv, minw;
} Metrix;

typedef struct WMaxMin_tag {
  double maxw, minw;
} WMaxMin;

int mkFromFile(VisData *vdp, const char *filename);
void freeBinHandler(VisData *vdp);
int readAndReshuffle(const VisData *vdp, double *amps, double *uvws, Metrix *mp,
                     WMaxMin *bl_ws);

#ifdef __cplusplus
}
#endif

#endif /* RCGRID_OSKAR_H */

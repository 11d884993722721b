/*
 * dsaxfl.h -- C ABI of the MI355X (gfx950) search-side RFI flagger.
 *
 * The reference is one process, src/gpu_flagger.cu: per block of a
 * [nbeam][ntime][nchan] uint8 filterbank (64 beams x 16384 times x 1024
 * channels, gpu_flagger.cu:69-75) it rescales the block against reference
 * spectra, measures per-(beam, channel) statistics on the GPU, derives
 * channel and time masks on the CPU and overwrites the masked data on the GPU
 * (loop body gpu_flagger.cu:1294-1400; kernels :113-520; CPU functions
 * :530-790).  The entry points below are those kernels and CPU functions,
 * plus a plan that runs the loop body.  Conventions are those of dsaxbf.h:
 * 0 is success, negative values are DSAXBF_E* codes, dsaxbf_last_error()
 * describes the last failure, d_* are device pointers, stream is a
 * hipStream_t as void* (NULL = the default stream, or the plan's).
 *
 * Shapes: ntime a multiple of 4096 (32 segments of whole 128-sample bins, as
 * the reference's 32-thread reductions and TBIN assume), nchan a multiple of
 * 16, one beam's ntime * nchan bytes below 2 GiB; the channel-mask functions
 * use channels 128 .. nchan-129 (nchan > 277).  A plan is not thread-safe.
 */
#ifndef DSAXFL_H
#define DSAXFL_H

#include <stddef.h>
#include <stdint.h>

#include "dsaxbf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference constants (gpu_flagger.cu:69-83) ----------------------- */
#define DSAXFL_NBEAMS 64
#define DSAXFL_NTIMES 16384
#define DSAXFL_NCHAN 1024
#define DSAXFL_MN 48.0          /* scaled mean and replacement level       */
#define DSAXFL_SIG 6.0          /* scaled standard deviation               */
#define DSAXFL_TBIN 128         /* ppspec time bin                          */
#define DSAXFL_FBIN 8           /* simple_channflag frequency bin           */
#define DSAXFL_ZERO_CHANNELS 128 /* band edges left out of the masks       */

/* ---- device kernels (asynchronous on stream) ---------------------------- */

/* calc_spectrum, calc_varspec, calc_maxspec and calc_ppspec
 * (gpu_flagger.cu:174-375) in one pass over the block; each output is
 * [nbeam][nchan] float and may be NULL.  d_spec_in is calc_varspec's
 * spectrum argument (NULL: the block's own mean).  spectrum, maxspec and
 * ppspec equal the reference's bit for bit (integer sums); varspec is the
 * exact variance rounded once, which the reference's float sums approximate
 * to < 1e-5 relative.  ppspec is what the reference's unsynchronised
 * max-then-min tree leaves: the largest 128-sample bin mean over the block
 * minus the smallest of segments 16..31's largest bin means (segment s =
 * times [s*ntime/32, (s+1)*ntime/32)). */
int dsaxfl_stats(const uint8_t *d_data, int nbeam, int ntime, int nchan, const float *d_spec_in,
                 float *d_spec, float *d_var, float *d_max, float *d_pp, void *stream);

/* scaley (gpu_flagger.cu:381-393) in place:
 * x <- (uint8)(rne(2 * fma(x - spec0, SIG / sqrtf(var0), MN)) / 2), and, in
 * the same pass, the statistics of the scaled block (outputs may be NULL). */
int dsaxfl_scaley(uint8_t *d_data, int nbeam, int ntime, int nchan, const float *d_spec0,
                  const float *d_var0, float *d_spec, float *d_var, float *d_max, float *d_pp,
                  void *stream);

/* make_ts (gpu_flagger.cu:413-452): ts[beam][time] = row sum / nchan, the
 * exact integer sum over all nchan channels divided once.  The reference's 32
 * threads take nchan/32 channels each (:419-429), so it sums them all only
 * when nchan is a multiple of 32 (its NCHAN_P is 1024); other widths drop the
 * last nchan % 32 channels there and not here. */
int dsaxfl_make_ts(const uint8_t *d_data, int nbeam, int ntime, int nchan, float *d_ts, void *stream);

/* flag (gpu_flagger.cu:458-474): every time of the n_mask columns
 * d_midx[i] = beam*nchan + channel set to (uint8)(MN * bpwr[beam]). */
int dsaxfl_flag(uint8_t *d_data, int nbeam, int ntime, int nchan, const int *d_midx, int n_mask,
                const float *d_bpwr, void *stream);

/* flag by a [nbeam][nchan] byte mask (nonzero = flagged) and, when d_ts is
 * not NULL, make_ts of the flagged block, in one pass (flag then make_ts in
 * the reference loop, gpu_flagger.cu:1379-1384). */
int dsaxfl_flag_mask(uint8_t *d_data, int nbeam, int ntime, int nchan, const uint8_t *d_mask,
                     const float *d_bpwr, float *d_ts, void *stream);

/* flagts (gpu_flagger.cu:478-494): channels 128 .. nchan-129 of the n_mask
 * rows d_midx[i] = beam*ntime + time set to (uint8)(MN * bpwr[beam]). */
int dsaxfl_flagts(uint8_t *d_data, int nbeam, int ntime, int nchan, const int *d_midx, int n_mask,
                  const float *d_bpwr, void *stream);

/* sumpulse (gpu_flagger.cu:398-406): x <- (uint8)(rne(2 (x + summand)) / 2). */
int dsaxfl_sumpulse(uint8_t *d_data, size_t n, const float *d_summand, void *stream);

/* genrand (gpu_flagger.cu:499-520): n noise bytes (uint8)(rne(2 (6 g + 48)) / 2)
 * with g standard normal by Box-Muller of two 14-bit uniforms, as the
 * reference, drawn from this library's counter-based generator (splitmix64 of
 * seed * 2^32 + i) instead of curand: statistically the same, not the same
 * bytes. */
int dsaxfl_genrand(uint8_t *d_repval, size_t n, uint32_t seed, void *stream);

/* The noise replacement (gpu_flagger.cu:1319-1320,1356-1358): every beam of
 * the block <- the [ntime][nchan] noise block. */
int dsaxfl_noise(uint8_t *d_data, int nbeam, int ntime, int nchan, const uint8_t *d_repval,
                 void *stream);

/* fix_zspec (gpu_flagger.cu:113-168): entry i < nbc of d_s0 and d_v0 becomes
 * the median of the naver stacked blocks d_s0[k*nbc + i] (the value with
 * exactly naver/2 others <= it; none with ties, and then block 0's stays). */
int dsaxfl_fix_zspec(float *d_s0, float *d_v0, int nbc, int naver, void *stream);

/* ---- CPU functions (gpu_flagger.cu:530-790) ----------------------------- */
/* channflag, simple_channflag and simple_tsflag split their beams over up to
 * 8 host threads (each beam's arithmetic unchanged, so the masks do not
 * depend on the split); they keep no state and may be called concurrently.
 * Like the reference's, they only set flagged entries to 1 and leave the
 * others as they are, so several calls accumulate into one mask: zero it
 * first, as the reference's loop does (gpu_flagger.cu:1339-1347,1388-1390). */

/* medval: element (n+1)/2-1 of a sorted copy (the lower median). */
float dsaxfl_medval(const float *a, int n);
/* channflag: 21-channel forward median filter, median and MAD over channels
 * 128..nchan-129, mask[b][c] = 1 where |residual| > thr * MAD, and the
 * permanent channels perm[0..nperm-1] inside that range. */
void dsaxfl_channflag(const float *spec, int nbeam, int nchan, float thr, const int *perm, int nperm,
                      int *mask);
/* simple_channflag, its FBIN loop (which adds spec[j] FBIN times) kept. */
void dsaxfl_simple_channflag(const float *spec, int nbeam, int nchan, float thr, int *mask);
/* simple_tsflag: median and MAD over the first ntime/16 samples of each
 * beam's series, mask[b][t] = 1 where ts - median > thr * MAD. */
void dsaxfl_simple_tsflag(const float *ts, int nbeam, int ntime, float thr, int *mask);
/* gather_mask / gather_tsmask: indices of the entries equal to 1. */
int dsaxfl_gather_mask(const int *mask, long n, int *idx);
/* calc_bpwr: bpwr[b] = (sum over channels of spec[b]) / (nchan - 256) / MN. */
void dsaxfl_calc_bpwr(const float *spec, int nbeam, int nchan, float *bpwr);
/* median_calc: each channel's values sorted over beams, in place. */
void dsaxfl_median_calc(float *arr, int nbeam, int nchan);

/* ---- plan: the flagger loop (gpu_flagger.cu:1233-1460) ------------------ */
typedef struct dsaxfl_config {
  int nbeam, ntime, nchan;
  float thresh;       /* -t, channel masks (5)                             */
  float tsthresh;     /* -s, time mask (10)                                */
  float mod_thresh;   /* -q, total-power jump that triggers noise (5e-4)   */
  int pwr;            /* -p, replacement level from the spectrum (0)       */
  int mkrand;         /* -m, replace the whole block with noise (0)        */
  int tsflag;         /* time-series flagging (set by -s; 0)               */
  int nperm;          /* permanent channels (1: channel 10)                */
  int perm[64];
  uint32_t seed;      /* noise generator seed                              */
  int device;
} dsaxfl_config;

typedef struct dsaxfl_step_info {
  float tpwr;         /* total power of the scaled block's spectrum        */
  int replaced;       /* 1: the power jump replaced the block with noise   */
  int n_mask;         /* flagged (beam, channel) columns (-1 if replaced)  */
  int n_tsmask;       /* flagged (beam, time) rows (-1 if replaced)        */
} dsaxfl_step_info;

typedef struct dsaxfl_plan dsaxfl_plan;

/* The reference's sizes and defaults (gpu_flagger.cu:69-92,932-955). */
void dsaxfl_config_default(dsaxfl_config *cfg);
/* Allocates the plan's device buffers and draws the noise block. */
int dsaxfl_create(dsaxfl_plan **plan, const dsaxfl_config *cfg);
void dsaxfl_destroy(dsaxfl_plan *plan);
void *dsaxfl_get_stream(dsaxfl_plan *plan);
/* The reference spectra spec0 / var0 ([nbeam][nchan]), as read from
 * specvar0.dat (gpu_flagger.cu:1420-1432). */
int dsaxfl_set_reference(dsaxfl_plan *plan, const float *h_spec0, const float *h_var0);
int dsaxfl_get_reference(dsaxfl_plan *plan, float *h_spec0, float *h_var0);
/* The start-up average (gpu_flagger.cu:1437-1458): spectrum and varspec of
 * naver unscaled device blocks, fix_zspec (naver > 1), median_calc. */
int dsaxfl_calibrate(dsaxfl_plan *plan, const uint8_t *const *d_blocks, int naver);
/* Previous block's total power for the jump test (0 after calibration, as in
 * the reference, so the first block after start-up is noise-replaced). */
int dsaxfl_set_total_power(dsaxfl_plan *plan, float prev_tpwr);
/* One block in place on the device (gpu_flagger.cu:1294-1400 with
 * started = 1, no pulse injection): scaley, statistics, the power-jump test,
 * channel masks on the CPU, flag (or noise with mkrand), time masks and
 * flagts with tsflag.  Synchronous; info may be NULL. */
int dsaxfl_step(dsaxfl_plan *plan, uint8_t *d_data, dsaxfl_step_info *info);

#ifdef __cplusplus
}
#endif
#endif

/*
 * bf_launch.h -- internal C ABI between the C host layer (host/bf_plan.c)
 * and the HIP kernels (csrc/ *.hip).  Not part of the public API.
 */
#ifndef BF_LAUNCH_H
#define BF_LAUNCH_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Derived dimensions shared by every kernel. */
typedef struct bfk_dims {
  int nant;       /* antennas in the block                                */
  int nant_pad;   /* antennas rounded up to 64 (MFMA K chunk)             */
  int nkc;        /* nant_pad / 64                                        */
  int nchan;      /* fine channels                                        */
  int ngroup;     /* nchan / 8                                            */
  int npkt;       /* packet-times (2 samples each)                        */
  int nsamp;      /* 2 * npkt                                             */
  int nbeam;      /* beams                                                */
  int nbeam_pad;  /* beams rounded up to the beam set (256)               */
  int nbt;        /* nbeam_pad / 32 beam tiles                            */
  int tavg;       /* samples per output                                   */
  int nout;       /* nsamp / tavg                                         */
  int ncal;       /* weight channels: ngroup (group mode) or nchan (fine) */
  int wmode;      /* 0 group, 1 fine                                      */
  int amode;      /* 0 theta, 1 sin(theta)                                */
  int prec;       /* 0 FP16, 1 FP8 hi+lo, 2 FP6 hi+lo, 3 FP6 hi+mid+lo    */
  int order;      /* 0 capture, 1 split -r                                */
  int stuff;      /* 0 none, 1 antennas, 2 incoherent beam 0              */
  int variant;    /* fused-kernel variant, BFK_VAR_*                      */
  int layout;     /* output layout: 0 [t][beam][g], 1 [beam][t][g]        */
} bfk_dims;

/* Variants of the fused beamformer (same bits, different speed):
 * CLASSIC_*    one tile per workgroup, two workgroups per CU; 256, 128 or 64
 *              beams per tile (128S: 2 beam tiles per wave, wave pairs split a
 *              tile's samples; 128: one beam tile per wave; 64: one beam tile
 *              per wave, split);
 * SHARED_256   waves of a tile that unpack each unit once into LDS
 *              operands all of them read (4 waves to 64 antennas, 8 above).
 * Value 1 (a persistent software-pipelined schedule, never faster) is
 * retired: no shape supports it. */
enum {
  BFK_VAR_CLASSIC_256 = 0,
  BFK_VAR_CLASSIC_128S = 2,
  BFK_VAR_CLASSIC_128 = 3,
  BFK_VAR_CLASSIC_64 = 4,
  BFK_VAR_SHARED_256 = 5,
  BFK_NVAR = 6
};

/* Bytes of the weight fragment buffer and of the per-(channel,pol,beam)
 * weight scale buffer for dims. */
size_t bfk_wfrag_bytes(const bfk_dims *d);
size_t bfk_wscale_bytes(const bfk_dims *d);

/* Beam weights from geometry, FP64 phase (device pointers). */
int bfk_weights(const bfk_dims *d, const float *d_antpos, const float *d_cal,
                const double *d_freqs, const double *d_theta, void *d_wfrag,
                float *d_wscale, void *stream);

/* Per-32-sample beam power sums written by every beamforming launch:
 * [nbeam][nsamp/32][ngroup] floats. */
size_t bfk_bpart_bytes(const bfk_dims *d);

/* Fused unpack -> MFMA beam GEMM -> |B|^2 -> average -> requantise.
 * d_power is optional; d_bpart receives the per-tile beam power sums. */
int bfk_beamform(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag,
                 const float *d_wscale, const float *d_scale, uint8_t *d_out,
                 float *d_power, float *d_bpart, void *stream);

/* 1 if variant v (BFK_VAR_*) can run dims d, else 0. */
int bfk_variant_supported(const bfk_dims *d, int v);

/* Batched classic launches: all blocks of a batch in one launch per kernel
 * instantiation (precision, tavg, antenna chunks, beam tiles per wave).
 * prepare builds the device job tables for n blocks (dims[i] with its own
 * variant; buffers as for bfk_beamform); launch enqueues the launches (no
 * allocation, no host synchronisation: capturable in a graph). */
typedef struct bfk_batch_bufs {
  const uint8_t *d_in;
  const void *d_wfrag;
  const float *d_wscale;
  const float *d_scale;
  uint8_t *d_out;
  float *d_power;       /* optional */
  float *d_bpart;
  float *d_beam_power;  /* optional: the block's per-beam power sums */
} bfk_batch_bufs;

/* One block's beam-power reduction in a batch: out[b] = sum of the n
 * partial sums bpart[b*n .. b*n + n-1] for b < nbeam. */
typedef struct bfk_bp_job {
  const float *bpart;
  float *out;
  int nbeam, n;
} bfk_bp_job;
int bfk_batch_prepare(const bfk_dims *dims, const bfk_batch_bufs *bufs, int n, void **handle);
int bfk_batch_launch(void *handle, void *stream);
int bfk_batch_launches(void *handle);   /* kernel launches per bfk_batch_launch */
void bfk_batch_free(void *handle);

/* Antenna-stuffing (-s) and incoherent-beam (-q) outputs (rewrites the
 * d_bpart entries of the beams it writes). */
int bfk_stuff(const bfk_dims *d, const uint8_t *d_in, const float *d_scale,
              uint8_t *d_out, float *d_power, float *d_bpart, void *stream);

/* beam_power[b] = sum of d_bpart[b][.][.] (deterministic order). */
int bfk_beam_power(const bfk_dims *d, const float *d_bpart, float *d_beam_power,
                   void *stream);

/* Name of the last HIP error seen by a launcher (thread-local). */
const char *bfk_last_error(void);

#ifdef __cplusplus
}
#endif
#endif

/*
 * dsaxbf.h -- C ABI of the MI355X (gfx950) DSA X-engine coherent beamformer.
 *
 * The reference beamformer is a single process, src/dsaX_beamformer.cu, that
 * reads psrdada ring REORDER_BLOCK_KEY2 (0xbada) and writes BF_BLOCK_KEY
 * (0xdcda).  It has no library interface of its own; every entry point below
 * replaces one step of its per-block loop (dsaX_beamformer.cu:939-1090) and is
 * what a host program (the dsaX_beamformer-compatible CLI in this package, or a
 * cgo/ctypes caller) binds.  Each declaration cites the reference code it
 * stands in for.
 *
 * Conventions
 *   - Return value 0 is success; negative values are DSAXBF_E* codes.  The
 *     library never prints and never exits; dsaxbf_last_error() gives a
 *     message for the last failure on the calling thread.
 *   - A plan owns its device buffers and weights.  Callers own host blocks.
 *   - A plan is not thread-safe: use one plan per host thread.
 *   - Pointers named d_* are device pointers (HBM); h_* are host pointers.
 *   - "stream" is a hipStream_t passed as void* (NULL = the plan's stream).
 *
 * Data layout (all byte orders are those of the reference pipeline)
 *   input  : capture order [npkt][nant][nchan][2 times][2 pols], one byte per
 *            complex sample, low nibble = real, high nibble = imaginary, both
 *            4-bit two's complement (dsaX_beamformer.cu:182-191,
 *            utils/gen_packet.py:196-206).  With input_order =
 *            DSAXBF_ORDER_SPLIT_R the block is [npkt][nchan/8][nant][8][2][2]
 *            (what dsaX_split -r writes, dsaX_split.c:86-101).
 *   output : [nsamp/tavg][nbeam][nchan/8] uint8 (BF_BLOCK_KEY layout,
 *            dsaX_beamformer.cu:150-156,986-996), nsamp = 2*npkt.
 *   power  : same index order as output, float32, the averaged |B|^2 before
 *            scaling (the reference's adder input sum, dsaX_beamformer.cu:139-147).
 */
#ifndef DSAXBF_H
#define DSAXBF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---------------------------------------------------- */
#define DSAXBF_OK        0
#define DSAXBF_EINVAL   -1   /* bad argument or unsupported shape            */
#define DSAXBF_ENOMEM   -2   /* host or device allocation failed              */
#define DSAXBF_EHIP     -3   /* a HIP runtime call failed                     */
#define DSAXBF_EIO      -4   /* file missing, short or malformed              */
#define DSAXBF_ESTATE   -5   /* call out of order (e.g. execute before weights)*/
#define DSAXBF_ENODEV   -6   /* no usable gfx950 device / kernels not loaded  */

/* ---- reference constants (dsaX_def.h:16,22,28-29,72,80-84) ---------- */
#define DSAXBF_REORDER_BLOCK_KEY2 0x0000bada
#define DSAXBF_BF_BLOCK_KEY       0x0000dcda
#define DSAXBF_CVAC               299792458.0
#define DSAXBF_NPACKETS           2048   /* packet-times per block           */
#define DSAXBF_NW                 48     /* output channels per 384          */
#define DSAXBF_NANT_FILE          64     /* antennas in the weights file     */
#define DSAXBF_NBP                8      /* AGC history length               */
#define DSAXBF_FAVG               8      /* fine channels per output channel */

/* ---- enumerations ---------------------------------------------------- */
enum dsaxbf_precision {
  DSAXBF_PREC_FP16    = 0,  /* FP16 weights x FP16 voltages, FP32 accumulate   */
  DSAXBF_PREC_FP8SPLIT = 1, /* FP8 E4M3 voltages x (hi+lo) FP8 weights, FP32   */
  DSAXBF_PREC_FP6SPLIT = 2, /* FP6 E2M3 voltages x (hi+lo) FP6 weights, FP32
                               accumulate, at twice the FP8 matrix rate (shared
                               schedule only).  No counterpart in the
                               reference.                                     */
  DSAXBF_PREC_FP6SPLIT3 = 3 /* FP6 E2M3 voltages x (hi+mid+lo) FP6 weights,
                               FP32 accumulate: three 32-cycle FP6 passes where
                               FP16 takes two 32-cycle passes' worth, with
                               weights good to ~12 bits (shared schedule, up
                               to 128 antennas).  No counterpart in the
                               reference.                                     */
};
enum dsaxbf_weight_mode {
  DSAXBF_WEIGHTS_GROUP = 0, /* one weight per 8-channel group at freqs[8g+4]
                               (dsaX_beamformer.cu:441)                       */
  DSAXBF_WEIGHTS_FINE  = 1  /* one weight per fine channel                    */
};
enum dsaxbf_angle_mode {
  DSAXBF_ANGLE_THETA = 0,   /* phase uses theta itself (dsaX_beamformer.cu:440-441) */
  DSAXBF_ANGLE_SIN   = 1    /* phase uses sin(theta) (dsaX_beamformer_offline.cu:422) */
};
enum dsaxbf_stuff {
  DSAXBF_STUFF_NONE = 0,    /* normal beams                                   */
  DSAXBF_STUFF_ANTS = 1,    /* -s: beam b carries one antenna's power
                               (dsaX_beamformer.cu:326-337,400-404)           */
  DSAXBF_STUFF_IB   = 2     /* -q: beam 0 carries the incoherent sum
                               (dsaX_beamformer.cu:307-323,405-414)           */
};
enum dsaxbf_input_order {
  DSAXBF_ORDER_CAPTURE = 0, /* [pkt][ant][chan][t][pol] (dsaX_split without -r) */
  DSAXBF_ORDER_SPLIT_R = 1  /* [pkt][chan/8][ant][8][t][pol] (dsaX_split -r)    */
};
enum dsaxbf_output_layout {
  DSAXBF_LAYOUT_TIME_MAJOR = 0, /* [nsamp/tavg][nbeam][nchan/8]: BF_BLOCK_KEY
                                   (dsaX_beamformer.cu:150-156)                */
  DSAXBF_LAYOUT_BEAM_MAJOR = 1  /* [nbeam][nsamp/tavg][nchan/8]: beams
                                   64s..64s+63 are the buffer dsaX_dbnic's
                                   transmit thread s builds with its transpose
                                   (dsaX_dbnic.c:111-116), so the corner turn
                                   sends slices of the block as they are     */
};

/* ---- configuration --------------------------------------------------- */
typedef struct dsaxbf_config {
  int nant;         /* antennas in the input block, 1..256                   */
  int nchan;        /* fine channels in the block, multiple of 8             */
  int npkt;         /* packet-times per block (2 samples each), mult. of 32  */
  int nbeam;        /* beams formed, 1..1024                                  */
  int tavg;         /* samples averaged per output, one of 8,16,32,64        */
  int precision;    /* enum dsaxbf_precision                                 */
  int weight_mode;  /* enum dsaxbf_weight_mode                               */
  int angle_mode;   /* enum dsaxbf_angle_mode                                */
  int stuffants;    /* enum dsaxbf_stuff                                     */
  int input_order;  /* enum dsaxbf_input_order                               */
  int device;       /* HIP device ordinal                                     */
  int output_layout;/* enum dsaxbf_output_layout                             */
} dsaxbf_config;

typedef struct dsaxbf_plan dsaxbf_plan;

/* Fill cfg with the reference's own compiled-in shape: 2048 packets, 384
 * channels, 256 beams, 8-sample averaging, group weights, theta angles,
 * capture order, FP16 (dsaX_beamformer.cu:139-156,420-449; dsaX_def.h:72,80). */
void dsaxbf_config_default(dsaxbf_config *cfg, int nant);

/* Bytes of one input block and one output block for cfg. */
size_t dsaxbf_input_bytes(const dsaxbf_config *cfg);
size_t dsaxbf_output_bytes(const dsaxbf_config *cfg);

/* Message describing the most recent failure on this thread ("" if none). */
const char *dsaxbf_last_error(void);

/* ---- plan lifecycle (replaces dsaX_beamformer.cu:880-925 allocations) --- */
int  dsaxbf_create(dsaxbf_plan **plan, const dsaxbf_config *cfg);
void dsaxbf_destroy(dsaxbf_plan *plan);     /* dsaX_beamformer.cu:1092-1122 */
int  dsaxbf_get_config(const dsaxbf_plan *plan, dsaxbf_config *cfg);
void *dsaxbf_get_stream(dsaxbf_plan *plan); /* the plan's hipStream_t        */

/* ---- weights (replaces init_weights + calc_weights, dsaX_beamformer.cu:
 *      420-449,559-602,948-951) ------------------------------------------ */

/* Read the reference weights file: 64 float antenna eastings (m) followed by
 * cal[64 ant][48 group][2 pol][re,im] float32, each complex normalised as
 * w/|w|^2 (zero stays zero), then zero every antenna listed in flagfile (one
 * integer per line; NULL or "nofile" = none).  Unlike the reference, a missing
 * or short file, or a flag index outside 0..63, is an error (DSAXBF_EIO /
 * DSAXBF_EINVAL) instead of silently leaving garbage.
 * antpos: 64 floats out; cal: 64*48*2*2 floats out. */
int dsaxbf_read_weights_file(const char *fnam, const char *flagfile,
                             float *antpos, float *cal);

/* Compute the beam weights on the device: W[b][a] = exp(i*phi)*cal[a], with
 * phi = -2*pi*f*ang_b*x_a/c evaluated in FP64 (ang_b = theta_b or
 * sin(theta_b) per angle_mode), then converted to the plan's MFMA operand
 * format (FP16, or FP8 E4M3 hi+lo with a power-of-two scale per beam).
 *   antpos   : [nant] eastings in metres
 *   cal      : [nant][ncal][2 pol][re,im], ncal = nchan/8 (group mode) or
 *              nchan (fine mode); already normalised, flagged antennas zero
 *   freqs_hz : [nchan] sky frequency of each fine channel (Hz)
 *   theta_rad: [nbeam] beam angles (radians); NULL selects the reference
 *              fan theta_b = sep_arcmin*(nbeam/2-1-b)*pi/10800
 *              (dsaX_beamformer.cu:440, sep = 1 arcmin there).
 * Host pointers.  Synchronous.  The new weights are written on the plan's
 * stream, after the executes queued there; executes queued on other streams
 * read the same weight buffer, so the caller lets them finish first. */
int dsaxbf_set_weights(dsaxbf_plan *plan, const float *antpos, const float *cal,
                       const double *freqs_hz, const double *theta_rad,
                       double sep_arcmin);

/* The beam weights exactly as the kernels apply them, read back from the
 * device and decoded on the host: the plan's precision encoding (binary16,
 * or the FP8 / FP6 split parts with their per-lane scales) times the power
 * factor of its row, i.e. the weights the reference's calc_weights leaves in
 * its half-precision host arrays before they go to the GPU
 * (dsaX_beamformer.cu:442-446, 948-951).  w: [ncal][2 pol][nbeam][nant]
 * complex as (re, im) doubles, ncal = nchan/8 (group) or nchan (fine).
 * For validation (an oracle fed the path's own weights); synchronous. */
int dsaxbf_get_weights(dsaxbf_plan *plan, double *w);

/* Reference channel frequencies: freqs[i] = (fch1_mhz - i*250/8192)*1e6
 * (dsaX_beamformer.cu:799-800). */
void dsaxbf_reference_freqs(double fch1_mhz, int nchan, double *freqs_hz);

/* ---- one block (replaces the 128-int loop, dsaX_beamformer.cu:955-1003,
 *      i.e. promoter + beamformer + badder + adder) ---------------------- */

/* Beamform one block already resident in device memory.
 *   d_in        : input block (dsaxbf_input_bytes bytes)
 *   d_scale     : [nbeam] float per-beam requantisation scale (the AGC
 *                 output; d_bp in the reference)
 *   d_out       : output block, uint8 (dsaxbf_output_bytes bytes)
 *   d_power     : optional float power before scaling, same shape as d_out
 *   d_beam_power: optional [nbeam] float, sum of power over the block (the
 *                 AGC input; d_added + calc_allbp in the reference)
 * Asynchronous on stream. */
int dsaxbf_execute(dsaxbf_plan *plan, const uint8_t *d_in, const float *d_scale,
                   uint8_t *d_out, float *d_power, float *d_beam_power,
                   void *stream);

/* dsaxbf_execute that allocates nothing and uses no plan-owned scratch.  No
 * counterpart in the reference: for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The fused kernel's per-32-sample beam-power partial sums go to the
 * library's static device array (DSAXBF_NOALLOC_SCRATCH_BYTES per device,
 * held until the process ends) instead of the plan's buffer; every call on a
 * GPU shares that GPU's array, so calls that overlap on one GPU race (run
 * them in order on one stream).  Same arguments, results and asynchrony as
 * dsaxbf_execute.  A block whose partial sums exceed the array
 * (nbeam * nsamp/32 * nchan/8 * 4 bytes; the full 256-beam DSA block needs
 * 6.3 MB) is DSAXBF_EINVAL.  Ask for d_beam_power in the call itself:
 * dsaxbf_beam_power reads the plan's own buffer, which this call leaves
 * alone. */
#define DSAXBF_NOALLOC_SCRATCH_BYTES (8u << 20)
int dsaxbf_execute_noalloc(dsaxbf_plan *plan, const uint8_t *d_in, const float *d_scale,
                           uint8_t *d_out, float *d_power, float *d_beam_power,
                           void *stream);

/* Per-beam power sum of the block most recently run by dsaxbf_execute on
 * this plan: d_beam_power[b] = sum over outputs and groups of the power
 * (badder + calc_allbp, dsaX_beamformer.cu:78-112,482-495), from partial
 * sums the beamforming kernel leaves in the plan.  dsaxbf_execute does this
 * itself when given d_beam_power.  Same stream as that execute. */
int dsaxbf_beam_power(dsaxbf_plan *plan, float *d_beam_power, void *stream);

/* Variants of the fused kernel: three schedules and three tile shapes.  All
 * give bit-identical results; which is fastest depends on the shape (few beams
 * favour small tiles) and the GPU.
 *   CLASSIC_256    one tile (256 beams x 64 samples x 8 channels) per
 *                  workgroup, two workgroups per CU (nant <= 64);
 *   CLASSIC_128S   128-beam tiles: 64 beams per wave, wave pairs split the
 *                  tile's samples (nant <= 64; default for 65-128 beams);
 *   CLASSIC_128    128-beam tiles: 32 beams per wave (default for 65-128
 *                  beams when nant > 64);
 *   CLASSIC_64     64-beam tiles: 32 beams per wave, samples split (default
 *                  for <= 64 beams);
 *   SHARED_256     256-beam tiles whose waves unpack each unit once into LDS
 *                  operands they all read: 4 waves of 64 beams (nant <= 64) or
 *                  8 waves of 32 beams (nant > 64) (default above 128 beams).
 * Value 1 is retired (a persistent schedule that never won): no shape
 * supports it, and dsaxbf_set_variant rejects it.
 * No counterpart in the reference (its five kernels per int have one fixed
 * launch shape, dsaX_beamformer.cu:955-1003). */
enum {
  DSAXBF_VARIANT_CLASSIC_256 = 0,
  DSAXBF_VARIANT_CLASSIC_128S = 2,
  DSAXBF_VARIANT_CLASSIC_128 = 3,
  DSAXBF_VARIANT_CLASSIC_64 = 4,
  DSAXBF_VARIANT_SHARED_256 = 5,
  DSAXBF_NVARIANT = 6
};

/* Select the variant used by later executes on this plan; DSAXBF_EINVAL if
 * it cannot run the plan's shape. */
int dsaxbf_set_variant(dsaxbf_plan *plan, int variant);
int dsaxbf_get_variant(const dsaxbf_plan *plan);

/* Autotuner: times every variant the shape supports (the plan's own device
 * buffers, current weights; reps launches each after one warm-up) and keeps
 * the fastest.  ms, if not NULL, receives DSAXBF_NVARIANT mean launch times in
 * milliseconds (-1 for a variant that cannot run this shape).  Synchronous;
 * requires dsaxbf_set_weights (or -s mode).  No counterpart in the reference. */
int dsaxbf_autotune(dsaxbf_plan *plan, int reps, float *ms);

/* Same as dsaxbf_execute with host pointers: copies the block in, runs it,
 * copies the output (and optional power / beam_power) back, then waits.
 * h_scale is [nbeam] floats on the host.  This is the psrdada-facing call
 * (ipcio_open_block_read pointer in, ipcio_write buffer out). */
int dsaxbf_process_block(dsaxbf_plan *plan, const uint8_t *h_in,
                         const float *h_scale, uint8_t *h_out,
                         float *h_power, float *h_beam_power);

/* ---- AGC (dsaX_beamformer.cu:482-530,1005-1066) ----------------------- */
typedef struct dsaxbf_agc {
  int    nbeam;
  int    ngroup;          /* output channels (48 in the reference)           */
  int    nints;           /* 32-sample integrations per block (128)          */
  int    enabled;         /* 0 = -g: hold the first block's median scale     */
  int    started;         /* 0 before the first block, then 1, then 2        */
  int    bpctr;           /* blocks seen                                      */
  float  medbp;           /* median of the first block's beam powers         */
  float *many_bp;         /* [DSAXBF_NBP][nbeam] history ring                */
  float *bp;              /* [nbeam] work                                     */
} dsaxbf_agc;

int  dsaxbf_agc_init(dsaxbf_agc *agc, int nbeam, int ngroup, int nints, int enabled);
void dsaxbf_agc_free(dsaxbf_agc *agc);
/* Feed one block's raw per-beam power sum; writes the scale to use for the
 * NEXT block into scale_out[nbeam].  *zero_output is set to 1 when the block
 * just processed must be written as zeros (the first block, :1021-1033). */
int  dsaxbf_agc_update(dsaxbf_agc *agc, const float *beam_power_raw,
                       float *scale_out, int *zero_output);

/* ---- requantisation of one value (dsaX_beamformer.cu:150-156) ---------- */
/* (uint8)(rint(x*scale) / 2) with C integer division and wrap modulo 256. */
uint8_t dsaxbf_requant(float x, float scale);

/* ---- synthetic inputs (one seeded generator for every harness) --------- */
/* Fill a block with 4-bit complex Gaussian samples (sigma per component),
 * rounded and clipped to [-clip, clip], in capture order; seeded
 * (splitmix64 -> xoshiro256**, Box-Muller in FP64).  If tone_snr > 0 a
 * narrowband tone from direction tone_theta (radians, sin geometry) is added
 * before quantisation to channels tone_chan[0..ntone-1], with per-antenna
 * delays from antpos and freqs, divided by the complex antenna gain
 * gain[nant][re,im] (NULL = 1) so that weights built with cal = gain steer
 * the tone coherently.  SNR is tone power over noise power per sample.
 * Multithreaded over packets; the result does not depend on thread count. */
int dsaxbf_synth_block(const dsaxbf_config *cfg, uint64_t seed, double sigma,
                       int clip, double tone_snr, double tone_theta,
                       const int *tone_chan, int ntone, const float *antpos,
                       const float *gain, const double *freqs_hz,
                       uint8_t *h_block);

/* Geometry of the benchmark workloads: antenna eastings uniform on
 * [0, length_m] (seeded), unit-amplitude-ish gains g_a = (1 +- amp_jitter)
 * * exp(i*U[0,2pi)) written as cal[nant][ncal][2][2] (same gain for every
 * channel and pol), and beam angles uniform in +-half_span_deg. */
int dsaxbf_synth_geometry(int nant, int ncal, int nbeam, uint64_t seed,
                          double length_m, double amp_jitter,
                          double half_span_deg, float *antpos, float *cal,
                          double *theta_rad);


/* ---- batches: many blocks of different shapes in one graph launch ------- */
/* A batch binds n plans (each with weights set; no -s/-q mode) to device
 * buffers: d_in[i] (input_bytes of plan i), d_scale[i] (nbeam floats),
 * d_out[i] (output_bytes), d_beam_power[i] (nbeam floats, or NULL for none;
 * the array itself may be NULL).  Blocks that share precision, tavg and
 * kernel shape run in one launch over all their tiles; the launches (and the
 * per-block beam-power sums) are captured once into a HIP graph, which
 * dsaxbf_batch_launch replays.  The plans must outlive the batch; their
 * variants at creation (or after dsaxbf_batch_autotune) fix the tile shapes.
 * Each block needs a plan of its own (a plan holds its block's beam-power
 * scratch): the same plan twice is DSAXBF_EINVAL.
 * No counterpart in the reference (one fixed-shape block per process). */
typedef struct dsaxbf_batch dsaxbf_batch;
int dsaxbf_batch_create(dsaxbf_batch **batch, dsaxbf_plan *const *plans, int n,
                        const uint8_t *const *d_in, const float *const *d_scale,
                        uint8_t *const *d_out, float *const *d_beam_power);
/* Autotuner for a batch: for each distinct block shape, time every classic
 * and shared variant on that block's buffers, set the fastest on the plans of
 * that shape, then rebuild the graph.  ms, if not NULL, receives n*DSAXBF_NVARIANT
 * times (row i for block i, -1 where not run). */
int dsaxbf_batch_autotune(dsaxbf_batch *batch, int reps, float *ms);
/* Replay the batch graph on stream (NULL = the batch's own stream). */
int dsaxbf_batch_launch(dsaxbf_batch *batch, void *stream);
/* Kernel launches in one replay (beamforming groups + beam-power sums). */
int dsaxbf_batch_kernels(const dsaxbf_batch *batch);
void dsaxbf_batch_destroy(dsaxbf_batch *batch);

/* ---- streaming (replaces the per-int copy/launch loop on 4 streams with
 *      pageable copies, dsaX_beamformer.cu:955-1003) -------------------- */
/* A pipe streams blocks through a plan with two pinned host input slots and
 * two device slots: a block's upload (copy stream) overlaps the previous
 * block's beamforming (plan stream).  Usage per block k:
 *   fill dsaxbf_pipe_input(pp, (k+1)%2); dsaxbf_pipe_upload(pp, (k+1)%2);
 *   dsaxbf_pipe_run(pp, k%2, scale, &out, &bp);   (then AGC -> next scale)
 * with slot 0 filled and uploaded before the first run.  A slot's input may be
 * refilled once dsaxbf_pipe_run on that slot has returned.  run blocks until
 * the block's outputs are in the pipe's pinned buffers, whose pointers it
 * returns (valid until the next run on the same slot). */
typedef struct dsaxbf_pipe dsaxbf_pipe;
int dsaxbf_pipe_create(dsaxbf_pipe **pipe, dsaxbf_plan *plan);
void dsaxbf_pipe_destroy(dsaxbf_pipe *pipe);
uint8_t *dsaxbf_pipe_input(dsaxbf_pipe *pipe, int slot);  /* pinned, input_bytes */
int dsaxbf_pipe_upload(dsaxbf_pipe *pipe, int slot);
int dsaxbf_pipe_run(dsaxbf_pipe *pipe, int slot, const float *h_scale,
                    const uint8_t **h_out, const float **h_beam_power);

/* ---- correlator X-engine on the same 4-bit input (replaces dcorrelator,
 *      dsaX_bfCorr.cu:472-569, with its reorder_input :316-356,
 *      reorder_output :392-457 and corr_output_copy :366-385) ---------- */
/* Visibilities of every antenna pair (i, j <= i), baseline number
 * b = i*(i+1)/2 + j (the reference's h_idxs order, :430-437), per channel
 * and pol (XX and YY, as the reference), summed over the whole block:
 *   V = sum over packets and both times of conj(d_j) * d_i
 * as vis[((b*nchan + c)*2 + pol)*2 + {0 re, 1 im}] float32, exact (integer
 * sums; the reference accumulates in FP16).  The input is the beamformer's
 * block (capture order, or dsaX_split -r order with input_order 1).
 * nant 1..96 (the reference's NANTS), nchan a multiple of 8, npkt a multiple
 * of 32. */
typedef struct dsaxbf_xengine dsaxbf_xengine;
int dsaxbf_xengine_create(dsaxbf_xengine **x, int nant, int nchan, int npkt,
                          int input_order, int device);
void dsaxbf_xengine_destroy(dsaxbf_xengine *x);
/* Bytes of the visibility output: nant*(nant+1)/2 * nchan * 2 pols * 8. */
size_t dsaxbf_xengine_vis_bytes(int nant, int nchan);
void *dsaxbf_xengine_stream(dsaxbf_xengine *x);   /* the plan's hipStream_t */
/* One block already in device memory (stream NULL = the plan's own).
 * Writes every entry of d_vis; no synchronisation.  Three launches: the
 * block turned into (channel, pol) planes of 64-sample units (plan-owned, the
 * input's size), the correlation over time slices, and the sum of the
 * slices' partial visibilities (plan-owned) when there is more than one
 * slice. */
int dsaxbf_xengine_execute(dsaxbf_xengine *x, const uint8_t *d_in, float *d_vis,
                           void *stream);
/* Time slices the correlation runs in (1 .. npkt/32; 0 = the library's
 * choice for the shape, the default).  Tuning only: results are identical
 * (exact integer sums).  Reallocates the partial sums; not while a launch
 * on the plan is in flight. */
int dsaxbf_xengine_set_slices(dsaxbf_xengine *x, int nslice);
int dsaxbf_xengine_get_slices(const dsaxbf_xengine *x);
/* Host block in, host visibilities out (copies, run, wait): the
 * psrdada-facing call, like dsaxbf_process_block. */
int dsaxbf_xengine_process_block(dsaxbf_xengine *x, const uint8_t *h_in,
                                 float *h_vis);

/* ---- device memory helpers (for bindings without a GPU array library) --- */
int dsaxbf_dev_alloc(void **d_ptr, size_t bytes, int device);
int dsaxbf_dev_free(void *d_ptr);
int dsaxbf_copy_h2d(void *d_dst, const void *h_src, size_t bytes, void *stream);
int dsaxbf_copy_d2h(void *h_dst, const void *d_src, size_t bytes, void *stream);
int dsaxbf_sync(void *stream);

/* ---- diagnostics (dsaX_beamformer.cu:217-238,971-972) ------------------ */
/* ANTPOL_RMS of packet pkt, channel group g: rms[pol*64+ant] =
 * sqrt(sum over 8 ch x 2 t of (re^2+im^2) / 32), host block pointer. */
int dsaxbf_antpol_rms(const dsaxbf_config *cfg, const uint8_t *h_block,
                      int pkt, int group, float *rms);

/* Library build identification, e.g. "dsaxbf 1.0 gfx950". */
const char *dsaxbf_version(void);

#ifdef __cplusplus
}
#endif
#endif /* DSAXBF_H */

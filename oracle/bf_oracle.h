/*
 * bf_oracle.h -- CPU restatement of the DSA X-engine beamformer math.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, never the product: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Every function follows the reference file:line it cites.  Arithmetic is
 * FP64 (orc_*); the multithreaded FP32 loop (cpu_*) is the reported CPU
 * baseline and follows the same op order in float.
 *
 * Pins (see tests/): unpack is checked exhaustively against the closed form
 * over all 256 bytes and against utils/packet.out decode statistics; weights
 * against closed forms and invariants (beam at angle 0 equals cal, |W| = |cal|,
 * phase linear in x and f); power against brute force on tiny shapes, an
 * independent numpy restatement, single-antenna and linearity invariants and a
 * steered-tone peak; requantisation against hand-computed edge values; the AGC
 * against hand-computed sequences.
 */
#ifndef BF_ORACLE_H
#define BF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_dims {
  int nant, nchan, npkt, nbeam, tavg;
  int input_order;   /* 0 capture [p][a][c][t][q]; 1 split -r [p][c/8][a][8][t][q] */
} orc_dims;

/* byte -> (re, im), 4-bit two's complement, low nibble real
 * (dsaX_beamformer.cu:190-191). */
void orc_unpack_byte(uint8_t v, int *re, int *im);

/* byte offset of sample (pkt p, ant a, chan c, time t, pol q) in a block
 * (dsaX_beamformer.cu:182-186 for order 0; dsaX_split.c:86-101 for order 1). */
long orc_in_index(const orc_dims *d, int p, int a, int c, int t, int q);

/* Beam weights in FP64 (dsaX_beamformer.cu:420-449):
 *   f     = freqs[8g+4] (wmode 0) or freqs[c] (wmode 1)
 *   ang   = theta_b (amode 0) or sin(theta_b) (amode 1)
 *   phi   = -2*pi*f*ang*x_a/c ;  W = (cos phi + i sin phi) * cal[a][wc][q]
 * W layout: [(wc*2+q)*nbeam + b]*nant + a, complex (re,im) doubles;
 * wc runs over nchan/8 (wmode 0) or nchan (wmode 1).
 * cal layout [nant][ncal][2][2]. */
void orc_weights(const orc_dims *d, int wmode, int amode, const double *antpos,
                 const double *cal, const double *freqs_hz,
                 const double *theta, double *W);

/* Reference fan: theta_b = sep*(nbeam/2-1-b)*pi/10800 (dsaX_beamformer.cu:440). */
void orc_ref_theta(int nbeam, double sep_arcmin, double *theta);

/* Averaged power before scaling, FP64, for output channel groups [g0, g1):
 * power[(n_out*nbeam + b)*ngroup + g] = sum over the 8 channels of group g,
 * both pols and the tavg samples of bin n_out of |sum_a W*D|^2
 * (dsaX_beamformer.cu:341-365 GEMM + detect, :383-395 channel sum,
 * :139-147 pol/time sum).  stuff = 1: beam b carries antenna
 * ((b/16)%4)*16 + b%16, unweighted (:326-337,400-404); stuff = 2: beam 0 is
 * the incoherent sum of every antenna (:307-323,405-414).  Entries of groups
 * outside [g0, g1) are left untouched.  Multithreaded over groups. */
void orc_power(const orc_dims *d, int wmode, const uint8_t *in, const double *W,
               int stuff, int g0, int g1, double *power);

/* (uint8)(rint(x*scale)/2): float product, round half to even, saturating
 * conversion to int, C division, wrap mod 256 (dsaX_beamformer.cu:150-156). */
uint8_t orc_requant(float x, float scale);

/* The corner turn's transpose (dsaX_dbnic.c:111-116): transmit thread
 * `slice` of dsaX_dbnic copies beams slice*nbps .. +nbps-1 of a BF_BLOCK_KEY
 * block out[nout][nbeam][nw] into op[nbps][nout][nw]
 * (op[j*nout*nw + i*nw + k] = out[i*nbeam*nw + slice*nbps*nw + j*nw + k],
 * NSAMPS_PER_TRANSMIT = nout, NBMS = nbeam, NBEAMS_PER_BLOCK = nbps). */
void orc_dbnic_op(const uint8_t *out, int nout, int nbeam, int nw, int nbps, int slice,
                  uint8_t *op);

/* Sum power over the block per beam (badder + calc_allbp,
 * dsaX_beamformer.cu:78-112,482-495). */
void orc_beam_power(const orc_dims *d, const double *power, double *bp);

/* AGC state machine, one block (dsaX_beamformer.cu:507-530,1005-1066).
 * state: started, bpctr, medbp; many_bp [8][nbeam].  bp_raw in, scale out
 * (for the next block).  Returns 1 if this block's output is zeroed. */
typedef struct orc_agc {
  int nbeam, ngroup, nints, enabled, started, bpctr;
  float medbp;
  float many_bp[8 * 1024];
} orc_agc;
void orc_agc_init(orc_agc *s, int nbeam, int ngroup, int nints, int enabled);
int  orc_agc_update(orc_agc *s, const float *bp_raw, float *scale);

/* ---- correlator X-engine ------------------------------------------------ */
/* Baseline number of antenna pair (i, j <= i): b = i*(i+1)/2 + j, the order
 * of the h_idxs loop in reorder_output (dsaX_bfCorr.cu:430-437, "upper
 * triangular order (column major) to match xGPU"). */
long orc_baseline(int i, int j);
/* dcorrelator (dsaX_bfCorr.cu:472-569) with reorder_input (:316-356),
 * reorder_output (:392-457) and corr_output_copy (:366-385): for every
 * channel c and pol q, the visibility of baseline b = (i, j <= i) is
 *   V = sum over the block's packets and both times of conj(d_j) * d_i
 * (the cublas products ac + bd and ad - bc of the column-major Gram matrix),
 * written vis[((b*nchan + c)*2 + q)*2 + {0 re, 1 im}].  Exact (integer sums
 * of 4-bit products) in int64, stored as double.  Multithreaded over
 * channels.  nbeam and tavg of d are unused. */
void orc_xcorr(const orc_dims *d, const uint8_t *in, double *vis);

/* ---- weight encodings of the GPU path (bf_quant.c) -------------------- */
/* Code values and round-to-nearest-even encoders, saturating at the largest
 * finite value (binary16: overflow to infinity): OCP E4M3, OCP MX E2M3
 * (6-bit code), IEEE binary16. */
double   orc_e4m3_decode(uint32_t c);
uint32_t orc_e4m3_encode(double v);
double   orc_e2m3_decode(uint32_t c);
uint32_t orc_e2m3_encode(double v);
double   orc_f16_decode(uint32_t h);
uint32_t orc_f16_encode(double v);
/* Weights (orc_weights layout) through the GPU path's encoding of precision
 * prec, as the values its matrix cores multiply (scales undone):
 *   0 FP16: w*2^e rounded to binary16, the row's largest component in
 *     [2^14, 2^15) (calc_weights' __float2half, dsaX_beamformer.cu:442-446);
 *   1 FP8 split: hi = E4M3(w*2^e), lo = E4M3(w*2^e - hi), largest in [128, 256);
 *   2 FP6 split: x = w*7.5/max; per 16 antennas (32 components) a part is
 *     E2M3 of the remainder boosted by the largest power of two keeping it
 *     within 7.5; hi and lo parts;
 *   3 FP6 three-part split: hi, mid and lo parts that way.
 * Returns 0, or < 0 for an unknown precision or no memory. */
int orc_quant_weights(const orc_dims *d, int wmode, int prec, const double *W, double *Wq);

/* ---- FP32 multithreaded CPU baseline (same op order, float) ---------- */
/* Weights as float complex in the same layout as orc_weights; writes float
 * power and uint8 output; returns wall seconds.  nthreads <= 0: all cores. */
double cpu_beamform(const orc_dims *d, int wmode, const uint8_t *in,
                    const float *W, const float *scale, int g0, int g1,
                    float *power, uint8_t *out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif

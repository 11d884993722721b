/*
 * bf_host.c -- host-side C of libdsaxbf: configuration, errors, the weights
 * file (init_weights), the AGC state machine, requantisation of one value,
 * reference frequencies, ANTPOL_RMS and the seeded synthetic generator.
 * No GPU code here.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "bf_internal.h"
#include "dsaxbf.h"

/* ---- errors ------------------------------------------------------------- */
static __thread char g_errmsg[512];

int dsaxbf_fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_errmsg, sizeof(g_errmsg), fmt, ap);
  va_end(ap);
  return code;
}
void dsaxbf_clear_error(void) { g_errmsg[0] = 0; }
const char *dsaxbf_last_error(void) { return g_errmsg; }
const char *dsaxbf_version(void) { return "dsaxbf 1.0 gfx950"; }

/* ---- configuration ------------------------------------------------------ */
/* dsaX_def.h:72,80 (NPACKETS 2048, NW 48), dsaX_beamformer.cu:799 (384
 * channels), :139-156 (8-sample adder), :420-449 (256 beams, group weights,
 * theta without sin). */
void dsaxbf_config_default(dsaxbf_config *cfg, int nant) {
  memset(cfg, 0, sizeof(*cfg));
  cfg->nant = nant;
  cfg->nchan = 384;
  cfg->npkt = DSAXBF_NPACKETS;
  cfg->nbeam = 256;
  cfg->tavg = 8;
  cfg->precision = DSAXBF_PREC_FP16;
  cfg->weight_mode = DSAXBF_WEIGHTS_GROUP;
  cfg->angle_mode = DSAXBF_ANGLE_THETA;
  cfg->stuffants = DSAXBF_STUFF_NONE;
  cfg->input_order = DSAXBF_ORDER_CAPTURE;
  cfg->device = 0;
  cfg->output_layout = DSAXBF_LAYOUT_TIME_MAJOR;
}

size_t dsaxbf_input_bytes(const dsaxbf_config *cfg) {
  return (size_t)cfg->npkt * cfg->nant * cfg->nchan * 4;
}
size_t dsaxbf_output_bytes(const dsaxbf_config *cfg) {
  if (cfg->tavg <= 0) return 0;
  return (size_t)(2 * cfg->npkt / cfg->tavg) * cfg->nbeam * (cfg->nchan / 8);
}

int dsaxbf_check_config(const dsaxbf_config *c) {
  if (!c) return dsaxbf_fail(DSAXBF_EINVAL, "config is NULL");
  if (c->nant < 1 || c->nant > 256)
    return dsaxbf_fail(DSAXBF_EINVAL, "nant=%d outside 1..256", c->nant);
  if (c->nchan < 8 || c->nchan % 8)
    return dsaxbf_fail(DSAXBF_EINVAL, "nchan=%d must be a positive multiple of 8", c->nchan);
  if (c->npkt < 32 || c->npkt % 32)
    return dsaxbf_fail(DSAXBF_EINVAL, "npkt=%d must be a positive multiple of 32", c->npkt);
  if (c->nbeam < 1 || c->nbeam > 1024)
    return dsaxbf_fail(DSAXBF_EINVAL, "nbeam=%d outside 1..1024", c->nbeam);
  if (c->tavg != 8 && c->tavg != 16 && c->tavg != 32 && c->tavg != 64)
    return dsaxbf_fail(DSAXBF_EINVAL, "tavg=%d must be 8, 16, 32 or 64", c->tavg);
  if (c->precision != DSAXBF_PREC_FP16 && c->precision != DSAXBF_PREC_FP8SPLIT &&
      c->precision != DSAXBF_PREC_FP6SPLIT && c->precision != DSAXBF_PREC_FP6SPLIT3)
    return dsaxbf_fail(DSAXBF_EINVAL, "precision=%d unknown", c->precision);
  if (c->precision == DSAXBF_PREC_FP6SPLIT3 && c->nant > 128)
    return dsaxbf_fail(DSAXBF_EINVAL, "nant=%d: the three-part FP6 split runs up to 128 antennas",
                       c->nant);
  if (c->weight_mode != 0 && c->weight_mode != 1)
    return dsaxbf_fail(DSAXBF_EINVAL, "weight_mode=%d unknown", c->weight_mode);
  if (c->angle_mode != 0 && c->angle_mode != 1)
    return dsaxbf_fail(DSAXBF_EINVAL, "angle_mode=%d unknown", c->angle_mode);
  if (c->stuffants < 0 || c->stuffants > 2)
    return dsaxbf_fail(DSAXBF_EINVAL, "stuffants=%d unknown", c->stuffants);
  if (c->input_order != 0 && c->input_order != 1)
    return dsaxbf_fail(DSAXBF_EINVAL, "input_order=%d unknown", c->input_order);
  if (c->output_layout != DSAXBF_LAYOUT_TIME_MAJOR && c->output_layout != DSAXBF_LAYOUT_BEAM_MAJOR)
    return dsaxbf_fail(DSAXBF_EINVAL, "output_layout=%d unknown", c->output_layout);
  return DSAXBF_OK;
}

/* dsaX_beamformer.cu:799-800 */
void dsaxbf_reference_freqs(double fch1_mhz, int nchan, double *freqs_hz) {
  for (int i = 0; i < nchan; i++) freqs_hz[i] = (fch1_mhz - i * 250.0 / 8192.0) * 1e6;
}

/* ---- weights file (init_weights, dsaX_beamformer.cu:559-602) ------------- */
int dsaxbf_read_weights_file(const char *fnam, const char *flagfile, float *antpos,
                             float *cal) {
  dsaxbf_clear_error();
  if (!fnam || !antpos || !cal) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  FILE *fin = fopen(fnam, "rb");
  if (!fin) return dsaxbf_fail(DSAXBF_EIO, "cannot open weights file %s: %s", fnam, strerror(errno));
  const size_t ncal = (size_t)DSAXBF_NANT_FILE * DSAXBF_NW * 2 * 2;
  size_t r1 = fread(antpos, sizeof(float), DSAXBF_NANT_FILE, fin);
  size_t r2 = fread(cal, sizeof(float), ncal, fin);
  fclose(fin);
  if (r1 != DSAXBF_NANT_FILE || r2 != ncal)
    return dsaxbf_fail(DSAXBF_EIO, "weights file %s is short (%zu of %zu floats)", fnam, r1 + r2,
                       (size_t)DSAXBF_NANT_FILE + ncal);
  /* w -> w/|w|^2, zero stays zero (:579-586) */
  for (size_t i = 0; i < ncal / 2; i++) {
    float wr = cal[2 * i], wi = cal[2 * i + 1];
    float wnorm = sqrtf(wr * wr + wi * wi);
    if (wnorm != 0.0f) {
      cal[2 * i] = wr / (wnorm * wnorm);
      cal[2 * i + 1] = wi / (wnorm * wnorm);
    }
  }
  /* flagged antennas -> 0 (:589-595) */
  if (flagfile && strcmp(flagfile, "nofile") != 0) {
    FILE *fa = fopen(flagfile, "r");
    if (!fa) return dsaxbf_fail(DSAXBF_EIO, "cannot open flag file %s: %s", flagfile, strerror(errno));
    char line[256];
    int lineno = 0;
    while (fgets(line, sizeof(line), fa)) {
      lineno++;
      char *p = line;
      while (*p == ' ' || *p == '\t') p++;
      if (*p == '\n' || *p == '\r' || *p == 0 || *p == '#') continue;
      char *end;
      long ant = strtol(p, &end, 10);
      if (end == p) {
        fclose(fa);
        return dsaxbf_fail(DSAXBF_EIO, "flag file %s line %d: not an integer", flagfile, lineno);
      }
      if (ant < 0 || ant >= DSAXBF_NANT_FILE) {
        fclose(fa);
        return dsaxbf_fail(DSAXBF_EINVAL, "flag file %s line %d: antenna %ld outside 0..63",
                           flagfile, lineno, ant);
      }
      memset(cal + (size_t)ant * DSAXBF_NW * 4, 0, sizeof(float) * DSAXBF_NW * 4);
    }
    fclose(fa);
  }
  return DSAXBF_OK;
}

/* ---- requantisation (dsaX_beamformer.cu:150-156) ------------------------- */
uint8_t dsaxbf_requant(float x, float scale) {
  float v = x * scale;
  float r = rintf(v);
  int iv;
  if (r != r) iv = 0;
  else if (r >= 2147483648.0f) iv = 2147483647;
  else if (r <= -2147483648.0f) iv = (-2147483647 - 1);
  else iv = (int)r;
  return (uint8_t)(unsigned)(iv / 2);
}

/* ---- AGC (dsaX_beamformer.cu:482-530,1005-1066) -------------------------- */
int dsaxbf_agc_init(dsaxbf_agc *agc, int nbeam, int ngroup, int nints, int enabled) {
  if (!agc || nbeam < 2 || ngroup < 1 || nints < 1)
    return dsaxbf_fail(DSAXBF_EINVAL, "bad AGC arguments");
  memset(agc, 0, sizeof(*agc));
  agc->nbeam = nbeam;
  agc->ngroup = ngroup;
  agc->nints = nints;
  agc->enabled = enabled;
  agc->many_bp = calloc((size_t)DSAXBF_NBP * nbeam, sizeof(float));
  agc->bp = calloc((size_t)nbeam, sizeof(float));
  if (!agc->many_bp || !agc->bp) {
    dsaxbf_agc_free(agc);
    return dsaxbf_fail(DSAXBF_ENOMEM, "AGC allocation failed");
  }
  return DSAXBF_OK;
}

void dsaxbf_agc_free(dsaxbf_agc *agc) {
  if (!agc) return;
  free(agc->many_bp);
  free(agc->bp);
  agc->many_bp = agc->bp = NULL;
}

static int cmp_float(const void *x, const void *y) {
  float a = *(const float *)x, b = *(const float *)y;
  return (a > b) - (a < b);
}

/* ret_med_bp (:507-514): every beam gets 0.5*(bp[n/2-1] + bp[n/2]) of the
 * sorted powers. */
static int median_fill(float *bp, int n) {
  float *tmp = malloc(sizeof(float) * n);
  if (!tmp) return dsaxbf_fail(DSAXBF_ENOMEM, "AGC median: allocation failed");
  memcpy(tmp, bp, sizeof(float) * n);
  qsort(tmp, n, sizeof(float), cmp_float);
  float med = 0.5f * (tmp[n / 2 - 1] + tmp[n / 2]);
  for (int i = 0; i < n; i++) bp[i] = med;
  free(tmp);
  return DSAXBF_OK;
}

int dsaxbf_agc_update(dsaxbf_agc *agc, const float *bp_raw, float *scale_out, int *zero_output) {
  if (!agc || !agc->bp || !bp_raw || !scale_out) return dsaxbf_fail(DSAXBF_EINVAL, "bad AGC arguments");
  const int nb = agc->nbeam;
  float *bp = agc->bp;
  int zero = 0;
  memcpy(bp, bp_raw, sizeof(float) * nb);
  memcpy(agc->many_bp + (size_t)(agc->bpctr % DSAXBF_NBP) * nb, bp, sizeof(float) * nb);
  if (agc->started == 0) {           /* :1021-1033 first block: median, zeros out */
    if (median_fill(bp, nb)) return DSAXBF_ENOMEM;
    agc->started = 1;
    agc->medbp = bp[0];
    zero = 1;
  }
  if (agc->started > 0 && agc->bpctr < DSAXBF_NBP && median_fill(bp, nb))   /* :1035-1036 */
    return DSAXBF_ENOMEM;
  if (agc->started > 0 && agc->bpctr >= DSAXBF_NBP) {                      /* :1039-1048 */
    for (int i = 0; i < nb; i++) {   /* ret_many_bp :516-530 */
      float s = 0.f;
      for (int j = 0; j < DSAXBF_NBP; j++) s += agc->many_bp[(size_t)j * nb + i];
      bp[i] = (float)((double)s / (1.0 * DSAXBF_NBP));
    }
    for (int i = 0; i < nb; i++)
      if (fabs(bp[i] - agc->medbp) / agc->medbp > 0.1) bp[i] = agc->medbp;
    agc->started = 2;
  }
  /* :1053-1063.  With AGC off every beam holds the first block's median (the
   * reference resets bp inside its beam loop, so only its last beam is
   * scaled; this is the intended behaviour). */
  for (int i = 0; i < nb; i++) {
    if (!agc->enabled) bp[i] = agc->medbp;
    if (bp[i] != 0.f) {
      bp[i] = (float)((double)bp[i] / (agc->ngroup * 1.0 * agc->nints));
      bp[i] = (float)(2.5 * 128.0 / (double)bp[i]);
    }
    scale_out[i] = bp[i];
  }
  agc->bpctr++;
  if (zero_output) *zero_output = zero;
  return DSAXBF_OK;
}

/* ---- ANTPOL_RMS (rms_printer, dsaX_beamformer.cu:217-238) ---------------- */
int dsaxbf_antpol_rms(const dsaxbf_config *cfg, const uint8_t *h_block, int pkt, int group,
                      float *rms) {
  if (!cfg || !h_block || !rms || pkt < 0 || pkt >= cfg->npkt || group < 0 ||
      group >= cfg->nchan / 8)
    return dsaxbf_fail(DSAXBF_EINVAL, "bad ANTPOL_RMS arguments");
  for (int q = 0; q < 2; q++)
    for (int a = 0; a < 64; a++) {
      float s = 0.f;
      if (a < cfg->nant)
        for (int ci = 0; ci < 8; ci++)
          for (int t = 0; t < 2; t++) {
            int c = group * 8 + ci;
            size_t off = cfg->input_order == 0
                             ? ((((size_t)pkt * cfg->nant + a) * cfg->nchan + c) * 2 + t) * 2 + q
                             : ((((size_t)pkt * (cfg->nchan / 8) + group) * cfg->nant + a) * 8 + ci) * 4 + t * 2 + q;
            uint8_t v = h_block[off];
            int lo = v & 15, hi = v >> 4;
            float re = (float)(lo >= 8 ? lo - 16 : lo), im = (float)(hi >= 8 ? hi - 16 : hi);
            s += re * re + im * im;
          }
      rms[q * 64 + a] = sqrtf(s / 32.f);
    }
  return DSAXBF_OK;
}

/* ---- seeded generator ----------------------------------------------------- */
static inline uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
typedef struct { uint64_t s[4]; int have; double spare; } rng_t;
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint64_t xoshiro(rng_t *r) {   /* xoshiro256** */
  uint64_t res = rotl(r->s[1] * 5, 7) * 9, t = r->s[1] << 17;
  r->s[2] ^= r->s[0]; r->s[3] ^= r->s[1]; r->s[1] ^= r->s[2]; r->s[0] ^= r->s[3];
  r->s[2] ^= t; r->s[3] = rotl(r->s[3], 45);
  return res;
}
static void rng_seed(rng_t *r, uint64_t seed) {
  uint64_t sm = seed;
  for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&sm);
  r->have = 0;
}
static inline double rng_uniform(rng_t *r) {  /* (0, 1) */
  return ((xoshiro(r) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
static inline double rng_normal(rng_t *r) {   /* Box-Muller, FP64 */
  if (r->have) { r->have = 0; return r->spare; }
  double u1 = rng_uniform(r), u2 = rng_uniform(r);
  double rad = sqrt(-2.0 * log(u1)), ang = 2.0 * M_PI * u2;
  r->spare = rad * sin(ang);
  r->have = 1;
  return rad * cos(ang);
}

typedef struct {
  const dsaxbf_config *cfg;
  uint64_t seed;
  double sigma, tone_amp, tone_theta, tone_w;
  int clip, ntone;
  const int *tone_chan;
  const float *antpos, *gain;
  const double *freqs;
  uint8_t *blk;
  int p0, p1;
} synth_job;

static inline int clip_round(double x, int clip) {
  long v = lround(x);
  if (v > clip) v = clip;
  if (v < -clip) v = -clip;
  return (int)v;
}

static void *synth_worker(void *arg) {
  synth_job *j = arg;
  const dsaxbf_config *c = j->cfg;
  rng_t r;
  for (int p = j->p0; p < j->p1; p++) {
    rng_seed(&r, j->seed * 0x9E3779B97F4A7C15ull + (uint64_t)p * 0xD1B54A32D192ED03ull + 1);
    for (int a = 0; a < c->nant; a++)
      for (int ch = 0; ch < c->nchan; ch++) {
        int tone = 0;
        for (int k = 0; k < j->ntone; k++) tone |= (j->tone_chan[k] == ch);
        for (int t = 0; t < 2; t++)
          for (int q = 0; q < 2; q++) {
            double re = j->sigma * rng_normal(&r), im = j->sigma * rng_normal(&r);
            if (tone) {
              /* geometric phase +2 pi f x sin(theta0)/c, slow drift in time */
              double ph = 2.0 * M_PI * j->freqs[ch] * (double)j->antpos[a] * sin(j->tone_theta) /
                              DSAXBF_CVAC + j->tone_w * (2 * p + t);
              double tr = j->tone_amp * cos(ph), ti = j->tone_amp * sin(ph);
              if (j->gain) { /* divide by the antenna gain so cal = gain re-phases it */
                double gr = j->gain[2 * a], gi = j->gain[2 * a + 1], g2 = gr * gr + gi * gi;
                if (g2 > 0) {
                  double xr = (tr * gr + ti * gi) / g2, xi = (ti * gr - tr * gi) / g2;
                  tr = xr;
                  ti = xi;
                }
              }
              re += tr;
              im += ti;
            }
            int ir = clip_round(re, j->clip), ii = clip_round(im, j->clip);
            size_t off = c->input_order == 0
                             ? ((((size_t)p * c->nant + a) * c->nchan + ch) * 2 + t) * 2 + q
                             : ((((size_t)p * (c->nchan / 8) + ch / 8) * c->nant + a) * 8 + ch % 8) * 4 + t * 2 + q;
            j->blk[off] = (uint8_t)(((unsigned)ii & 15u) << 4 | ((unsigned)ir & 15u));
          }
      }
  }
  return NULL;
}

int dsaxbf_synth_block(const dsaxbf_config *cfg, uint64_t seed, double sigma, int clip,
                       double tone_snr, double tone_theta, const int *tone_chan, int ntone,
                       const float *antpos, const float *gain, const double *freqs_hz,
                       uint8_t *h_block) {
  if (!cfg || !h_block || clip < 0 || clip > 8) return dsaxbf_fail(DSAXBF_EINVAL, "bad synth arguments");
  if (tone_snr > 0 && (!tone_chan || ntone <= 0 || !antpos || !freqs_hz))
    return dsaxbf_fail(DSAXBF_EINVAL, "tone needs channels, antpos and freqs");
  if (clip == 8) clip = 7; /* symmetric clip keeps -8 out unless asked by data */
  int nthr = (int)sysconf(_SC_NPROCESSORS_ONLN);
  if (nthr > 16) nthr = 16;
  if (nthr < 1) nthr = 1;
  if (nthr > cfg->npkt) nthr = cfg->npkt;
  pthread_t tid[16];
  synth_job job[16];
  for (int i = 0; i < nthr; i++) {
    job[i].cfg = cfg;
    job[i].seed = seed;
    job[i].sigma = sigma;
    job[i].tone_amp = tone_snr > 0 ? sigma * sqrt(2.0 * tone_snr) : 0.0;
    job[i].tone_theta = tone_theta;
    job[i].tone_w = 0.05;
    job[i].clip = clip;
    job[i].ntone = tone_snr > 0 ? ntone : 0;
    job[i].tone_chan = tone_chan;
    job[i].antpos = antpos;
    job[i].gain = gain;
    job[i].freqs = freqs_hz;
    job[i].blk = h_block;
    job[i].p0 = (int)((long)cfg->npkt * i / nthr);
    job[i].p1 = (int)((long)cfg->npkt * (i + 1) / nthr);
  }
  int started = 0;   /* workers that run in threads; the rest run here */
  for (; started < nthr; started++)
    if (pthread_create(&tid[started], NULL, synth_worker, &job[started])) break;
  for (int i = started; i < nthr; i++) synth_worker(&job[i]);
  for (int i = 0; i < started; i++) pthread_join(tid[i], NULL);
  return DSAXBF_OK;
}

int dsaxbf_synth_geometry(int nant, int ncal, int nbeam, uint64_t seed, double length_m,
                          double amp_jitter, double half_span_deg, float *antpos, float *cal,
                          double *theta_rad) {
  if (nant < 1 || ncal < 1 || nbeam < 1) return dsaxbf_fail(DSAXBF_EINVAL, "bad geometry arguments");
  rng_t r;
  rng_seed(&r, seed ^ 0xA5A5A5A55A5A5A5Aull);
  for (int a = 0; a < nant; a++) {
    double x = length_m * rng_uniform(&r);
    double amp = 1.0 + amp_jitter * (2.0 * rng_uniform(&r) - 1.0);
    double ph = 2.0 * M_PI * rng_uniform(&r);
    if (antpos) antpos[a] = (float)x;
    if (cal)
      for (int w = 0; w < ncal; w++)
        for (int q = 0; q < 2; q++) {
          float *cw = cal + (((size_t)a * ncal + w) * 2 + q) * 2;
          cw[0] = (float)(amp * cos(ph));
          cw[1] = (float)(amp * sin(ph));
        }
  }
  if (theta_rad) {
    double span = half_span_deg * M_PI / 180.0;
    for (int b = 0; b < nbeam; b++)
      theta_rad[b] = nbeam == 1 ? 0.0 : -span + 2.0 * span * b / (nbeam - 1);
  }
  return DSAXBF_OK;
}

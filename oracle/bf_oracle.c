/*
 * bf_oracle.c -- FP64 CPU restatement of the reference beamformer math.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Loaded only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Reference: src/dsaX_beamformer.cu (cited per function), src/dsaX_def.h.
 * Math is restated in FP64 in the reference's order: unpack -> complex weight
 * x voltage -> |B|^2 -> sum over 8 channels x 2 pols x tavg samples ->
 * scale -> rint -> /2 -> uint8 (wrap).  The reference accumulates the GEMM in
 * FP16 (dsaX_beamformer.cu:291-294); the oracle is exact FP64 by design.
 */
#include "bf_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define ORC_CVAC 299792458.0

/* dsaX_beamformer.cu:190-191: re = (char)((v & 15) << 4) >> 4,
 * im = (char)(v & 240) >> 4 (arithmetic shifts of a signed char). */
void orc_unpack_byte(uint8_t v, int *re, int *im) {
  int lo = v & 15, hi = (v >> 4) & 15;
  *re = lo >= 8 ? lo - 16 : lo;
  *im = hi >= 8 ? hi - 16 : hi;
}

/* dsaX_beamformer.cu:182-186 (promoter reads [time][ant][chan][t][pol]) and
 * dsaX_split.c:94-95 (-r writes [time][chan/8][ant][32 B]). */
long orc_in_index(const orc_dims *d, int p, int a, int c, int t, int q) {
  if (d->input_order == 1) {
    long g = c / 8, ci = c % 8;
    return ((((long)p * (d->nchan / 8) + g) * d->nant + a) * 8 + ci) * 4 + t * 2 + q;
  }
  return ((((long)p * d->nant + a) * d->nchan + c) * 2 + t) * 2 + q;
}

/* dsaX_beamformer.cu:440: theta = sep*(127-beam)*PI/10800 for 256 beams. */
void orc_ref_theta(int nbeam, double sep_arcmin, double *theta) {
  for (int b = 0; b < nbeam; b++)
    theta[b] = sep_arcmin * (double)(nbeam / 2 - 1 - b) * M_PI / 10800.0;
}

/* dsaX_beamformer.cu:420-449 (calc_weights) in FP64.
 * afac = -2*PI*freqs[f*8+4]*theta/CVAC; tw = exp(i*afac*antpos[ant]);
 * wr = twr*cr - twi*ci; wi = twi*cr + twr*ci. */
void orc_weights(const orc_dims *d, int wmode, int amode, const double *antpos,
                 const double *cal, const double *freqs_hz,
                 const double *theta, double *W) {
  int ncal = wmode == 0 ? d->nchan / 8 : d->nchan;
  for (int wc = 0; wc < ncal; wc++) {
    double f = wmode == 0 ? freqs_hz[wc * 8 + 4] : freqs_hz[wc];
    for (int q = 0; q < 2; q++)
      for (int b = 0; b < d->nbeam; b++) {
        double ang = amode == 0 ? theta[b] : sin(theta[b]);
        double afac = -2.0 * M_PI * f * ang / ORC_CVAC;
        for (int a = 0; a < d->nant; a++) {
          double ph = afac * antpos[a];
          double twr = cos(ph), twi = sin(ph);
          const double *cw = cal + (((long)a * ncal + wc) * 2 + q) * 2;
          double *w = W + ((((long)wc * 2 + q) * d->nbeam + b) * d->nant + a) * 2;
          w[0] = twr * cw[0] - twi * cw[1];
          w[1] = twi * cw[0] + twr * cw[1];
        }
      }
  }
}

typedef struct {
  const orc_dims *d;
  int wmode, stuff, g;
  const uint8_t *in;
  const double *W;
  double *power;
} orc_job;

/* One output channel group: dsaX_beamformer.cu:267-416 per (time, group, pol)
 * then :120-158 summation.  Work arrays are per call. */
static void orc_group(const orc_job *j) {
  const orc_dims *d = j->d;
  int nant = d->nant, nbeam = d->nbeam, ng = d->nchan / 8;
  int nsamp = 2 * d->npkt, nout = nsamp / d->tavg;
  double *dr = malloc(sizeof(double) * nant), *di = malloc(sizeof(double) * nant);
  double *acc = calloc((size_t)nout * nbeam, sizeof(double));
  for (int ci = 0; ci < 8; ci++) {
    int c = j->g * 8 + ci;
    int wc = j->wmode == 0 ? j->g : c;
    for (int q = 0; q < 2; q++) {
      const double *Wcq = j->W + ((long)wc * 2 + q) * nbeam * nant * 2;
      for (int p = 0; p < d->npkt; p++)
        for (int t = 0; t < 2; t++) {
          int n = 2 * p + t, o = n / d->tavg;
          double ib = 0.0;
          for (int a = 0; a < nant; a++) {
            int re, im;
            orc_unpack_byte(j->in[orc_in_index(d, p, a, c, t, q)], &re, &im);
            dr[a] = re;
            di[a] = im;
            ib += (double)re * re + (double)im * im;
          }
          for (int b = 0; b < nbeam; b++) {
            double P;
            if (j->stuff == 1) {
              int ant = ((b / 16) % 4) * 16 + b % 16;
              P = ant < nant ? dr[ant] * dr[ant] + di[ant] * di[ant] : 0.0;
            } else if (j->stuff == 2 && b == 0) {
              P = ib;
            } else {
              const double *w = Wcq + (long)b * nant * 2;
              double br = 0.0, bi = 0.0;
              for (int a = 0; a < nant; a++) {
                br += w[2 * a] * dr[a] - w[2 * a + 1] * di[a];
                bi += w[2 * a + 1] * dr[a] + w[2 * a] * di[a];
              }
              P = br * br + bi * bi;
            }
            acc[(long)o * nbeam + b] += P;
          }
        }
    }
  }
  for (int o = 0; o < nout; o++)
    for (int b = 0; b < nbeam; b++)
      j->power[((long)o * nbeam + b) * ng + j->g] = acc[(long)o * nbeam + b];
  free(dr);
  free(di);
  free(acc);
}

typedef struct {
  orc_job base;
  int g0, g1, stride, first;
} orc_thr;

static void *orc_worker(void *arg) {
  orc_thr *t = arg;
  for (int g = t->g0 + t->first; g < t->g1; g += t->stride) {
    orc_job j = t->base;
    j.g = g;
    orc_group(&j);
  }
  return NULL;
}

void orc_power(const orc_dims *d, int wmode, const uint8_t *in, const double *W,
               int stuff, int g0, int g1, double *power) {
  /* groups outside the block are not computed (nothing is read or written for them) */
  if (g0 < 0) g0 = 0;
  if (g1 > d->nchan / 8) g1 = d->nchan / 8;
  if (g1 <= g0) return;
  int nthr = (int)sysconf(_SC_NPROCESSORS_ONLN);
  if (nthr < 1) nthr = 1;
  if (nthr > g1 - g0) nthr = g1 - g0;
  if (nthr > 64) nthr = 64;
  pthread_t tid[64];
  orc_thr arg[64];
  for (int i = 0; i < nthr; i++) {
    arg[i].base.d = d;
    arg[i].base.wmode = wmode;
    arg[i].base.stuff = stuff;
    arg[i].base.in = in;
    arg[i].base.W = W;
    arg[i].base.power = power;
    arg[i].g0 = g0;
    arg[i].g1 = g1;
    arg[i].stride = nthr;
    arg[i].first = i;
    pthread_create(&tid[i], NULL, orc_worker, &arg[i]);
  }
  for (int i = 0; i < nthr; i++) pthread_join(tid[i], NULL);
}

/* dsaX_beamformer.cu:150-156: (unsigned char)(__float2int_rn(data*bp)/2).
 * __float2int_rn rounds half to even and saturates to the int range. */
uint8_t orc_requant(float x, float scale) {
  float v = x * scale;
  double r = rint((double)v);
  long iv;
  if (r != r) iv = 0;
  else if (r >= 2147483647.0) iv = 2147483647L;
  else if (r <= -2147483648.0) iv = -2147483648L;
  else iv = (long)r;
  int h = (int)iv / 2;
  return (uint8_t)(unsigned)h;
}

/* dsaX_dbnic.c:111-116, the "yes transpose" loop of transmit(). */
void orc_dbnic_op(const uint8_t *out, int nout, int nbeam, int nw, int nbps, int slice,
                  uint8_t *op) {
  for (int i = 0; i < nout; i++)
    for (int j = 0; j < nbps; j++)
      for (int k = 0; k < nw; k++)
        op[(long)j * nout * nw + (long)i * nw + k] =
            out[(long)i * nbeam * nw + (long)slice * nbps * nw + (long)j * nw + k];
}

/* badder (dsaX_beamformer.cu:78-112) accumulates every |B|^2 of the block per
 * (beam, group); calc_allbp (:482-495) sums the groups. */
void orc_beam_power(const orc_dims *d, const double *power, double *bp) {
  int ng = d->nchan / 8, nout = 2 * d->npkt / d->tavg;
  for (int b = 0; b < d->nbeam; b++) bp[b] = 0.0;
  for (int o = 0; o < nout; o++)
    for (int b = 0; b < d->nbeam; b++)
      for (int g = 0; g < ng; g++) bp[b] += power[((long)o * d->nbeam + b) * ng + g];
}

static int orc_cmpf(const void *x, const void *y) {
  float a = *(const float *)x, b = *(const float *)y;
  return (a > b) - (a < b);
}

void orc_agc_init(orc_agc *s, int nbeam, int ngroup, int nints, int enabled) {
  memset(s, 0, sizeof(*s));
  s->nbeam = nbeam;
  s->ngroup = ngroup;
  s->nints = nints;
  s->enabled = enabled;
}

/* dsaX_beamformer.cu:1005-1066 with ret_med_bp :507-514 and ret_many_bp
 * :516-530.  The -g path is restated as intended (hold the first block's
 * median for every beam); the reference resets bp inside its beam loop so
 * only the last beam ends up scaled. */
int orc_agc_update(orc_agc *s, const float *bp_raw, float *scale) {
  int nb = s->nbeam, zero = 0;
  float bp[1024], tmp[1024];
  for (int i = 0; i < nb; i++) bp[i] = bp_raw[i];
  for (int i = 0; i < nb; i++) s->many_bp[(s->bpctr % 8) * nb + i] = bp[i];
  if (s->started == 0) {
    s->started = 1;
    memcpy(tmp, bp, sizeof(float) * nb);
    qsort(tmp, nb, sizeof(float), orc_cmpf);
    float med = 0.5f * (tmp[nb / 2 - 1] + tmp[nb / 2]);
    for (int i = 0; i < nb; i++) bp[i] = med;
    s->medbp = med;
    zero = 1;
  }
  if (s->started > 0 && s->bpctr < 8) {
    memcpy(tmp, bp, sizeof(float) * nb);
    qsort(tmp, nb, sizeof(float), orc_cmpf);
    float med = 0.5f * (tmp[nb / 2 - 1] + tmp[nb / 2]);
    for (int i = 0; i < nb; i++) bp[i] = med;
  }
  if (s->started > 0 && s->bpctr >= 8) {
    for (int i = 0; i < nb; i++) {
      bp[i] = 0.f;
      for (int j = 0; j < 8; j++) bp[i] += s->many_bp[j * nb + i];
      bp[i] = (float)((double)bp[i] / (1.0 * 8));
    }
    for (int i = 0; i < nb; i++)
      if (fabs(bp[i] - s->medbp) / s->medbp > 0.1) bp[i] = s->medbp;
    s->started = 2;
  }
  for (int i = 0; i < nb; i++) {
    if (!s->enabled) bp[i] = s->medbp;
    if (bp[i] != 0.f) {
      bp[i] = (float)((double)bp[i] / (s->ngroup * 1.0 * s->nints));
      bp[i] = (float)(2.5 * 128.0 / (double)bp[i]);
    }
    scale[i] = bp[i];
  }
  s->bpctr++;
  return zero;
}

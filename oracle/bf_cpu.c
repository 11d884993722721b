/*
 * bf_cpu.c -- multithreaded FP32 CPU restatement of the beamformer, timed as
 * the reported CPU baseline (kind "port": the reference has no CPU
 * beamformer of its own).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by bench.py's cpu_baseline leg and tests.
 *
 * Same op order as the reference (dsaX_beamformer.cu): unpack (:190-191) ->
 * complex weight x voltage summed over antennas (:341-358) -> |B|^2
 * (:361-365) -> sum over 8 channels (:383-395), 2 pols and tavg samples
 * (:139-147) -> scale, rint, /2, uint8 (:150-156).  pthreads over (output
 * channel group, time chunk) tasks; the antenna loop is outermost per sample
 * and the beam loop innermost, so the complex multiply-accumulate vectorises
 * across beams without reassociating any sum.
 */
#include "bf_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

typedef struct {
  const orc_dims *d;
  int wmode;
  const uint8_t *in;
  const float *W, *scale;
  float *power;
  uint8_t *out;
  int g0, ntask, tchunk;
  int next;                 /* shared task counter */
  pthread_mutex_t *mu;
} cpu_ctx;

static int take_task(cpu_ctx *c) {
  pthread_mutex_lock(c->mu);
  int t = c->next < c->ntask ? c->next++ : -1;
  pthread_mutex_unlock(c->mu);
  return t;
}

static void *cpu_worker(void *arg) {
  cpu_ctx *cx = arg;
  const orc_dims *d = cx->d;
  const int nant = d->nant, nbeam = d->nbeam, ng = d->nchan / 8;
  const int nsamp = 2 * d->npkt, nchunk = nsamp / cx->tchunk;
  float *wr = malloc(sizeof(float) * (size_t)nbeam * nant);   /* [a][b] */
  float *wi = malloc(sizeof(float) * (size_t)nbeam * nant);
  float *br = malloc(sizeof(float) * nbeam), *bi = malloc(sizeof(float) * nbeam);
  float *acc = malloc(sizeof(float) * (size_t)(cx->tchunk / d->tavg) * nbeam);
  float *xr = malloc(sizeof(float) * nant), *xi = malloc(sizeof(float) * nant);
  for (int task; (task = take_task(cx)) >= 0;) {
    const int g = cx->g0 + task / nchunk, n0 = (task % nchunk) * cx->tchunk;
    const int nbin = cx->tchunk / d->tavg;
    memset(acc, 0, sizeof(float) * (size_t)nbin * nbeam);
    for (int ci = 0; ci < 8; ci++) {
      const int c = g * 8 + ci, wc = cx->wmode == 0 ? g : c;
      for (int q = 0; q < 2; q++) {
        const float *Wcq = cx->W + ((long)wc * 2 + q) * nbeam * nant * 2;
        for (int b = 0; b < nbeam; b++)
          for (int a = 0; a < nant; a++) {
            wr[(long)a * nbeam + b] = Wcq[((long)b * nant + a) * 2];
            wi[(long)a * nbeam + b] = Wcq[((long)b * nant + a) * 2 + 1];
          }
        for (int n = n0; n < n0 + cx->tchunk; n++) {
          const int p = n / 2, t = n % 2;
          for (int a = 0; a < nant; a++) {
            const uint8_t v = cx->in[orc_in_index(d, p, a, c, t, q)];
            const int lo = v & 15, hi = v >> 4;
            xr[a] = (float)(lo >= 8 ? lo - 16 : lo);
            xi[a] = (float)(hi >= 8 ? hi - 16 : hi);
          }
          for (int b = 0; b < nbeam; b++) br[b] = bi[b] = 0.f;
          for (int a = 0; a < nant; a++) {
            const float dr = xr[a], di = xi[a];
            const float *vr = wr + (long)a * nbeam, *vi = wi + (long)a * nbeam;
            for (int b = 0; b < nbeam; b++) {
              br[b] += vr[b] * dr - vi[b] * di;
              bi[b] += vi[b] * dr + vr[b] * di;
            }
          }
          float *arow = acc + (long)((n - n0) / d->tavg) * nbeam;
          for (int b = 0; b < nbeam; b++) arow[b] += br[b] * br[b] + bi[b] * bi[b];
        }
      }
    }
    for (int o = 0; o < nbin; o++)
      for (int b = 0; b < nbeam; b++) {
        const float x = acc[(long)o * nbeam + b];
        const long idx = ((long)(n0 / d->tavg + o) * nbeam + b) * ng + g;
        if (cx->power) cx->power[idx] = x;
        if (cx->out) cx->out[idx] = orc_requant(x, cx->scale[b]);
      }
  }
  free(wr); free(wi); free(br); free(bi); free(acc); free(xr); free(xi);
  return NULL;
}

double cpu_beamform(const orc_dims *d, int wmode, const uint8_t *in,
                    const float *W, const float *scale, int g0, int g1,
                    float *power, uint8_t *out, int nthreads) {
  int nthr = nthreads > 0 ? nthreads : (int)sysconf(_SC_NPROCESSORS_ONLN);
  if (nthr < 1) nthr = 1;
  if (nthr > 256) nthr = 256;
  const int nsamp = 2 * d->npkt;
  int tchunk = nsamp;   /* split time so every thread has work */
  while (tchunk > d->tavg && tchunk % 2 == 0 && (long)(g1 - g0) * (nsamp / tchunk) < 4L * nthr &&
         (tchunk / 2) % d->tavg == 0)
    tchunk /= 2;
  pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
  cpu_ctx cx = {d, wmode, in, W, scale, power, out, g0, (g1 - g0) * (nsamp / tchunk), tchunk, 0, &mu};
  pthread_t *tid = malloc(sizeof(pthread_t) * nthr);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int i = 0; i < nthr; i++) pthread_create(&tid[i], NULL, cpu_worker, &cx);
  for (int i = 0; i < nthr; i++) pthread_join(tid[i], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  free(tid);
  return (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}

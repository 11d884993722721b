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
 * (:139-147) -> scale, rint, /2, uint8 (:150-156).  pthreads over output
 * channel groups; the antenna sum is the vectorised inner loop.
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
  int g0, g1, stride, first;
} cpu_thr;

static void *cpu_worker(void *arg) {
  cpu_thr *t = arg;
  const orc_dims *d = t->d;
  int nant = d->nant, nbeam = d->nbeam, ng = d->nchan / 8;
  int nsamp = 2 * d->npkt, nout = nsamp / d->tavg;
  float *dr = malloc(sizeof(float) * (size_t)nsamp * nant);
  float *di = malloc(sizeof(float) * (size_t)nsamp * nant);
  float *wr = malloc(sizeof(float) * (size_t)nbeam * nant);
  float *wi = malloc(sizeof(float) * (size_t)nbeam * nant);
  float *acc = malloc(sizeof(float) * (size_t)nout * nbeam);
  for (int g = t->g0 + t->first; g < t->g1; g += t->stride) {
    memset(acc, 0, sizeof(float) * (size_t)nout * nbeam);
    for (int ci = 0; ci < 8; ci++) {
      int c = g * 8 + ci, wc = t->wmode == 0 ? g : c;
      for (int q = 0; q < 2; q++) {
        const float *Wcq = t->W + ((long)wc * 2 + q) * nbeam * nant * 2;
        for (int b = 0; b < nbeam; b++)
          for (int a = 0; a < nant; a++) {
            wr[(long)b * nant + a] = Wcq[((long)b * nant + a) * 2];
            wi[(long)b * nant + a] = Wcq[((long)b * nant + a) * 2 + 1];
          }
        for (int p = 0; p < d->npkt; p++)
          for (int tt = 0; tt < 2; tt++)
            for (int a = 0; a < nant; a++) {
              uint8_t v = t->in[orc_in_index(d, p, a, c, tt, q)];
              int lo = v & 15, hi = v >> 4;
              dr[(long)(2 * p + tt) * nant + a] = (float)(lo >= 8 ? lo - 16 : lo);
              di[(long)(2 * p + tt) * nant + a] = (float)(hi >= 8 ? hi - 16 : hi);
            }
        for (int n = 0; n < nsamp; n++) {
          const float *xr = dr + (long)n * nant, *xi = di + (long)n * nant;
          float *arow = acc + (long)(n / d->tavg) * nbeam;
          for (int b = 0; b < nbeam; b++) {
            const float *vr = wr + (long)b * nant, *vi = wi + (long)b * nant;
            float br = 0.f, bi = 0.f;
            for (int a = 0; a < nant; a++) {
              br += vr[a] * xr[a] - vi[a] * xi[a];
              bi += vi[a] * xr[a] + vr[a] * xi[a];
            }
            arow[b] += br * br + bi * bi;
          }
        }
      }
    }
    for (int o = 0; o < nout; o++)
      for (int b = 0; b < nbeam; b++) {
        float x = acc[(long)o * nbeam + b];
        long idx = ((long)o * nbeam + b) * ng + g;
        if (t->power) t->power[idx] = x;
        if (t->out) t->out[idx] = orc_requant(x, t->scale[b]);
      }
  }
  free(dr);
  free(di);
  free(wr);
  free(wi);
  free(acc);
  return NULL;
}

double cpu_beamform(const orc_dims *d, int wmode, const uint8_t *in,
                    const float *W, const float *scale, int g0, int g1,
                    float *power, uint8_t *out, int nthreads) {
  int nthr = nthreads > 0 ? nthreads : (int)sysconf(_SC_NPROCESSORS_ONLN);
  if (nthr < 1) nthr = 1;
  if (nthr > g1 - g0) nthr = g1 - g0;
  if (nthr > 256) nthr = 256;
  pthread_t *tid = malloc(sizeof(pthread_t) * nthr);
  cpu_thr *arg = malloc(sizeof(cpu_thr) * nthr);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int i = 0; i < nthr; i++) {
    arg[i].d = d;
    arg[i].wmode = wmode;
    arg[i].in = in;
    arg[i].W = W;
    arg[i].scale = scale;
    arg[i].power = power;
    arg[i].out = out;
    arg[i].g0 = g0;
    arg[i].g1 = g1;
    arg[i].stride = nthr;
    arg[i].first = i;
    pthread_create(&tid[i], NULL, cpu_worker, &arg[i]);
  }
  for (int i = 0; i < nthr; i++) pthread_join(tid[i], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  free(tid);
  free(arg);
  return (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}

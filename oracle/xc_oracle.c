/*
 * xc_oracle.c -- CPU restatement of the DSA correlator X-engine
 * (src/dsaX_bfCorr.cu dcorrelator) on the beamformer's 4-bit input.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Loaded only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * The reference reorders the block [packet][antenna][channel][2 times]
 * [2 pols] (dsaX_bfCorr.cu:314-356) into [channel][time][pol][packet]
 * [antenna] half planes, forms per (channel, time, pol) the column-major Gram
 * matrices C = r r^T + i i^T and C' = r i^T - i r^T with four strided batched
 * GEMMs (:512-556, i.e. C(a, a') = sum conj(d_a) d_a'), transposes them
 * (:392-403) and keeps the upper triangle (i, j <= i) at C(j, i), summing
 * both times and the four packet quarters (:366-385, :430-441).  Output
 * [baseline][channel][2 pols][re, im] float.  Here the sums are exact
 * integers (the reference accumulates in FP16).
 */
#include <pthread.h>
#include <stdlib.h>
#include <unistd.h>

#include "bf_oracle.h"

long orc_baseline(int i, int j) { return (long)i * (i + 1) / 2 + j; }

typedef struct {
  const orc_dims *d;
  const uint8_t *in;
  double *vis;
  int c0, c1;
} xc_job;

static void *xc_worker(void *p) {
  const xc_job *j = p;
  const orc_dims *d = j->d;
  const int na = d->nant;
  int *re = malloc(sizeof(int) * na), *im = malloc(sizeof(int) * na);
  long long *sr = malloc(sizeof(long long) * na * na), *si = malloc(sizeof(long long) * na * na);
  for (int c = j->c0; c < j->c1; c++)
    for (int q = 0; q < 2; q++) {
      for (long k = 0; k < (long)na * na; k++) sr[k] = si[k] = 0;
      for (int p = 0; p < d->npkt; p++)
        for (int t = 0; t < 2; t++) {
          for (int a = 0; a < na; a++) orc_unpack_byte(j->in[orc_in_index(d, p, a, c, t, q)], &re[a], &im[a]);
          /* conj(d_jj) * d_ii for jj <= ii */
          for (int ii = 0; ii < na; ii++)
            for (int jj = 0; jj <= ii; jj++) {
              sr[(long)ii * na + jj] += (long long)re[jj] * re[ii] + (long long)im[jj] * im[ii];
              si[(long)ii * na + jj] += (long long)re[jj] * im[ii] - (long long)im[jj] * re[ii];
            }
        }
      for (int ii = 0; ii < na; ii++)
        for (int jj = 0; jj <= ii; jj++) {
          double *v = j->vis + ((orc_baseline(ii, jj) * d->nchan + c) * 2 + q) * 2;
          v[0] = (double)sr[(long)ii * na + jj];
          v[1] = (double)si[(long)ii * na + jj];
        }
    }
  free(re);
  free(im);
  free(sr);
  free(si);
  return NULL;
}

void orc_xcorr(const orc_dims *d, const uint8_t *in, double *vis) {
  int nthr = (int)sysconf(_SC_NPROCESSORS_ONLN);
  if (nthr < 1) nthr = 1;
  if (nthr > 64) nthr = 64;
  if (nthr > d->nchan) nthr = d->nchan;
  pthread_t tid[64];
  xc_job job[64];
  for (int k = 0; k < nthr; k++) {
    job[k].d = d;
    job[k].in = in;
    job[k].vis = vis;
    job[k].c0 = (int)((long)d->nchan * k / nthr);
    job[k].c1 = (int)((long)d->nchan * (k + 1) / nthr);
    pthread_create(&tid[k], NULL, xc_worker, &job[k]);
  }
  for (int k = 0; k < nthr; k++) pthread_join(tid[k], NULL);
}

/*
 * fl_host.c -- the flagger's CPU functions (gpu_flagger.cu:530-790): medians,
 * median absolute deviations and the channel / time masks derived from the
 * GPU statistics.
 *
 * The reference sorts a copy by bubble sort for every median (medval,
 * :532-552), 64 x 3 x 768-element sorts plus 49k 21-element sorts per block;
 * here a median is a selection (Hoare partitioning, insertion sort for short
 * runs), or for channflag's 21-channel windows a sorted window that slides
 * along the band, each returning the same element of the sorted order.  Every float and
 * double operation of the mask tests is kept as the reference writes it.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "dsaxfl.h"

#define FL_Z DSAXFL_ZERO_CHANNELS
#define FL_FILT 21

/* k-th smallest of a[0..n-1] (reorders a). */
static float select_k(float *a, int n, int k) {
  int lo = 0, hi = n - 1;
  while (hi - lo > 16) {
    const int mid = lo + (hi - lo) / 2;
    /* median of three as pivot */
    float x = a[lo], y = a[mid], z = a[hi];
    float p = x < y ? (y < z ? y : (x < z ? z : x)) : (x < z ? x : (y < z ? z : y));
    int i = lo, j = hi;
    while (i <= j) {
      while (a[i] < p) i++;
      while (a[j] > p) j--;
      if (i <= j) {
        float t = a[i];
        a[i] = a[j];
        a[j] = t;
        i++;
        j--;
      }
    }
    if (k <= j) hi = j;
    else if (k >= i) lo = i;
    else return a[k];
  }
  for (int i = lo + 1; i <= hi; i++) {
    float v = a[i];
    int j = i - 1;
    while (j >= lo && a[j] > v) {
      a[j + 1] = a[j];
      j--;
    }
    a[j + 1] = v;
  }
  return a[k];
}

static float median_of(const float *a, int n, float *tmp) {
  memcpy(tmp, a, sizeof(float) * n);
  return select_k(tmp, n, (n + 1) / 2 - 1);
}

/* Beams [0, nbeam) in up to 8 host threads, each a contiguous range: every
 * beam's arithmetic is its own, so the masks do not depend on the split. */
typedef void (*fl_beam_fn)(const void *ctx, int b0, int b1);
struct fl_range {
  fl_beam_fn fn;
  const void *ctx;
  int b0, b1;
};
static void *fl_range_run(void *p) {
  const struct fl_range *r = p;
  r->fn(r->ctx, r->b0, r->b1);
  return NULL;
}
static void fl_for_beams(int nbeam, fl_beam_fn fn, const void *ctx) {
  const long ncpu = sysconf(_SC_NPROCESSORS_ONLN);
  int nt = nbeam / 4 < 8 ? nbeam / 4 : 8;
  if (ncpu > 0 && nt > ncpu) nt = (int)ncpu;
  if (nt <= 1) {
    fn(ctx, 0, nbeam);
    return;
  }
  pthread_t th[8];
  struct fl_range r[8];
  int live[8] = {0};
  for (int t = 0; t < nt; t++) r[t] = (struct fl_range){fn, ctx, (int)((long)nbeam * t / nt), (int)((long)nbeam * (t + 1) / nt)};
  for (int t = 1; t < nt; t++) live[t] = pthread_create(&th[t], NULL, fl_range_run, &r[t]) == 0;
  fn(ctx, r[0].b0, r[0].b1);
  for (int t = 1; t < nt; t++) {
    if (live[t]) pthread_join(th[t], NULL);
    else fn(ctx, r[t].b0, r[t].b1);            /* no thread: this one does it */
  }
}

float dsaxfl_medval(const float *a, int n) {
  if (n < 1) return 0.f;
  float *tmp = malloc(sizeof(float) * n);
  if (!tmp) return NAN;
  float r = median_of(a, n, tmp);
  free(tmp);
  return r;
}

/* The 21-channel window medians channflag subtracts: window j starts at
 * channel j, or at nchan - 128 - 21 for the windows that would pass channel
 * nchan - 129.  Consecutive windows differ by one channel, so one sorted
 * window slides across the band (one removal and one insertion a step, where
 * selecting each window afresh costs a partition of all 21); its element 10
 * is the element medval picks. */
static void window_medians(const float *row, int nchan, float *med) {
  float w[FL_FILT];
  const int last = nchan - FL_Z - FL_FILT;
  for (int i = 0; i < FL_FILT; i++) {          /* the first window, sorted */
    const float v = row[FL_Z + i];
    int k = i - 1;
    while (k >= 0 && w[k] > v) {
      w[k + 1] = w[k];
      k--;
    }
    w[k + 1] = v;
  }
  int start = FL_Z;
  for (int j = FL_Z; j < nchan - FL_Z; j++) {
    const int w0 = j <= last ? j : last;
    for (; start < w0; start++) {
      /* drop row[start] (any element equal to it), then insert row[start + 21] */
      const float out = row[start], in = row[start + FL_FILT];
      int i = 0;
      while (i < FL_FILT - 1 && w[i] != out) i++;
      for (; i < FL_FILT - 1; i++) w[i] = w[i + 1];
      int k = FL_FILT - 2;
      while (k >= 0 && w[k] > in) {
        w[k + 1] = w[k];
        k--;
      }
      w[k + 1] = in;
    }
    med[j - FL_Z] = w[(FL_FILT + 1) / 2 - 1];
  }
}

struct fl_chan_ctx {
  const float *spec;
  int nchan;
  float thr;
  const int *perm;
  int nperm;
  int *mask;
};
static void channflag_beams(const void *vc, int b0, int b1) {
  const struct fl_chan_ctx *c = vc;
  const int nchan = c->nchan, nin = nchan - 2 * FL_Z;
  float *cs = malloc(sizeof(float) * nin), *tmp = malloc(sizeof(float) * nin);
  if (!cs || !tmp) {
    free(cs);
    free(tmp);
    return;
  }
  for (int b = b0; b < b1; b++) {
    const float *row = c->spec + (size_t)b * nchan;
    window_medians(row, nchan, tmp);
    for (int j = FL_Z; j < nchan - FL_Z; j++) cs[j - FL_Z] = row[j] - tmp[j - FL_Z];
    const float med = median_of(cs, nin, tmp);
    for (int j = 0; j < nin; j++) tmp[j] = fabsf(cs[j] - med);
    const float mad = select_k(tmp, nin, (nin + 1) / 2 - 1);
    const float lim = c->thr * mad;
    int *m = c->mask + (size_t)b * nchan;
    for (int j = FL_Z; j < nchan - FL_Z; j++) {
      const float v = cs[j - FL_Z];
      if (v > lim || v < -lim) m[j] = 1;
    }
    for (int k = 0; k < c->nperm; k++)
      if (c->perm[k] >= FL_Z && c->perm[k] < nchan - FL_Z) m[c->perm[k]] = 1;
  }
  free(cs);
  free(tmp);
}

void dsaxfl_channflag(const float *spec, int nbeam, int nchan, float thr, const int *perm, int nperm,
                      int *mask) {
  if (nchan - 2 * FL_Z < FL_FILT) return;
  const struct fl_chan_ctx c = {spec, nchan, thr, perm, nperm, mask};
  fl_for_beams(nbeam, channflag_beams, &c);
}

static void simple_channflag_beams(const void *vc, int b0, int b1) {
  const struct fl_chan_ctx *c = vc;
  const int nchan = c->nchan, nin = nchan - 2 * FL_Z;
  float *tmp = malloc(sizeof(float) * nin);
  if (!tmp) return;
  const float mythr = (float)(c->thr / sqrt(1. * DSAXFL_FBIN));
  for (int b = b0; b < b1; b++) {
    const float *row = c->spec + (size_t)b * nchan;
    const float med = median_of(row + FL_Z, nin, tmp);
    for (int j = 0; j < nin; j++) tmp[j] = fabsf(row[FL_Z + j] - med);
    const float mad = select_k(tmp, nin, (nin + 1) / 2 - 1);
    const float lim = mythr * mad;
    for (int j = FL_Z; j < nchan - FL_Z - DSAXFL_FBIN; j++) {
      float vv = 0.f;
      for (int k = 0; k < DSAXFL_FBIN; k++) vv += row[j];   /* the reference adds spec[j] FBIN times */
      vv = (float)((double)vv / (1. * DSAXFL_FBIN) - (double)med);
      if (vv > lim) c->mask[(size_t)b * nchan + j] = 1;
    }
  }
  free(tmp);
}

void dsaxfl_simple_channflag(const float *spec, int nbeam, int nchan, float thr, int *mask) {
  if (nchan - 2 * FL_Z < 1) return;
  const struct fl_chan_ctx c = {spec, nchan, thr, NULL, 0, mask};
  fl_for_beams(nbeam, simple_channflag_beams, &c);
}

static void tsflag_beams(const void *vc, int b0, int b1) {
  const struct fl_chan_ctx *c = vc;                  /* nchan = ntime here */
  const int ntime = c->nchan, nm = ntime / 16;
  float *tmp = malloc(sizeof(float) * nm);
  if (!tmp) return;
  for (int b = b0; b < b1; b++) {
    const float *row = c->spec + (size_t)b * ntime;
    const float med = median_of(row, nm, tmp);
    for (int j = 0; j < nm; j++) tmp[j] = fabsf(row[j] - med);
    const float lim = c->thr * select_k(tmp, nm, (nm + 1) / 2 - 1);
    for (int j = 0; j < ntime; j++)
      if (row[j] - med > lim) c->mask[(size_t)b * ntime + j] = 1;
  }
  free(tmp);
}

void dsaxfl_simple_tsflag(const float *ts, int nbeam, int ntime, float thr, int *mask) {
  if (ntime / 16 < 1) return;
  const struct fl_chan_ctx c = {ts, ntime, thr, NULL, 0, mask};
  fl_for_beams(nbeam, tsflag_beams, &c);
}

int dsaxfl_gather_mask(const int *mask, long n, int *idx) {
  int k = 0;
  for (long i = 0; i < n; i++)
    if (mask[i] == 1) idx[k++] = (int)i;
  return k;
}

void dsaxfl_calc_bpwr(const float *spec, int nbeam, int nchan, float *bpwr) {
  for (int b = 0; b < nbeam; b++) {
    float s = 0.f;
    for (int j = 0; j < nchan; j++) s += spec[(size_t)b * nchan + j];
    bpwr[b] = (float)(((double)s / (1. * (nchan - 2 * FL_Z))) / DSAXFL_MN);
  }
}

void dsaxfl_median_calc(float *arr, int nbeam, int nchan) {
  float *col = malloc(sizeof(float) * nbeam);
  if (!col) return;
  for (int c = 0; c < nchan; c++) {
    for (int b = 0; b < nbeam; b++) col[b] = arr[(size_t)b * nchan + c];
    for (int i = 1; i < nbeam; i++) {        /* insertion sort: stable, as the bubble sort */
      float v = col[i];
      int j = i - 1;
      while (j >= 0 && col[j] > v) {
        col[j + 1] = col[j];
        j--;
      }
      col[j + 1] = v;
    }
    for (int b = 0; b < nbeam; b++) arr[(size_t)b * nchan + c] = col[b];
  }
  free(col);
}

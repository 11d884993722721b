/*
 * fl_oracle.c -- CPU restatement of the search-side RFI flagger
 * (gpu_flagger.cu) on a [nbeam][ntime][nchan] uint8 filterbank block.
 *
 * TEST INFRASTRUCTURE ONLY: the checker, never the product.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * The kernels are restated thread by thread as the reference launches them
 * (32 threads per (beam, channel) or per (beam, time), each over a contiguous
 * run, then the shared-memory trees).  The trees run in warp lockstep: in a
 * step every active thread reads both operands before any writes, which is
 * what one 32-wide warp does and what fixes the results of the
 * unsynchronised max/min trees of calc_maxspec and calc_ppspec.  The host
 * functions (medval, channflag, simple_channflag, simple_tsflag, gather_mask,
 * calc_bpwr, median_calc) are restated statement by statement, in the same
 * float/double mix.  The reference's sizes (64 beams, 16384 times, 1024
 * channels) are arguments here.
 *
 * Pins (tests/test_fl_oracle.py): spectrum and make_ts against numpy integer
 * sums (exact); varspec against the exact FP64 variance and a float64 numpy
 * restatement; maxspec against numpy max; the lockstep ppspec against a
 * separate pure-Python warp simulation and against hand cases whose result
 * the tree fixes; scaley / sumpulse / flag / flagts against numpy restatements
 * and hand edge values (rounding ties, wrap, NaN, negatives); fix_zspec
 * against numpy medians and the tie cases where the reference writes nothing;
 * the host functions against numpy restatements and hand cases.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fl_oracle.h"

#define FL_MN 48.0
#define FL_SIG 6.0
#define FL_TBIN 128
#define FL_FBIN 8
#define FL_RMAX 16384
#define FL_NT 32        /* threads per block in every reference launch */

/* One shared-memory tree over 32 slots, threads 0..15 active in every step,
 * lockstep (all reads of a step before its writes). op 0 sum, 1 max (replace
 * when csum[t] < csum[t+a]), 2 min (replace when csum[t] > csum[t+a]). */
static void lockstep_tree(float *c, int op) {
  for (int a = 16; a > 0; a /= 2) {
    float nv[16];
    for (int t = 0; t < 16; t++) {
      float x = c[t], y = c[t + a];
      nv[t] = op == 0 ? x + y : op == 1 ? (x < y ? y : x) : (x > y ? y : x);
    }
    for (int t = 0; t < 16; t++) c[t] = nv[t];
  }
}

/* (float)rint-to-even conversion to unsigned with the PTX cvt saturation
 * (NaN -> 0, negative -> 0, >= 2^32 -> 2^32-1), then /2 and the uchar store:
 * (unsigned char)((__float2uint_rn(2.*val))/2) of gpu_flagger.cu:390,405. */
static uint8_t fl_u8(float val) {
  float v2 = (float)(2.0 * (double)val);
  float r = rintf(v2);
  uint32_t u;
  if (!(r > 0.f)) u = 0;
  else if (r >= 4294967296.f) u = 0xFFFFFFFFu;
  else u = (uint32_t)r;
  return (uint8_t)(u / 2);
}

/* calc_spectrum (gpu_flagger.cu:174-222). */
void orc_fl_spectrum(const uint8_t *data, int nb, int nt, int nc, float *spec) {
  const int per = nt / FL_NT;
  for (int bm = 0; bm < nb; bm++)
    for (int ch = 0; ch < nc; ch++) {
      float c[FL_NT];
      for (int t = 0; t < FL_NT; t++) {
        c[t] = 0.f;
        long idx = ((long)bm * nt + (long)t * per) * nc + ch;
        for (int k = 0; k < per; k++, idx += nc) c[t] += (float)data[idx];
      }
      lockstep_tree(c, 0);
      spec[(long)bm * nc + ch] = (float)((double)c[0] / (1. * nt));
    }
}

/* calc_varspec (gpu_flagger.cu:228-278): val = data - spectrum in float and
 * csum += val*val, contracted to one fused multiply-add as nvcc does. */
void orc_fl_varspec(const uint8_t *data, int nb, int nt, int nc, const float *spec, float *var) {
  const int per = nt / FL_NT;
  for (int bm = 0; bm < nb; bm++)
    for (int ch = 0; ch < nc; ch++) {
      const float m = spec[(long)bm * nc + ch];
      float c[FL_NT];
      for (int t = 0; t < FL_NT; t++) {
        c[t] = 0.f;
        long idx = ((long)bm * nt + (long)t * per) * nc + ch;
        for (int k = 0; k < per; k++, idx += nc) {
          float val = (float)data[idx] - m;
          c[t] = fmaf(val, val, c[t]);
        }
      }
      lockstep_tree(c, 0);
      var[(long)bm * nc + ch] = (float)((double)c[0] / (1. * nt));
    }
}

/* The variance the reference's varspec approximates, from exact integer
 * sums: sum (x - s)^2 = [(n Q - S^2) + (S - s n)^2] / n with S = sum x,
 * Q = sum x^2; exact (then rounded once) when s is the mean S/n. */
void orc_fl_varspec_exact(const uint8_t *data, int nb, int nt, int nc, const float *spec, float *var) {
  for (int bm = 0; bm < nb; bm++)
    for (int ch = 0; ch < nc; ch++) {
      int64_t S = 0, Q = 0;
      for (int t = 0; t < nt; t++) {
        int64_t x = data[((long)bm * nt + t) * nc + ch];
        S += x;
        Q += x * x;
      }
      const double n = (double)nt;
      volatile double sn = (double)spec[(long)bm * nc + ch] * n;
      volatile double dl = (double)S - sn;
      volatile double d2 = dl * dl;
      volatile double num = (double)((int64_t)nt * Q - S * S) + d2;
      var[(long)bm * nc + ch] = (float)(num / n / n);
    }
}

/* calc_maxspec (gpu_flagger.cu:282-318). */
void orc_fl_maxspec(const uint8_t *data, int nb, int nt, int nc, float *mx) {
  const int per = nt / FL_NT;
  for (int bm = 0; bm < nb; bm++)
    for (int ch = 0; ch < nc; ch++) {
      float c[FL_NT];
      for (int t = 0; t < FL_NT; t++) {
        float val = 0.f;
        long idx = ((long)bm * nt + (long)t * per) * nc + ch;
        for (int k = 0; k < per; k++, idx += nc)
          if ((float)data[idx] > val) val = (float)data[idx];
        c[t] = val;
      }
      lockstep_tree(c, 1);
      mx[(long)bm * nc + ch] = c[0];
    }
}

/* calc_ppspec (gpu_flagger.cu:323-375): each thread's largest 128-sample bin
 * mean; the max tree; then the min tree over the array the max tree left. */
void orc_fl_ppspec(const uint8_t *data, int nb, int nt, int nc, float *pp) {
  const int per = nt / FL_NT;
  for (int bm = 0; bm < nb; bm++)
    for (int ch = 0; ch < nc; ch++) {
      float c[FL_NT];
      for (int t = 0; t < FL_NT; t++) {
        float val = 0.f;
        for (int j = 0; j < per / FL_TBIN; j++) {
          long idx = ((long)bm * nt + (long)t * per + (long)j * FL_TBIN) * nc + ch;
          float vv = 0.f;
          for (int i = 0; i < FL_TBIN; i++, idx += nc) vv += (float)data[idx];
          vv = (float)((double)vv / (1. * FL_TBIN));
          if (vv > val) val = vv;
        }
        c[t] = val;
      }
      lockstep_tree(c, 1);
      const float v1 = c[0];
      lockstep_tree(c, 2);
      pp[(long)bm * nc + ch] = v1 - c[0];
    }
}

/* scaley (gpu_flagger.cu:381-393): val = (x - spec) * (SIG / sqrtf(var)) + MN
 * with SIG and MN double literals, so the scale and the multiply-add are
 * double (the multiply-add fused as nvcc contracts it), x - spec is float, and
 * the result is stored to a float. */
void orc_fl_scaley(uint8_t *data, int nb, int nt, int nc, const float *spec, const float *var) {
  const long n = (long)nb * nt * nc;
  for (long idx = 0; idx < n; idx++) {
    const long bm = idx / ((long)nt * nc), ch = idx % nc, sp = bm * nc + ch;
    const double k = FL_SIG / (double)sqrtf(var[sp]);
    const float val = (float)fma((double)((float)data[idx] - spec[sp]), k, FL_MN);
    data[idx] = fl_u8(val);
  }
}

/* sumpulse (gpu_flagger.cu:398-406). */
void orc_fl_sumpulse(uint8_t *data, long n, const float *summand) {
  for (long i = 0; i < n; i++) data[i] = fl_u8((float)data[i] + summand[i]);
}

/* make_ts (gpu_flagger.cu:413-452). */
void orc_fl_make_ts(const uint8_t *data, int nb, int nt, int nc, float *ts) {
  const int per = nc / FL_NT;
  for (int bm = 0; bm < nb; bm++)
    for (int tm = 0; tm < nt; tm++) {
      float c[FL_NT];
      for (int t = 0; t < FL_NT; t++) {
        c[t] = 0.f;
        const long idx = ((long)bm * nt + tm) * nc + (long)t * per;
        for (int k = 0; k < per; k++) c[t] += (float)data[idx + k];
      }
      lockstep_tree(c, 0);
      ts[(long)bm * nt + tm] = (float)((double)c[0] / (1. * nc));
    }
}

/* The replacement byte of flag/flagts: MN * bpwr[bm] is a double stored to an
 * unsigned char (truncation; out of [0, 256) clamped, NaN -> 0). */
static uint8_t fl_rep(float bpwr) {
  const double v = FL_MN * (double)bpwr;
  if (!(v > 0.)) return 0;
  if (v >= 255.) return 255;
  return (uint8_t)v;
}

/* flag (gpu_flagger.cu:458-474): every time of each listed (beam, channel). */
void orc_fl_flag(uint8_t *data, int nb, int nt, int nc, const int *midx, int n_mask, const float *bpwr) {
  (void)nb;
  for (int m = 0; m < n_mask; m++) {
    const int bm = midx[m] / nc, ch = midx[m] % nc;
    const uint8_t v = fl_rep(bpwr[bm]);
    for (int tm = 0; tm < nt; tm++) data[((long)bm * nt + tm) * nc + ch] = v;
  }
}

/* flagts (gpu_flagger.cu:478-494): channels 128 .. nchan-129 of each listed
 * (beam, time). */
void orc_fl_flagts(uint8_t *data, int nb, int nt, int nc, const int *midx, int n_mask, const float *bpwr) {
  (void)nb;
  for (int m = 0; m < n_mask; m++) {
    const int bm = midx[m] / nt, tm = midx[m] % nt;
    const uint8_t v = fl_rep(bpwr[bm]);
    for (int ch = 128; ch < nc - 128; ch++) data[((long)bm * nt + tm) * nc + ch] = v;
  }
}

/* genrand (gpu_flagger.cu:499-520) with this build's own counter-based
 * generator in place of curand: r = splitmix64 finaliser of
 * (seed * 2^32 + i); u1, u2 = (low, high 32 bits % 16384) / 16384; Box-Muller
 * in FP64; byte = (uchar)(rne(2 fma(va, 6, 48)) / 2) with the saturating
 * conversion. */
uint64_t orc_fl_mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void orc_fl_genrand(uint8_t *repval, long n, uint32_t seed) {
  for (long i = 0; i < n; i++) {
    const uint64_t r = orc_fl_mix(((uint64_t)seed << 32) + (uint64_t)i);
    const double u1 = (double)((uint32_t)r % FL_RMAX) / (1. * FL_RMAX);
    const double u2 = (double)((uint32_t)(r >> 32) % FL_RMAX) / (1. * FL_RMAX);
    const double va = sqrt(-2. * log(u1)) * cos(2. * M_PI * u2);
    repval[i] = fl_u8((float)fma(va, FL_SIG, FL_MN));
  }
}

/* fix_zspec (gpu_flagger.cu:113-168): for each of the nbc (beam, channel)
 * entries, the value among s0[k*nbc + i], k < naver, with exactly naver/2
 * others <= it replaces s0[i] (likewise v0); when none qualifies (ties),
 * s0[i] keeps block 0's value. */
void orc_fl_fix_zspec(float *s0, float *v0, int nbc, int naver) {
  const int place = naver / 2;
  float *vec = malloc(sizeof(float) * 2 * naver);
  for (int i = 0; i < nbc; i++) {
    for (int k = 0; k < naver; k++) {
      vec[k] = s0[(long)k * nbc + i];
      vec[naver + k] = v0[(long)k * nbc + i];
    }
    for (int h = 0; h < 2; h++) {
      float *w = vec + h * naver;
      for (int t = 0; t < naver; t++) {
        int ct = 0;
        for (int k = 0; k < naver; k++)
          if (k != t && w[k] <= w[t]) ct++;
        if (ct == place) (h ? v0 : s0)[i] = w[t];
      }
    }
  }
  free(vec);
}

/* ---- host functions ---------------------------------------------------- */

/* medval (gpu_flagger.cu:532-552): bubble sort, element (n+1)/2 - 1. */
float orc_fl_medval(const float *a, int n) {
  float *tmp = malloc(sizeof(float) * (n > 0 ? n : 1));
  for (int i = 0; i < n; i++) tmp[i] = a[i];
  for (int i = 0; i < n - 1; i++)
    for (int j = 0; j < n - i - 1; j++)
      if (tmp[j] > tmp[j + 1]) {
        float tt = tmp[j + 1];
        tmp[j + 1] = tmp[j];
        tmp[j] = tt;
      }
  float r = tmp[(n + 1) / 2 - 1];
  free(tmp);
  return r;
}

/* simple_channflag (gpu_flagger.cu:558-605), its FBIN loop adding the same
 * element FBIN times included. */
void orc_fl_simple_channflag(const float *spec, int nb, int nc, float thr, int *mask) {
  const int Z = 128, nin = nc - 2 * Z;
  float *med = malloc(sizeof(float) * nb), *mad = malloc(sizeof(float) * nb);
  float *norm = malloc(sizeof(float) * nin);
  for (int i = 0; i < nb; i++) med[i] = orc_fl_medval(spec + (long)i * nc + Z, nin);
  for (int i = 0; i < nb; i++) {
    for (int j = Z; j < nc - Z; j++) norm[j - Z] = fabsf(spec[(long)i * nc + j] - med[i]);
    mad[i] = orc_fl_medval(norm, nin);
  }
  const float mythr = (float)(thr / sqrt(1. * FL_FBIN));
  for (int i = 0; i < nb; i++)
    for (int j = Z; j < nc - Z - FL_FBIN; j++) {
      float vv = 0.f;
      for (int k = 0; k < FL_FBIN; k++) vv += spec[(long)i * nc + j];
      vv = (float)((double)vv / (1. * FL_FBIN) - (double)med[i]);
      if (vv > mythr * mad[i]) mask[(long)i * nc + j] = 1;
    }
  free(med);
  free(mad);
  free(norm);
}

/* simple_tsflag (gpu_flagger.cu:607-648): median and MAD of the first
 * ntime/16 samples of each beam's series, one-sided threshold. */
void orc_fl_simple_tsflag(const float *ts, int nb, int nt, float thr, int *mask) {
  const int nm = nt / 16;
  float *med = malloc(sizeof(float) * nb), *mad = malloc(sizeof(float) * nb);
  float *norm = malloc(sizeof(float) * nm);
  for (int i = 0; i < nb; i++) med[i] = orc_fl_medval(ts + (long)i * nt, nm);
  for (int i = 0; i < nb; i++) {
    for (int j = 0; j < nm; j++) norm[j] = fabsf(ts[(long)i * nt + j] - med[i]);
    mad[i] = orc_fl_medval(norm, nm);
  }
  for (int i = 0; i < nb; i++)
    for (int j = 0; j < nt; j++) {
      float vv = ts[(long)i * nt + j] - med[i];
      if (vv > thr * mad[i]) mask[(long)i * nt + j] = 1;
    }
  free(med);
  free(mad);
  free(norm);
}

/* channflag (gpu_flagger.cu:651-718): forward 21-channel median filter
 * (the last full window near the top edge), median and MAD over channels
 * 128 .. nchan-129, two-sided threshold, permanent channels. */
void orc_fl_channflag(const float *spec, int nb, int nc, float thr, const int *perm, int nperm,
                      int *mask) {
  const int Z = 128, F = 21, nin = nc - 2 * Z;
  float *cs = calloc((size_t)nb * nc, sizeof(float));
  float *med = malloc(sizeof(float) * nb), *mad = malloc(sizeof(float) * nb);
  float *norm = malloc(sizeof(float) * nin);
  for (int i = 0; i < nb; i++)
    for (int j = Z; j < nc - Z; j++) {
      const float *row = spec + (long)i * nc;
      if (nc - Z - j >= F) cs[(long)i * nc + j] = row[j] - orc_fl_medval(row + j, F);
      else cs[(long)i * nc + j] = row[j] - orc_fl_medval(row + nc - Z - F, F);
    }
  for (int i = 0; i < nb; i++) med[i] = orc_fl_medval(cs + (long)i * nc + Z, nin);
  for (int i = 0; i < nb; i++) {
    for (int j = Z; j < nc - Z; j++) norm[j - Z] = fabsf(cs[(long)i * nc + j] - med[i]);
    mad[i] = orc_fl_medval(norm, nin);
  }
  for (int i = 0; i < nb; i++)
    for (int j = Z; j < nc - Z; j++) {
      const float v = cs[(long)i * nc + j];
      if (v > thr * mad[i] || v < -thr * mad[i]) mask[(long)i * nc + j] = 1;
      for (int k = 0; k < nperm; k++)
        if (j == perm[k]) mask[(long)i * nc + j] = 1;
    }
  free(cs);
  free(med);
  free(mad);
  free(norm);
}

/* gather_mask / gather_tsmask (gpu_flagger.cu:722-748). */
int orc_fl_gather(const int *mask, long n, int *idx) {
  int k = 0;
  for (long i = 0; i < n; i++)
    if (mask[i] == 1) idx[k++] = (int)i;
  return k;
}

/* calc_bpwr (gpu_flagger.cu:752-763). */
void orc_fl_calc_bpwr(const float *spec, int nb, int nc, float *bpwr) {
  for (int i = 0; i < nb; i++) {
    bpwr[i] = 0.f;
    for (int j = 0; j < nc; j++) bpwr[i] += spec[(long)i * nc + j];
    bpwr[i] = (float)(((double)bpwr[i] / (1. * (nc - 256))) / FL_MN);
  }
}

/* median_calc (gpu_flagger.cu:767-790): each channel's values sorted over
 * beams, in place. */
void orc_fl_median_calc(float *arr, int nb, int nc) {
  for (int ch = 0; ch < nc; ch++)
    for (int i = 0; i < nb - 1; i++)
      for (int j = 0; j < nb - i - 1; j++)
        if (arr[(long)j * nc + ch] > arr[(long)(j + 1) * nc + ch]) {
          float tt = arr[(long)(j + 1) * nc + ch];
          arr[(long)(j + 1) * nc + ch] = arr[(long)j * nc + ch];
          arr[(long)j * nc + ch] = tt;
        }
}

/* The steady-state block step of the flagger loop (gpu_flagger.cu:1294-1400,
 * started == 1, no pulse): scaley with spec0/var0; spectrum and varspec; the
 * total-power test (noise replacement when |tpwr - prev|/prev >= mod_thresh);
 * otherwise the channel masks -- channflag on spectrum - oldspec (oldspec
 * stays zero in the reference), on varspec (thresh + 0.5) and on the "max"
 * spectrum (which the loop fills with calc_spectrum), simple_channflag on
 * ppspec -- bpwr (calc_bpwr when pwr, else the previous values), the random
 * replacement (mkrand) or flag, then make_ts / simple_tsflag / flagts when
 * tsflag.  exact_var: varspec from exact integer sums instead of the
 * reference's float order.  Returns the total power; *n_mask and *n_tsmask
 * report the counts (-1 when the block was noise-replaced). */
float orc_fl_step(uint8_t *data, int nb, int nt, int nc, const float *spec0, const float *var0,
                  const uint8_t *repval, float prev_tpwr, float mod_thresh, float thresh,
                  float tsthresh, int pwr, int mkrand, int tsflag, const int *perm, int nperm,
                  int exact_var, float *bpwr, int *n_mask, int *n_tsmask) {
  const long nbc = (long)nb * nc, nbt = (long)nb * nt;
  float *spec = malloc(sizeof(float) * nbc), *var = malloc(sizeof(float) * nbc);
  float *pp = malloc(sizeof(float) * nbc);
  orc_fl_scaley(data, nb, nt, nc, spec0, var0);
  orc_fl_spectrum(data, nb, nt, nc, spec);
  if (exact_var) orc_fl_varspec_exact(data, nb, nt, nc, spec, var);
  else orc_fl_varspec(data, nb, nt, nc, spec, var);
  float tpwr = 0.f;
  for (long i = 0; i < nbc; i++) tpwr += spec[i];
  *n_mask = *n_tsmask = -1;
  if (fabs(tpwr - prev_tpwr) / prev_tpwr >= mod_thresh) {
    for (int b = 0; b < nb; b++) memcpy(data + (long)b * nt * nc, repval, (size_t)nt * nc);
  } else {
    orc_fl_ppspec(data, nb, nt, nc, pp);
    int *mask = calloc(nbc, sizeof(int)), *idx = malloc(sizeof(int) * nbc);
    orc_fl_channflag(spec, nb, nc, thresh, perm, nperm, mask);
    orc_fl_channflag(var, nb, nc, thresh + 0.5f, perm, nperm, mask);
    orc_fl_channflag(spec, nb, nc, thresh, perm, nperm, mask);
    orc_fl_simple_channflag(pp, nb, nc, thresh, mask);
    if (pwr) orc_fl_calc_bpwr(spec, nb, nc, bpwr);
    *n_mask = orc_fl_gather(mask, nbc, idx);
    if (mkrand)
      for (int b = 0; b < nb; b++) memcpy(data + (long)b * nt * nc, repval, (size_t)nt * nc);
    else
      orc_fl_flag(data, nb, nt, nc, idx, *n_mask, bpwr);
    *n_tsmask = 0;
    if (tsflag) {
      float *ts = malloc(sizeof(float) * nbt);
      int *tmask = calloc(nbt, sizeof(int)), *tidx = malloc(sizeof(int) * nbt);
      orc_fl_make_ts(data, nb, nt, nc, ts);
      orc_fl_simple_tsflag(ts, nb, nt, tsthresh, tmask);
      *n_tsmask = orc_fl_gather(tmask, nbt, tidx);
      orc_fl_flagts(data, nb, nt, nc, tidx, *n_tsmask, bpwr);
      free(ts);
      free(tmask);
      free(tidx);
    }
    free(mask);
    free(idx);
  }
  free(spec);
  free(var);
  free(pp);
  return tpwr;
}

// bf_weights.hip -- beam weights on the device (replaces calc_weights,
// dsaX_beamformer.cu:420-449): FP64 phase exp(-2 pi i f ang x / c) times the
// calibration, scaled by a power of two per (channel, pol, beam) and written
// as FP16, FP8 E4M3 hi+lo or FP6 E2M3 hi+lo (or hi+mid+lo) MFMA B-operand
// fragments.
#include <stdarg.h>

#include "bf_common.h"

static thread_local char g_err[256];
extern "C" const char *bfk_last_error(void) { return g_err; }
void bfk_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
int bfk_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return 0;
  bfk_set_error("%s: %s", what, hipGetErrorString(e));
  return -3; /* DSAXBF_EHIP */
}

extern "C" size_t bfk_wfrag_bytes(const bfk_dims *d) {
  return (size_t)d->ncal * 2 * d->nbt * d->nkc * 64 * (d->prec == 3 ? 3 * WFRAG_LANE_BYTES / 2 : WFRAG_LANE_BYTES);
}
extern "C" size_t bfk_wscale_bytes(const bfk_dims *d) {
  return (size_t)d->ncal * 2 * d->nbeam_pad * sizeof(float);
}

// The largest k in 0..kmax with m * 2^k <= 7.5 (0 for m = 0).
__device__ static inline int lane_boost(double m, int kmax) {
  if (!(m > 0.0)) return 0;
  int k = (int)floor(log2(7.5 / m));
  k = k < 0 ? 0 : k > kmax ? kmax : k;
  while (k > 0 && ldexp(m, k) > 7.5) k--;
  while (k < kmax && ldexp(m, k + 1) <= 7.5) k++;
  return k;
}

// One thread per (weight channel wc, pol, beam).  The scale puts the largest
// component of the beam's weights in [2^14, 2^15) (FP16) or [128, 256) (FP8),
// or at exactly 7.5, the top of E2M3 (FP6); wscale holds the power factor
// that undoes it and the voltage code's unit (2^-9 for FP8, 1/8 for FP6).
// E2M3 has only two exponent bits, so FP6 also boosts each lane's 32 slots
// (16 antennas of one beam and K step) by the largest power of two that keeps
// them within 7.5, and each further part (lo; or mid, then lo) by the
// largest that keeps the lane's remaining residuals within 7.5; the lane's
// E8M0 MFMA operand scales take them back (their exponents stay >= 0).
__global__ void bf_weights_kernel(bfk_dims d, const float *__restrict__ antpos,
                                  const float *__restrict__ cal,
                                  const double *__restrict__ freqs,
                                  const double *__restrict__ theta,
                                  uint8_t *__restrict__ wfrag,
                                  float *__restrict__ wscale) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long total = (long)d.ncal * 2 * d.nbeam_pad;
  if (idx >= total) return;
  int b = (int)(idx % d.nbeam_pad);
  int q = (int)((idx / d.nbeam_pad) & 1);
  int wc = (int)(idx / (2L * d.nbeam_pad));
  int cq = wc * 2 + q;
  double f = d.wmode ? freqs[wc] : freqs[wc * 8 + 4];
  bool live = b < d.nbeam;
  double ang = live ? theta[b] : 0.0;
  if (d.amode) ang = sin(ang);
  double afac = -2.0 * M_PI * f * ang / 299792458.0;

  double mx = 0.0;
  if (live)
    for (int a = 0; a < d.nant; a++) {
      const float *cw = cal + (((long)a * d.ncal + wc) * 2 + q) * 2;
      double s, c;
      sincos(afac * (double)antpos[a], &s, &c);
      double wr = c * cw[0] - s * cw[1], wi = s * cw[0] + c * cw[1];
      mx = fmax(mx, fmax(fabs(wr), fabs(wi)));
    }
  int k = 0;
  if (mx > 0.0) frexp(mx, &k);
  int e = (d.prec == 0 ? 15 : 8) - k;
  double sc = ldexp(1.0, e);
  if (d.prec >= 2) sc = mx > 0.0 ? 7.5 / mx : 1.0;
  wscale[idx] = (float)(d.prec == 0   ? ldexp(1.0, -2 * e)
                        : d.prec == 1 ? ldexp(1.0, 18 - 2 * e)
                                      : 64.0 / (sc * sc));

  int bt = b >> 5, n = b & 31;
  for (int kc = 0; kc < d.nkc; kc++)
    for (int h = 0; h < 2; h++) {
      int lane = h * 32 + n;
      double wr[32], wi[32];
      for (int j = 0; j < 32; j++) {
        int a = kc * 64 + h * 32 + j;
        wr[j] = wi[j] = 0.0;
        if (live && a < d.nant) {
          const float *cw = cal + (((long)a * d.ncal + wc) * 2 + q) * 2;
          double s, c;
          sincos(afac * (double)antpos[a], &s, &c);
          wr[j] = (c * cw[0] - s * cw[1]) * sc;
          wi[j] = (s * cw[0] + c * cw[1]) * sc;
        }
      }
      if (d.prec == 0) {
        for (int s = 0; s < 8; s++) {
          h8v v;
          int j = 4 * s;
          v[0] = (_Float16)wr[j];     v[1] = (_Float16)wr[j + 1];
          v[2] = (_Float16)wi[j];     v[3] = (_Float16)wi[j + 1];
          v[4] = (_Float16)wr[j + 2]; v[5] = (_Float16)wr[j + 3];
          v[6] = (_Float16)wi[j + 2]; v[7] = (_Float16)wi[j + 3];
          long rec = (((long)(cq * d.nbt + bt) * d.nkc + kc) * 8 + s) * 64 + lane;
          *(h8v *)(wfrag + rec * 16) = v;
        }
      } else if (d.prec >= 2) {
        const int nparts = d.prec == 3 ? 3 : 2;
        for (int s2 = 0; s2 < 2; s2++) {
          double r[32];
          uint32_t code[3][8];
          int ktot = 0, sexp[3];
          for (int j = 0; j < 32; j++)      // slot j: Wr (j < 16) or Wi of antenna j % 16
            r[j] = j < 16 ? wr[16 * s2 + j] : wi[16 * s2 + j - 16];
          for (int pt = 0; pt < nparts; pt++) {
            double m = 0.0;
            for (int j = 0; j < 32; j++) m = fmax(m, fabs(r[j]));
            const int k = lane_boost(m, 127 - ktot < 60 ? 127 - ktot : 60);
            ktot += k;
            sexp[pt] = 127 - ktot;          // E8M0 scale 2^-ktot of this part
            for (int u = 0; u < 8; u++) code[pt][u] = 0;
            for (int j = 0; j < 32; j++) {
              const double rb = ldexp(r[j], k);
              const uint32_t c = e2m3_encode(rb);
              r[j] = rb - e2m3_decode(c);
              const int bit = 6 * j, dw = bit >> 5, sh = bit & 31;
              code[pt][dw] |= c << sh;
              if (sh > 26) code[pt][dw + 1] |= c >> (32 - sh);
            }
          }
          code[0][6] = (uint32_t)sexp[0];
          code[0][7] = (uint32_t)sexp[1];
          if (nparts == 3) code[1][6] = (uint32_t)sexp[2];
          long rec = ((((long)(cq * d.nbt + bt) * d.nkc + kc) * 2 + s2) * nparts) * 64 + lane;
          for (int pt = 0; pt < nparts; pt++) {
            uint32_t *pp = (uint32_t *)(wfrag + (rec + 64 * pt) * 32);
            for (int u = 0; u < 8; u++) pp[u] = code[pt][u];
          }
        }
      } else {
        for (int s2 = 0; s2 < 2; s2++) {
          uint32_t hi[8], lo[8];
          for (int u = 0; u < 4; u++) {
            uint32_t hr = 0, hi_i = 0, lr = 0, li = 0;
            for (int i = 0; i < 4; i++) {
              int j = 16 * s2 + 4 * u + i;
              uint32_t br = e4m3_encode(wr[j]), bi = e4m3_encode(wi[j]);
              uint32_t cr = e4m3_encode(wr[j] - e4m3_decode(br));
              uint32_t ci = e4m3_encode(wi[j] - e4m3_decode(bi));
              hr |= br << (8 * i);  hi_i |= bi << (8 * i);
              lr |= cr << (8 * i);  li |= ci << (8 * i);
            }
            hi[2 * u] = hr; hi[2 * u + 1] = hi_i;
            lo[2 * u] = lr; lo[2 * u + 1] = li;
          }
          long rec = ((((long)(cq * d.nbt + bt) * d.nkc + kc) * 2 + s2) * 2) * 64 + lane;
          uint32_t *ph = (uint32_t *)(wfrag + rec * 32);
          uint32_t *pl = (uint32_t *)(wfrag + (rec + 64) * 32);
          for (int u = 0; u < 8; u++) { ph[u] = hi[u]; pl[u] = lo[u]; }
        }
      }
    }
}

extern "C" int bfk_weights(const bfk_dims *d, const float *d_antpos, const float *d_cal,
                           const double *d_freqs, const double *d_theta, void *d_wfrag,
                           float *d_wscale, void *stream) {
  long total = (long)d->ncal * 2 * d->nbeam_pad;
  int nblk = (int)((total + 127) / 128);
  hipLaunchKernelGGL(bf_weights_kernel, dim3(nblk), dim3(128), 0, (hipStream_t)stream, *d,
                     d_antpos, d_cal, d_freqs, d_theta, (uint8_t *)d_wfrag, d_wscale);
  return bfk_check(hipGetLastError(), "bf_weights_kernel");
}

/*
 * bf_quant.c -- CPU restatement of the weight encodings the GPU path feeds
 * its matrix cores, for an oracle "fed the path's own quantised weights".
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Loaded only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * The reference stores its beam weights as FP16 (calc_weights,
 * dsaX_beamformer.cu:442-446, __float2half of the float weight) and feeds
 * them to FP16 wmma (:291-294).  The GPU path here keeps that form (precision
 * 0) and adds split forms in which a weight is the sum of two or three
 * narrow floating-point parts.  This file restates each encoding from the
 * format definitions alone:
 *   - code values from the OCP 8-bit (E4M3) and MX 6-bit (E2M3) format
 *     tables and IEEE binary16;
 *   - encoding as the nearest representable value by binary search over the
 *     monotone table of positive codes, ties to the even code, saturating at
 *     the largest finite value (binary16: overflow to infinity);
 *   - the power-of-two (or, for FP6, 7.5/max) row scale and the FP6 per-lane
 *     power-of-two boosts found by counting upwards.
 * It shares no code with the library's encoders.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "bf_oracle.h"

/* ---- code tables (positive codes, increasing values) ---------------- */
/* E4M3 (OCP "E4M3FN"): bias 7, exponent 0 subnormal m/8 * 2^-6, code 0x7F NaN. */
double orc_e4m3_decode(uint32_t c) {
  double s = (c & 0x80u) ? -1.0 : 1.0;
  unsigned e = (c >> 3) & 15u, m = c & 7u;
  if (e == 15 && m == 7) return NAN;
  return s * (e == 0 ? ldexp((double)m / 8.0, -6) : ldexp(1.0 + m / 8.0, (int)e - 7));
}

/* E2M3 (OCP MX FP6): bias 1, exponent 0 subnormal m/8, no NaN or infinity. */
double orc_e2m3_decode(uint32_t c) {
  double s = (c & 0x20u) ? -1.0 : 1.0;
  unsigned e = (c >> 3) & 3u, m = c & 7u;
  return s * (e == 0 ? (double)m / 8.0 : ldexp(1.0 + m / 8.0, (int)e - 1));
}

/* IEEE binary16: bias 15, exponent 0 subnormal m * 2^-24, 31 inf/NaN. */
double orc_f16_decode(uint32_t h) {
  double s = (h & 0x8000u) ? -1.0 : 1.0;
  unsigned e = (h >> 10) & 31u, m = h & 1023u;
  if (e == 31) return m ? NAN : s * INFINITY;
  return s * (e == 0 ? ldexp((double)m, -24) : ldexp(1.0 + m / 1024.0, (int)e - 15));
}

/* Nearest of the positive codes 0..cmax (values increasing with the code) to
 * a >= 0, ties to the even code; a beyond the largest value gives cmax. */
static uint32_t nearest_code(double a, uint32_t cmax, double (*dec)(uint32_t)) {
  if (!(a < dec(cmax))) return cmax;
  uint32_t lo = 0, hi = cmax;           /* dec(lo) <= a < dec(hi) */
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) / 2;
    if (dec(mid) <= a) lo = mid;
    else hi = mid;
  }
  double dl = a - dec(lo), dh = dec(hi) - a;
  if (dl < dh) return lo;
  if (dh < dl) return hi;
  return (lo & 1u) ? hi : lo;
}

uint32_t orc_e4m3_encode(double v) {
  uint32_t s = signbit(v) ? 0x80u : 0u;
  return s | nearest_code(fabs(v), 0x7Eu, orc_e4m3_decode);
}

uint32_t orc_e2m3_encode(double v) {
  uint32_t s = signbit(v) ? 0x20u : 0u;
  return s | nearest_code(fabs(v), 0x1Fu, orc_e2m3_decode);
}

uint32_t orc_f16_encode(double v) {
  uint32_t s = signbit(v) ? 0x8000u : 0u;
  double a = fabs(v);
  /* beyond the largest finite value plus half its spacing: infinity */
  if (a >= 65520.0) return s | 0x7C00u;
  return s | nearest_code(a, 0x7BFFu, orc_f16_decode);
}

/* ---- the path's weight encodings ------------------------------------ */
/* The power of two 2^e with 2^lo_exp <= mx * 2^e < 2^(lo_exp+1) (mx > 0). */
static double pow2_scale(double mx, int lo_exp) {
  int e = 0;
  while (ldexp(mx, e) >= ldexp(1.0, lo_exp + 1)) e--;
  while (ldexp(mx, e) < ldexp(1.0, lo_exp)) e++;
  return ldexp(1.0, e);
}

/* The largest k in 0..kmax with m * 2^k <= 7.5 (0 for m = 0). */
static int boost(double m, int kmax) {
  if (!(m > 0.0)) return 0;
  int k = 0;
  while (k < kmax && ldexp(m, k + 1) <= 7.5) k++;
  return k;
}

/* One FP6 lane: n values x (already row-scaled) -> their value through
 * nparts E2M3 parts, part p boosted by 2^k[p] over the previous one
 * (k[0] over x), the boosts capped so that the total stays within the
 * E8M0 exponent range (<= 127). */
static void fp6_lane(const double *x, int n, int nparts, double *xq) {
  double r[64], acc[64];
  int ktot = 0;
  for (int j = 0; j < n; j++) {
    r[j] = x[j];
    acc[j] = 0.0;
  }
  for (int p = 0; p < nparts; p++) {
    double m = 0.0;
    for (int j = 0; j < n; j++) m = fmax(m, fabs(r[j]));
    int cap = 127 - ktot < 60 ? 127 - ktot : 60;
    int k = boost(m, cap);
    ktot += k;
    for (int j = 0; j < n; j++) {
      double rb = ldexp(r[j], k);
      double q = orc_e2m3_decode(orc_e2m3_encode(rb));
      acc[j] += ldexp(q, -ktot);
      r[j] = rb - q;
    }
  }
  for (int j = 0; j < n; j++) xq[j] = acc[j];
}

int orc_quant_weights(const orc_dims *d, int wmode, int prec, const double *W, double *Wq) {
  if (prec < 0 || prec > 3) return -1;
  int ncal = wmode == 0 ? d->nchan / 8 : d->nchan;
  int nant = d->nant, nant_pad = (nant + 15) / 16 * 16;
  long nrow = (long)ncal * 2 * d->nbeam;
  double *x = malloc(sizeof(double) * 2 * nant_pad), *xq = malloc(sizeof(double) * 32);
  if (!x || !xq) {
    free(x);
    free(xq);
    return -2;
  }
  for (long row = 0; row < nrow; row++) {
    const double *w = W + row * nant * 2;
    double *wq = Wq + row * nant * 2;
    double mx = 0.0;
    for (int i = 0; i < 2 * nant; i++) mx = fmax(mx, fabs(w[i]));
    if (!(mx > 0.0)) {
      for (int i = 0; i < 2 * nant; i++) wq[i] = 0.0;
      continue;
    }
    if (prec == 0 || prec == 1) {
      double sc = pow2_scale(mx, prec == 0 ? 14 : 7);
      for (int i = 0; i < 2 * nant; i++) {
        double v = w[i] * sc, q;
        if (prec == 0) {
          q = orc_f16_decode(orc_f16_encode(v));
        } else {
          double hi = orc_e4m3_decode(orc_e4m3_encode(v));
          q = hi + orc_e4m3_decode(orc_e4m3_encode(v - hi));
        }
        wq[i] = q / sc;
      }
      continue;
    }
    /* FP6: largest component at 7.5; lanes of 16 antennas (their 16 real
     * and 16 imaginary parts) share the boosts */
    double sc = 7.5 / mx;
    for (int a0 = 0; a0 < nant; a0 += 16) {
      for (int j = 0; j < 16; j++) {
        int a = a0 + j;
        x[j] = a < nant ? w[2 * a] * sc : 0.0;
        x[16 + j] = a < nant ? w[2 * a + 1] * sc : 0.0;
      }
      fp6_lane(x, 32, prec == 2 ? 2 : 3, xq);
      for (int j = 0; j < 16 && a0 + j < nant; j++) {
        wq[2 * (a0 + j)] = xq[j] / sc;
        wq[2 * (a0 + j) + 1] = xq[16 + j] / sc;
      }
    }
  }
  free(x);
  free(xq);
  return 0;
}

/*
 * bf_plan.c -- the plan: device buffers, weights upload and the per-block
 * sequence of kernel launches (C, HIP runtime C API).
 *
 * Replaces the reference's allocation block (dsaX_beamformer.cu:880-925),
 * weight refresh (:948-951) and per-int launch loop (:955-1003).  One block
 * is one fused beamformer launch (plus the -s/-q kernel when asked and a
 * per-beam power reduction), instead of 5 launches per 32-sample int.
 */
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../csrc/bf_launch.h"
#include "bf_internal.h"
#include "dsaxbf.h"

struct dsaxbf_plan {
  dsaxbf_config cfg;
  bfk_dims dims;
  hipStream_t stream;
  void *d_wfrag;
  float *d_wscale;
  uint8_t *d_in;        /* staging for dsaxbf_process_block */
  uint8_t *d_out;
  float *d_power;       /* internal power buffer (always available) */
  float *d_beam_power;
  float *d_scale;
  float *d_bpart;       /* per-tile beam power sums of the last execute */
  void *d_scratch;      /* the device's static scratch array (dsaxbf_execute_noalloc) */
  int have_weights;
};

size_t bfk_scratch_bytes(void);
int bfk_scratch_init(int device, void **ptr);

#define HIPCHK(call, what)                                                      \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess)                                                       \
      return dsaxbf_fail(DSAXBF_EHIP, "%s: %s", what, hipGetErrorString(e_));   \
  } while (0)

static void fill_dims(const dsaxbf_config *c, bfk_dims *d) {
  memset(d, 0, sizeof(*d));
  d->nant = c->nant;
  d->nant_pad = c->nant <= 64 ? 64 : c->nant <= 128 ? 128 : 256;   /* 1, 2 or 4 chunks */
  d->nkc = d->nant_pad / 64;
  d->nchan = c->nchan;
  d->ngroup = c->nchan / 8;
  d->npkt = c->npkt;
  d->nsamp = 2 * c->npkt;
  d->nbeam = c->nbeam;
  d->nbeam_pad = (c->nbeam + 255) / 256 * 256;
  d->nbt = d->nbeam_pad / 32;
  d->tavg = c->tavg;
  d->nout = d->nsamp / c->tavg;
  d->ncal = c->weight_mode == DSAXBF_WEIGHTS_GROUP ? d->ngroup : c->nchan;
  d->wmode = c->weight_mode;
  d->amode = c->angle_mode;
  d->prec = c->precision;
  d->order = c->input_order;
  d->stuff = c->stuffants;
  d->layout = c->output_layout;
  /* the autotuner's usual picks on MI355X: shared 256-beam tiles from 129
   * beams (full block: 64 antennas, 256 beams; 256 antennas, 1024 beams),
   * 128- or 64-beam classic tiles below; FP6 runs shared tiles only */
  d->variant = c->nbeam > 128 || c->precision >= DSAXBF_PREC_FP6SPLIT ? BFK_VAR_SHARED_256
               : c->nbeam > 64 ? (d->nkc == 1 ? BFK_VAR_CLASSIC_128S : BFK_VAR_CLASSIC_128)
                               : BFK_VAR_CLASSIC_64;
}

void dsaxbf_destroy(dsaxbf_plan *p) {
  if (!p) return;
  hipSetDevice(p->cfg.device);
  if (p->stream) hipStreamSynchronize(p->stream);
  hipFree(p->d_wfrag);
  hipFree(p->d_wscale);
  hipFree(p->d_in);
  hipFree(p->d_out);
  hipFree(p->d_power);
  hipFree(p->d_beam_power);
  hipFree(p->d_scale);
  hipFree(p->d_bpart);
  if (p->stream) hipStreamDestroy(p->stream);
  free(p);
}

int dsaxbf_create(dsaxbf_plan **out, const dsaxbf_config *cfg) {
  dsaxbf_clear_error();
  if (!out) return dsaxbf_fail(DSAXBF_EINVAL, "plan pointer is NULL");
  *out = NULL;
  int rc = dsaxbf_check_config(cfg);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device || cfg->device < 0)
    return dsaxbf_fail(DSAXBF_ENODEV, "no HIP device %d (found %d)", cfg->device, ndev);
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, cfg->device), "hipGetDeviceProperties");
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return dsaxbf_fail(DSAXBF_ENODEV, "device %d is %s; kernels are built for gfx950", cfg->device,
                       prop.gcnArchName);
  HIPCHK(hipSetDevice(cfg->device), "hipSetDevice");
  dsaxbf_plan *p = calloc(1, sizeof(*p));
  if (!p) return dsaxbf_fail(DSAXBF_ENOMEM, "plan allocation failed");
  p->cfg = *cfg;
  fill_dims(cfg, &p->dims);
  const bfk_dims *d = &p->dims;
  size_t nin = dsaxbf_input_bytes(cfg), nout = dsaxbf_output_bytes(cfg);
  hipError_t e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&p->d_wfrag, bfk_wfrag_bytes(d));
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_wscale, bfk_wscale_bytes(d));
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_in, nin);
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_out, nout);
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_power, nout * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_beam_power, (size_t)cfg->nbeam * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_scale, (size_t)cfg->nbeam * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_bpart, bfk_bpart_bytes(d));
  if (e != hipSuccess) {
    dsaxbf_destroy(p);
    return dsaxbf_fail(DSAXBF_ENOMEM, "device allocation failed: %s", hipGetErrorString(e));
  }
  /* set-up of the allocation-free entry point: the static scratch array's
     address on this device (looked up once per device, device current) */
  rc = bfk_scratch_init(cfg->device, &p->d_scratch);
  if (rc) {
    dsaxbf_destroy(p);
    return dsaxbf_fail(rc, "%s", bfk_last_error());
  }
  *out = p;
  return DSAXBF_OK;
}

int dsaxbf_get_config(const dsaxbf_plan *p, dsaxbf_config *cfg) {
  if (!p || !cfg) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  *cfg = p->cfg;
  return DSAXBF_OK;
}

void *dsaxbf_get_stream(dsaxbf_plan *p) { return p ? (void *)p->stream : NULL; }

int dsaxbf_set_weights(dsaxbf_plan *p, const float *antpos, const float *cal,
                       const double *freqs_hz, const double *theta_rad, double sep_arcmin) {
  dsaxbf_clear_error();
  if (!p || !antpos || !cal || !freqs_hz) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  const bfk_dims *d = &p->dims;
  HIPCHK(hipSetDevice(p->cfg.device), "hipSetDevice");
  size_t ncal_f = (size_t)d->nant * d->ncal * 4;
  double *theta = malloc(sizeof(double) * d->nbeam);
  if (!theta) return dsaxbf_fail(DSAXBF_ENOMEM, "host allocation failed");
  for (int b = 0; b < d->nbeam; b++)
    theta[b] = theta_rad ? theta_rad[b] : sep_arcmin * (double)(d->nbeam / 2 - 1 - b) * M_PI / 10800.0;
  float *d_antpos = NULL, *d_cal = NULL;
  double *d_freqs = NULL, *d_theta = NULL;
  hipError_t e = hipMalloc((void **)&d_antpos, sizeof(float) * d->nant);
  if (e == hipSuccess) e = hipMalloc((void **)&d_cal, sizeof(float) * ncal_f);
  if (e == hipSuccess) e = hipMalloc((void **)&d_freqs, sizeof(double) * d->nchan);
  if (e == hipSuccess) e = hipMalloc((void **)&d_theta, sizeof(double) * d->nbeam);
  if (e == hipSuccess) e = hipMemcpy(d_antpos, antpos, sizeof(float) * d->nant, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_cal, cal, sizeof(float) * ncal_f, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_freqs, freqs_hz, sizeof(double) * d->nchan, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_theta, theta, sizeof(double) * d->nbeam, hipMemcpyHostToDevice);
  free(theta);
  int rc = DSAXBF_OK;
  if (e != hipSuccess) {
    rc = dsaxbf_fail(DSAXBF_EHIP, "weights upload: %s", hipGetErrorString(e));
  } else {
    rc = bfk_weights(d, d_antpos, d_cal, d_freqs, d_theta, p->d_wfrag, p->d_wscale, p->stream);
    if (rc) rc = dsaxbf_fail(rc, "%s", bfk_last_error());
    else {
      e = hipStreamSynchronize(p->stream);
      if (e != hipSuccess) rc = dsaxbf_fail(DSAXBF_EHIP, "weights kernel: %s", hipGetErrorString(e));
    }
  }
  hipFree(d_antpos);
  hipFree(d_cal);
  hipFree(d_freqs);
  hipFree(d_theta);
  if (rc == DSAXBF_OK) p->have_weights = 1;
  return rc;
}

/* ---- the weights as the kernels apply them ----------------------------- */
static double f16_value(uint16_t h) {
  const int e = (h >> 10) & 31, m = h & 1023;
  const double v = e == 0 ? ldexp((double)m, -24) : ldexp(1024.0 + m, e - 25);
  return (h & 0x8000u) ? -v : v;
}
static double e4m3_value(uint8_t c) {
  const int e = (c >> 3) & 15, m = c & 7;
  const double v = e == 0 ? ldexp((double)m, -9) : ldexp(8.0 + m, e - 10);
  return (c & 0x80u) ? -v : v;
}
static double e2m3_value(uint32_t c) {
  const int e = (c >> 3) & 3, m = c & 7;
  const double v = e == 0 ? ldexp((double)m, -3) : ldexp(8.0 + m, e - 4);
  return (c & 0x20u) ? -v : v;
}
/* 6-bit slot j of an FP6 operand's six dwords */
static uint32_t f6_slot(const uint32_t *dw, int j) {
  const int bit = 6 * j, k = bit >> 5, sh = bit & 31;
  uint32_t v = dw[k] >> sh;
  if (sh > 26) v |= dw[k + 1] << (32 - sh);
  return v & 63u;
}

int dsaxbf_get_weights(dsaxbf_plan *p, double *w) {
  dsaxbf_clear_error();
  if (!p || !w) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  if (!p->have_weights) return dsaxbf_fail(DSAXBF_ESTATE, "plan has no weights");
  const bfk_dims *d = &p->dims;
  HIPCHK(hipSetDevice(p->cfg.device), "hipSetDevice");
  const size_t nf = bfk_wfrag_bytes(d), ns = bfk_wscale_bytes(d);
  uint8_t *frag = malloc(nf);
  float *wsc = malloc(ns);
  if (!frag || !wsc) {
    free(frag);
    free(wsc);
    return dsaxbf_fail(DSAXBF_ENOMEM, "host allocation failed");
  }
  hipError_t e = hipMemcpy(frag, p->d_wfrag, nf, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(wsc, p->d_wscale, ns, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    free(frag);
    free(wsc);
    return dsaxbf_fail(DSAXBF_EHIP, "weights download: %s", hipGetErrorString(e));
  }
  /* voltage code unit folded into the power factor: FP16 1, FP8 2^-9, FP6 1/8 */
  const double vunit = d->prec == 0 ? 1.0 : d->prec == 1 ? ldexp(1.0, -9) : 0.125;
  const int nparts = d->prec == 1 || d->prec == 2 ? 2 : 3;
  for (int cq = 0; cq < d->ncal * 2; cq++)
    for (int b = 0; b < d->nbeam; b++) {
      const int bt = b >> 5, n = b & 31;
      const double f = sqrt((double)wsc[(size_t)cq * d->nbeam_pad + b]) * vunit;
      double *row = w + ((size_t)cq * d->nbeam + b) * d->nant * 2;
      for (int a = 0; a < d->nant; a++) {
        const int kc = a >> 6, h = (a >> 5) & 1, lane = h * 32 + n, a32 = a & 31;
        double re, im;
        if (d->prec == 0) {
          const int s = a32 >> 2, i = a32 & 3;   /* slots [r0 r1 i0 i1 r2 r3 i2 i3] */
          const size_t rec = ((((size_t)cq * d->nbt + bt) * d->nkc + kc) * 8 + s) * 64 + lane;
          const uint16_t *v = (const uint16_t *)(frag + rec * 16);
          const int sr = (i >> 1) * 4 + (i & 1);
          re = f16_value(v[sr]);
          im = f16_value(v[sr + 2]);
        } else {
          const int s2 = a32 >> 4, j = a32 & 15;
          const size_t rec = ((((size_t)cq * d->nbt + bt) * d->nkc + kc) * 2 + s2) * nparts * 64 + lane;
          const uint32_t *hi = (const uint32_t *)(frag + rec * 32);
          if (d->prec == 1) {
            const uint32_t *lo = (const uint32_t *)(frag + (rec + 64) * 32);
            const int u = j >> 2, sh = 8 * (j & 3);
            re = e4m3_value((uint8_t)(hi[2 * u] >> sh)) + e4m3_value((uint8_t)(lo[2 * u] >> sh));
            im = e4m3_value((uint8_t)(hi[2 * u + 1] >> sh)) + e4m3_value((uint8_t)(lo[2 * u + 1] >> sh));
          } else {
            /* parts 0..nparts-1 with E8M0 scales: hi dwords 6, 7; mid dword 6 */
            re = im = 0.0;
            for (int pt = 0; pt < nparts; pt++) {
              const uint32_t *dw = (const uint32_t *)(frag + (rec + 64 * pt) * 32);
              const uint32_t sexp = pt < 2 ? hi[6 + pt] : ((const uint32_t *)(frag + (rec + 64) * 32))[6];
              const double s = ldexp(1.0, (int)sexp - 127);
              re += e2m3_value(f6_slot(dw, j)) * s;
              im += e2m3_value(f6_slot(dw, 16 + j)) * s;
            }
          }
        }
        row[2 * a] = re * f;
        row[2 * a + 1] = im * f;
      }
    }
  free(frag);
  free(wsc);
  return DSAXBF_OK;
}

/* One block on device buffers with the per-unit beam-power partial sums in
 * bpart (kernel launches only). */
static int execute_with(dsaxbf_plan *p, float *bpart, const uint8_t *d_in, const float *d_scale,
                        uint8_t *d_out, float *d_power, float *d_beam_power, void *stream) {
  if (!p || !d_in || !d_scale || !d_out) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  const bfk_dims *d = &p->dims;
  if (!p->have_weights && d->stuff != DSAXBF_STUFF_ANTS)
    return dsaxbf_fail(DSAXBF_ESTATE, "execute before dsaxbf_set_weights");
  void *s = stream ? stream : (void *)p->stream;
  int rc;
  if (d->stuff != DSAXBF_STUFF_ANTS) {
    rc = bfk_beamform(d, d_in, p->d_wfrag, p->d_wscale, d_scale, d_out, d_power, bpart, s);
    if (rc) return dsaxbf_fail(rc, "%s", bfk_last_error());
  }
  if (d->stuff != DSAXBF_STUFF_NONE) {
    rc = bfk_stuff(d, d_in, d_scale, d_out, d_power, bpart, s);
    if (rc) return dsaxbf_fail(rc, "%s", bfk_last_error());
  }
  if (d_beam_power) {
    rc = bfk_beam_power(d, bpart, d_beam_power, s);
    if (rc) return dsaxbf_fail(rc, "%s", bfk_last_error());
  }
  return DSAXBF_OK;
}

int dsaxbf_execute(dsaxbf_plan *p, const uint8_t *d_in, const float *d_scale, uint8_t *d_out,
                   float *d_power, float *d_beam_power, void *stream) {
  if (!p) return dsaxbf_fail(DSAXBF_EINVAL, "NULL plan");
  return execute_with(p, p->d_bpart, d_in, d_scale, d_out, d_power, d_beam_power, stream);
}

int dsaxbf_execute_noalloc(dsaxbf_plan *p, const uint8_t *d_in, const float *d_scale, uint8_t *d_out,
                           float *d_power, float *d_beam_power, void *stream) {
  if (!p) return dsaxbf_fail(DSAXBF_EINVAL, "NULL plan");
  const size_t need = bfk_bpart_bytes(&p->dims), have = bfk_scratch_bytes();
  if (need > have)
    return dsaxbf_fail(DSAXBF_EINVAL,
                       "block needs %zu bytes of beam-power scratch, above the %zu-byte static "
                       "maximum (nbeam x nsamp/32 x nchan/8 x 4 bytes); use dsaxbf_execute",
                       need, have);
  /* every partial sum is overwritten by this launch before it is read: the
     array needs no clearing between calls */
  return execute_with(p, (float *)p->d_scratch, d_in, d_scale, d_out, d_power, d_beam_power, stream);
}

int dsaxbf_beam_power(dsaxbf_plan *p, float *d_beam_power, void *stream) {
  if (!p || !d_beam_power) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  int rc = bfk_beam_power(&p->dims, p->d_bpart, d_beam_power, stream ? stream : (void *)p->stream);
  return rc ? dsaxbf_fail(rc, "%s", bfk_last_error()) : DSAXBF_OK;
}

int dsaxbf_process_block(dsaxbf_plan *p, const uint8_t *h_in, const float *h_scale,
                         uint8_t *h_out, float *h_power, float *h_beam_power) {
  dsaxbf_clear_error();
  if (!p || !h_in || !h_scale || !h_out) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  HIPCHK(hipSetDevice(p->cfg.device), "hipSetDevice");
  size_t nin = dsaxbf_input_bytes(&p->cfg), nout = dsaxbf_output_bytes(&p->cfg);
  hipStream_t s = p->stream;
  HIPCHK(hipMemcpyAsync(p->d_in, h_in, nin, hipMemcpyHostToDevice, s), "H2D block");
  HIPCHK(hipMemcpyAsync(p->d_scale, h_scale, sizeof(float) * p->cfg.nbeam, hipMemcpyHostToDevice, s),
         "H2D scale");
  int rc = dsaxbf_execute(p, p->d_in, p->d_scale, p->d_out, h_power ? p->d_power : NULL,
                          h_beam_power ? p->d_beam_power : NULL, s);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(h_out, p->d_out, nout, hipMemcpyDeviceToHost, s), "D2H output");
  if (h_power)
    HIPCHK(hipMemcpyAsync(h_power, p->d_power, nout * sizeof(float), hipMemcpyDeviceToHost, s), "D2H power");
  if (h_beam_power)
    HIPCHK(hipMemcpyAsync(h_beam_power, p->d_beam_power, sizeof(float) * p->cfg.nbeam,
                          hipMemcpyDeviceToHost, s), "D2H beam power");
  HIPCHK(hipStreamSynchronize(s), "block");
  return DSAXBF_OK;
}

/* ---- device memory helpers ------------------------------------------------ */
int dsaxbf_dev_alloc(void **d_ptr, size_t bytes, int device) {
  if (!d_ptr) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  HIPCHK(hipSetDevice(device), "hipSetDevice");
  HIPCHK(hipMalloc(d_ptr, bytes), "hipMalloc");
  return DSAXBF_OK;
}
int dsaxbf_dev_free(void *d_ptr) {
  HIPCHK(hipFree(d_ptr), "hipFree");
  return DSAXBF_OK;
}
int dsaxbf_copy_h2d(void *d_dst, const void *h_src, size_t bytes, void *stream) {
  HIPCHK(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream), "H2D");
  HIPCHK(hipStreamSynchronize((hipStream_t)stream), "H2D sync");
  return DSAXBF_OK;
}
int dsaxbf_copy_d2h(void *h_dst, const void *d_src, size_t bytes, void *stream) {
  HIPCHK(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream), "D2H");
  HIPCHK(hipStreamSynchronize((hipStream_t)stream), "D2H sync");
  return DSAXBF_OK;
}
int dsaxbf_sync(void *stream) {
  HIPCHK(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
  return DSAXBF_OK;
}

int dsaxbf_set_variant(dsaxbf_plan *p, int v) {
  if (!p) return dsaxbf_fail(DSAXBF_EINVAL, "NULL plan");
  if (v < 0 || v >= DSAXBF_NVARIANT) return dsaxbf_fail(DSAXBF_EINVAL, "variant %d out of range", v);
  if (!bfk_variant_supported(&p->dims, v))
    return dsaxbf_fail(DSAXBF_EINVAL, "variant %d cannot run this shape (nant %d)", v, p->cfg.nant);
  p->dims.variant = v;
  return DSAXBF_OK;
}

int dsaxbf_get_variant(const dsaxbf_plan *p) { return p ? p->dims.variant : -1; }

int dsaxbf_plan_internals(dsaxbf_plan *p, bfk_dims *dims, void **d_wfrag, float **d_wscale,
                          float **d_bpart) {
  if (!p) return dsaxbf_fail(DSAXBF_EINVAL, "NULL plan");
  if (!p->have_weights) return dsaxbf_fail(DSAXBF_ESTATE, "plan has no weights");
  if (dims) *dims = p->dims;
  if (d_wfrag) *d_wfrag = p->d_wfrag;
  if (d_wscale) *d_wscale = p->d_wscale;
  if (d_bpart) *d_bpart = p->d_bpart;
  return DSAXBF_OK;
}

int dsaxbf_time_variants(dsaxbf_plan *p, const uint8_t *d_in, const float *d_scale, uint8_t *d_out,
                         int reps, float *ms, int *best_out) {
  if (!p || !d_in || !d_scale || !d_out) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  if (reps < 1) return dsaxbf_fail(DSAXBF_EINVAL, "reps must be >= 1");
  if (!p->have_weights && p->dims.stuff != DSAXBF_STUFF_ANTS)
    return dsaxbf_fail(DSAXBF_ESTATE, "autotune before dsaxbf_set_weights");
  HIPCHK(hipSetDevice(p->cfg.device), "hipSetDevice");
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0), "hipEventCreate");
  hipError_t he = hipEventCreate(&e1);
  if (he != hipSuccess) {
    hipEventDestroy(e0);
    return dsaxbf_fail(DSAXBF_EHIP, "hipEventCreate: %s", hipGetErrorString(he));
  }
  const int keep = p->dims.variant;
  int best = keep, rc = DSAXBF_OK;
  float best_ms = -1.f;
  for (int v = 0; v < DSAXBF_NVARIANT && rc == DSAXBF_OK; v++) {
    if (!bfk_variant_supported(&p->dims, v)) {
      if (ms) ms[v] = -1.f;   /* not a candidate for this shape */
      continue;
    }
    p->dims.variant = v;
    float t = 0.f;
    he = hipSuccess;
    for (int r = -1; r < reps && rc == DSAXBF_OK && he == hipSuccess; r++) {
      if (r == 0) he = hipEventRecord(e0, p->stream);
      if (he == hipSuccess) rc = dsaxbf_execute(p, d_in, d_scale, d_out, NULL, NULL, p->stream);
    }
    if (rc) break;
    if (he == hipSuccess) he = hipEventRecord(e1, p->stream);
    if (he == hipSuccess) he = hipEventSynchronize(e1);
    if (he == hipSuccess) he = hipEventElapsedTime(&t, e0, e1);
    if (he != hipSuccess) {
      rc = dsaxbf_fail(DSAXBF_EHIP, "autotune timing: %s", hipGetErrorString(he));
      break;
    }
    t /= (float)reps;
    if (ms) ms[v] = t;
    if (best_ms < 0.f || t < best_ms) {
      best_ms = t;
      best = v;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  p->dims.variant = keep;
  if (best_out) *best_out = best;
  return rc;
}

int dsaxbf_autotune(dsaxbf_plan *p, int reps, float *ms) {
  dsaxbf_clear_error();
  if (!p) return dsaxbf_fail(DSAXBF_EINVAL, "NULL plan");
  HIPCHK(hipSetDevice(p->cfg.device), "hipSetDevice");
  /* any voltages do: the work does not depend on the values */
  HIPCHK(hipMemsetAsync(p->d_in, 0x5a, dsaxbf_input_bytes(&p->cfg), p->stream), "hipMemsetAsync");
  HIPCHK(hipMemsetAsync(p->d_scale, 0, sizeof(float) * p->cfg.nbeam, p->stream), "hipMemsetAsync");
  int best;
  int rc = dsaxbf_time_variants(p, p->d_in, p->d_scale, p->d_out, reps, ms, &best);
  if (rc == DSAXBF_OK) p->dims.variant = best;
  return rc;
}

/*
 * fl_plan.c -- the flagger's device entry points and the plan that runs one
 * block of the flagger loop (C, HIP runtime C API).
 *
 * Replaces gpu_flagger.cu:1233-1460.  Per block the reference launches
 * scaley, calc_spectrum, calc_varspec, calc_spectrum (its "maxspec"),
 * calc_ppspec, flag, make_ts and flagts -- five full reads of the 1 GiB block
 * and two writes -- with blocking copies of every statistic between them.
 * The plan runs two passes over the block (scaley fused with every
 * statistic; flag fused with make_ts) plus flagts on the masked rows, and
 * copies only the spectra, the masks and the time series.
 */
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../csrc/bf_launch.h"
#include "bf_internal.h"
#include "dsaxfl.h"

int flk_stats(const uint8_t *d_data, uint8_t *d_scaled, const float *d_spec0, const float *d_var0,
              const float *d_spec_in, float *d_spec, float *d_var, float *d_max, float *d_pp, int nb, int nt,
              int nc, void *stream);
int flk_flag_ts(uint8_t *d_data, const uint8_t *d_mask, const float *d_bpwr, float *d_ts, int nb, int nt, int nc,
                void *stream);
int flk_flagts(uint8_t *d_data, const int *d_midx, int n, const float *d_bpwr, int nb, int nt, int nc,
               void *stream);
int flk_flag_cols(uint8_t *d_data, const int *d_midx, int n, const float *d_bpwr, int nb, int nt, int nc,
                  void *stream);
int flk_sumpulse(uint8_t *d_data, const float *d_summand, size_t n, void *stream);
int flk_genrand(uint8_t *d_rep, size_t n, uint32_t seed, void *stream);
int flk_noise(uint8_t *d_data, const uint8_t *d_rep, int nb, int nt, int nc, void *stream);
int flk_fix_zspec(float *d_s0, float *d_v0, int nbc, int naver, void *stream);

/* launcher codes: -1 a shape or argument the kernels do not take, -3 a HIP
   launch error */
#define KCALL(expr)                                                                    \
  do {                                                                                 \
    dsaxbf_clear_error();                                                              \
    int rc_ = (expr);                                                                  \
    if (rc_) return dsaxbf_fail(rc_ == -3 ? DSAXBF_EHIP : DSAXBF_EINVAL, "%s", bfk_last_error()); \
  } while (0)

static int nonnull(const void *p) { return p != NULL; }

/* ---- device entry points ------------------------------------------------ */

int dsaxfl_stats(const uint8_t *d_data, int nb, int nt, int nc, const float *d_spec_in, float *d_spec,
                 float *d_var, float *d_max, float *d_pp, void *stream) {
  if (!nonnull(d_data)) return dsaxbf_fail(DSAXBF_EINVAL, "NULL block");
  KCALL(flk_stats(d_data, NULL, NULL, NULL, d_spec_in, d_spec, d_var, d_max, d_pp, nb, nt, nc, stream));
  return DSAXBF_OK;
}

int dsaxfl_scaley(uint8_t *d_data, int nb, int nt, int nc, const float *d_spec0, const float *d_var0,
                  float *d_spec, float *d_var, float *d_max, float *d_pp, void *stream) {
  if (!d_data || !d_spec0 || !d_var0) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  KCALL(flk_stats(d_data, d_data, d_spec0, d_var0, NULL, d_spec, d_var, d_max, d_pp, nb, nt, nc, stream));
  return DSAXBF_OK;
}

int dsaxfl_make_ts(const uint8_t *d_data, int nb, int nt, int nc, float *d_ts, void *stream) {
  if (!d_data || !d_ts) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  /* no mask: the block is only read */
  KCALL(flk_flag_ts((uint8_t *)d_data, NULL, NULL, d_ts, nb, nt, nc, stream));
  return DSAXBF_OK;
}

int dsaxfl_flag(uint8_t *d_data, int nb, int nt, int nc, const int *d_midx, int n_mask, const float *d_bpwr,
                void *stream) {
  if (!d_data || (n_mask > 0 && (!d_midx || !d_bpwr))) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  if (n_mask < 0 || (long)n_mask > (long)nb * nc) return dsaxbf_fail(DSAXBF_EINVAL, "n_mask=%d", n_mask);
  KCALL(flk_flag_cols(d_data, d_midx, n_mask, d_bpwr, nb, nt, nc, stream));
  return DSAXBF_OK;
}

int dsaxfl_flag_mask(uint8_t *d_data, int nb, int nt, int nc, const uint8_t *d_mask, const float *d_bpwr,
                     float *d_ts, void *stream) {
  if (!d_data || !d_mask || !d_bpwr) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  KCALL(flk_flag_ts(d_data, d_mask, d_bpwr, d_ts, nb, nt, nc, stream));
  return DSAXBF_OK;
}

int dsaxfl_flagts(uint8_t *d_data, int nb, int nt, int nc, const int *d_midx, int n_mask, const float *d_bpwr,
                  void *stream) {
  if (!d_data || (n_mask > 0 && (!d_midx || !d_bpwr))) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  if (n_mask < 0 || (long)n_mask > (long)nb * nt) return dsaxbf_fail(DSAXBF_EINVAL, "n_mask=%d", n_mask);
  KCALL(flk_flagts(d_data, d_midx, n_mask, d_bpwr, nb, nt, nc, stream));
  return DSAXBF_OK;
}

int dsaxfl_sumpulse(uint8_t *d_data, size_t n, const float *d_summand, void *stream) {
  if (!d_data || !d_summand) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  KCALL(flk_sumpulse(d_data, d_summand, n, stream));
  return DSAXBF_OK;
}

int dsaxfl_genrand(uint8_t *d_repval, size_t n, uint32_t seed, void *stream) {
  if (!d_repval) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  KCALL(flk_genrand(d_repval, n, seed, stream));
  return DSAXBF_OK;
}

int dsaxfl_noise(uint8_t *d_data, int nb, int nt, int nc, const uint8_t *d_repval, void *stream) {
  if (!d_data || !d_repval) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  KCALL(flk_noise(d_data, d_repval, nb, nt, nc, stream));
  return DSAXBF_OK;
}

int dsaxfl_fix_zspec(float *d_s0, float *d_v0, int nbc, int naver, void *stream) {
  if (!d_s0 || !d_v0) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  KCALL(flk_fix_zspec(d_s0, d_v0, nbc, naver, stream));
  return DSAXBF_OK;
}

/* ---- plan ------------------------------------------------------------------ */

struct dsaxfl_plan {
  dsaxfl_config cfg;
  hipStream_t stream;
  float *d_spec0, *d_var0, *d_spec, *d_var, *d_pp, *d_ts, *d_bpwr;
  uint8_t *d_mask, *d_repval;
  int *d_tsidx;
  float *h_spec, *h_var, *h_pp, *h_ts, *h_bpwr;
  uint8_t *h_cmask;
  int *h_mask, *h_tsmask, *h_tsidx;
  float prev_tpwr;
  int have_ref;
};

void dsaxfl_config_default(dsaxfl_config *c) {
  if (!c) return;
  memset(c, 0, sizeof(*c));
  c->nbeam = DSAXFL_NBEAMS;
  c->ntime = DSAXFL_NTIMES;
  c->nchan = DSAXFL_NCHAN;
  c->thresh = 5.0f;
  c->tsthresh = 10.0f;
  c->mod_thresh = 0.0005f;
  c->nperm = 1;
  c->perm[0] = 10;
  c->seed = 1;
}

void dsaxfl_destroy(dsaxfl_plan *p) {
  if (!p) return;
  hipSetDevice(p->cfg.device);
  if (p->stream) hipStreamSynchronize(p->stream);
  hipFree(p->d_spec0);
  hipFree(p->d_var0);
  hipFree(p->d_spec);
  hipFree(p->d_var);
  hipFree(p->d_pp);
  hipFree(p->d_ts);
  hipFree(p->d_bpwr);
  hipFree(p->d_mask);
  hipFree(p->d_repval);
  hipFree(p->d_tsidx);
  hipHostFree(p->h_spec);
  hipHostFree(p->h_var);
  hipHostFree(p->h_pp);
  hipHostFree(p->h_ts);
  hipHostFree(p->h_bpwr);
  hipHostFree(p->h_cmask);
  hipHostFree(p->h_tsidx);
  free(p->h_mask);
  free(p->h_tsmask);
  if (p->stream) hipStreamDestroy(p->stream);
  free(p);
}

void *dsaxfl_get_stream(dsaxfl_plan *p) { return p ? (void *)p->stream : NULL; }

int dsaxfl_create(dsaxfl_plan **out, const dsaxfl_config *cfg) {
  dsaxbf_clear_error();
  if (!out || !cfg) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  *out = NULL;
  const int nb = cfg->nbeam, nt = cfg->ntime, nc = cfg->nchan;
  if (nb < 1 || nt < 4096 || nt % 4096 || nc % 16 || nc <= 2 * DSAXFL_ZERO_CHANNELS + 21 ||
      (long)nb * nt * nc > (1L << 34) || (long)nt * nc >= (1L << 31))
    return dsaxbf_fail(DSAXBF_EINVAL,
                       "shape %d x %d x %d: ntime a multiple of 4096, nchan a multiple of 16 above 277", nb, nt,
                       nc);
  if (cfg->nperm < 0 || cfg->nperm > 64) return dsaxbf_fail(DSAXBF_EINVAL, "nperm=%d outside 0..64", cfg->nperm);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device < 0 || cfg->device >= ndev)
    return dsaxbf_fail(DSAXBF_ENODEV, "no HIP device %d (found %d)", cfg->device, ndev);
  if (hipSetDevice(cfg->device) != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "hipSetDevice");
  dsaxfl_plan *p = calloc(1, sizeof(*p));
  if (!p) return dsaxbf_fail(DSAXBF_ENOMEM, "host allocation failed");
  p->cfg = *cfg;
  const size_t nbc = (size_t)nb * nc, nbt = (size_t)nb * nt;
  hipError_t e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_spec0, nbc * 4);
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_var0, nbc * 4);
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_spec, nbc * 4);
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_var, nbc * 4);
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_pp, nbc * 4);
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_ts, nbt * 4);
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_bpwr, (size_t)nb * 4);
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_mask, nbc);
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_repval, (size_t)nt * nc);
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_tsidx, nbt * 4);
  if (e == hipSuccess) e = hipHostMalloc((void **)&p->h_spec, nbc * 4, 0);
  if (e == hipSuccess) e = hipHostMalloc((void **)&p->h_var, nbc * 4, 0);
  if (e == hipSuccess) e = hipHostMalloc((void **)&p->h_pp, nbc * 4, 0);
  if (e == hipSuccess) e = hipHostMalloc((void **)&p->h_ts, nbt * 4, 0);
  if (e == hipSuccess) e = hipHostMalloc((void **)&p->h_bpwr, (size_t)nb * 4, 0);
  if (e == hipSuccess) e = hipHostMalloc((void **)&p->h_cmask, nbc, 0);
  if (e == hipSuccess) e = hipHostMalloc((void **)&p->h_tsidx, nbt * 4, 0);
  p->h_mask = malloc(nbc * sizeof(int));
  p->h_tsmask = malloc(nbt * sizeof(int));
  if (e != hipSuccess || !p->h_mask || !p->h_tsmask) {
    dsaxfl_destroy(p);
    return dsaxbf_fail(DSAXBF_ENOMEM, "flagger allocation failed: %s", hipGetErrorString(e));
  }
  for (int b = 0; b < nb; b++) p->h_bpwr[b] = 1.f;       /* gpu_flagger.cu:1234 */
  e = hipMemcpyAsync(p->d_bpwr, p->h_bpwr, (size_t)nb * 4, hipMemcpyHostToDevice, p->stream);
  int rc = e == hipSuccess ? flk_genrand(p->d_repval, (size_t)nt * nc, cfg->seed, p->stream) : -1;
  if (rc == 0) e = hipStreamSynchronize(p->stream);
  if (e != hipSuccess || rc) {
    dsaxfl_destroy(p);
    return dsaxbf_fail(DSAXBF_EHIP, "flagger set-up: %s", e != hipSuccess ? hipGetErrorString(e) : bfk_last_error());
  }
  *out = p;
  return DSAXBF_OK;
}

int dsaxfl_set_reference(dsaxfl_plan *p, const float *h_spec0, const float *h_var0) {
  if (!p || !h_spec0 || !h_var0) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  const size_t n = (size_t)p->cfg.nbeam * p->cfg.nchan * 4;
  hipError_t e = hipMemcpyAsync(p->d_spec0, h_spec0, n, hipMemcpyHostToDevice, p->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(p->d_var0, h_var0, n, hipMemcpyHostToDevice, p->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
  if (e != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "reference upload: %s", hipGetErrorString(e));
  p->have_ref = 1;
  return DSAXBF_OK;
}

int dsaxfl_get_reference(dsaxfl_plan *p, float *h_spec0, float *h_var0) {
  if (!p || !h_spec0 || !h_var0) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  if (!p->have_ref) return dsaxbf_fail(DSAXBF_ESTATE, "no reference spectra yet");
  const size_t n = (size_t)p->cfg.nbeam * p->cfg.nchan * 4;
  hipError_t e = hipMemcpyAsync(h_spec0, p->d_spec0, n, hipMemcpyDeviceToHost, p->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(h_var0, p->d_var0, n, hipMemcpyDeviceToHost, p->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
  if (e != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "reference download: %s", hipGetErrorString(e));
  return DSAXBF_OK;
}

int dsaxfl_calibrate(dsaxfl_plan *p, const uint8_t *const *d_blocks, int naver) {
  dsaxbf_clear_error();
  if (!p || !d_blocks) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  if (naver < 1 || naver > 16) return dsaxbf_fail(DSAXBF_EINVAL, "naver=%d outside 1..16", naver);
  if (hipSetDevice(p->cfg.device) != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "hipSetDevice");
  const int nb = p->cfg.nbeam, nt = p->cfg.ntime, nc = p->cfg.nchan;
  const size_t nbc = (size_t)nb * nc;
  float *d_s = NULL, *d_v = NULL;
  hipError_t e = hipMalloc((void **)&d_s, nbc * naver * 4);
  if (e == hipSuccess) e = hipMalloc((void **)&d_v, nbc * naver * 4);
  int rc = DSAXBF_OK;
  if (e != hipSuccess) rc = dsaxbf_fail(DSAXBF_ENOMEM, "calibration buffers: %s", hipGetErrorString(e));
  for (int k = 0; k < naver && rc == DSAXBF_OK; k++) {
    if (!d_blocks[k]) rc = dsaxbf_fail(DSAXBF_EINVAL, "block %d is NULL", k);
    else if (flk_stats(d_blocks[k], NULL, NULL, NULL, NULL, d_s + k * nbc, d_v + k * nbc, NULL, NULL, nb, nt, nc,
                       p->stream))
      rc = dsaxbf_fail(DSAXBF_EHIP, "%s", bfk_last_error());
  }
  if (rc == DSAXBF_OK && naver > 1 && flk_fix_zspec(d_s, d_v, (int)nbc, naver, p->stream))
    rc = dsaxbf_fail(DSAXBF_EHIP, "%s", bfk_last_error());
  if (rc == DSAXBF_OK) {
    e = hipMemcpyAsync(p->h_spec, d_s, nbc * 4, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(p->h_var, d_v, nbc * 4, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) rc = dsaxbf_fail(DSAXBF_EHIP, "calibration: %s", hipGetErrorString(e));
  }
  hipFree(d_s);
  hipFree(d_v);
  if (rc) return rc;
  dsaxfl_median_calc(p->h_spec, nb, nc);
  dsaxfl_median_calc(p->h_var, nb, nc);
  rc = dsaxfl_set_reference(p, p->h_spec, p->h_var);
  p->prev_tpwr = 0.f;
  return rc;
}

int dsaxfl_set_total_power(dsaxfl_plan *p, float prev_tpwr) {
  if (!p) return dsaxbf_fail(DSAXBF_EINVAL, "NULL plan");
  p->prev_tpwr = prev_tpwr;
  return DSAXBF_OK;
}

#define HCALL(expr, what)                                                                 \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "%s: %s", what, hipGetErrorString(e_)); \
  } while (0)
#define FCALL(expr)                                                                     \
  do {                                                                                  \
    int rc_ = (expr);                                                                   \
    if (rc_) return dsaxbf_fail(rc_ == -3 ? DSAXBF_EHIP : DSAXBF_EINVAL, "%s", bfk_last_error()); \
  } while (0)

int dsaxfl_step(dsaxfl_plan *p, uint8_t *d_data, dsaxfl_step_info *info) {
  dsaxbf_clear_error();
  if (!p || !d_data) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  if (!p->have_ref) return dsaxbf_fail(DSAXBF_ESTATE, "step before dsaxfl_set_reference / dsaxfl_calibrate");
  HCALL(hipSetDevice(p->cfg.device), "hipSetDevice");
  const dsaxfl_config *c = &p->cfg;
  const int nb = c->nbeam, nt = c->ntime, nc = c->nchan;
  const size_t nbc = (size_t)nb * nc, nbt = (size_t)nb * nt;
  hipStream_t s = p->stream;
  /* scaley + spectrum, varspec, ppspec of the scaled block, one pass */
  FCALL(flk_stats(d_data, d_data, p->d_spec0, p->d_var0, NULL, p->d_spec, p->d_var, NULL, p->d_pp, nb, nt, nc, s));
  HCALL(hipMemcpyAsync(p->h_spec, p->d_spec, nbc * 4, hipMemcpyDeviceToHost, s), "D2H spectrum");
  HCALL(hipMemcpyAsync(p->h_var, p->d_var, nbc * 4, hipMemcpyDeviceToHost, s), "D2H varspec");
  HCALL(hipMemcpyAsync(p->h_pp, p->d_pp, nbc * 4, hipMemcpyDeviceToHost, s), "D2H ppspec");
  HCALL(hipStreamSynchronize(s), "statistics");
  float tpwr = 0.f;
  for (size_t i = 0; i < nbc; i++) tpwr += p->h_spec[i];
  dsaxfl_step_info inf = {tpwr, 0, -1, -1};
  if (fabs(tpwr - p->prev_tpwr) / p->prev_tpwr >= c->mod_thresh) {
    inf.replaced = 1;
    FCALL(flk_noise(d_data, p->d_repval, nb, nt, nc, s));
  } else {
    /* channflag on spectrum - oldspec (oldspec is never updated from zero in
       the reference), on varspec at thresh + 0.5 and on its "maxspec" (a
       second calc_spectrum: the same mask again); simple_channflag on ppspec */
    memset(p->h_mask, 0, nbc * sizeof(int));
    dsaxfl_channflag(p->h_spec, nb, nc, c->thresh, c->perm, c->nperm, p->h_mask);
    dsaxfl_channflag(p->h_var, nb, nc, (float)((double)c->thresh + 0.5), c->perm, c->nperm, p->h_mask);
    dsaxfl_simple_channflag(p->h_pp, nb, nc, c->thresh, p->h_mask);
    if (c->pwr) dsaxfl_calc_bpwr(p->h_spec, nb, nc, p->h_bpwr);
    int nm = 0;
    for (size_t i = 0; i < nbc; i++) {
      p->h_cmask[i] = p->h_mask[i] == 1;
      nm += p->h_mask[i] == 1;
    }
    inf.n_mask = nm;
    HCALL(hipMemcpyAsync(p->d_bpwr, p->h_bpwr, (size_t)nb * 4, hipMemcpyHostToDevice, s), "H2D bpwr");
    float *d_ts = c->tsflag ? p->d_ts : NULL;
    if (c->mkrand) {
      FCALL(flk_noise(d_data, p->d_repval, nb, nt, nc, s));
      if (d_ts) FCALL(flk_flag_ts(d_data, NULL, p->d_bpwr, d_ts, nb, nt, nc, s));
    } else {
      HCALL(hipMemcpyAsync(p->d_mask, p->h_cmask, nbc, hipMemcpyHostToDevice, s), "H2D mask");
      FCALL(flk_flag_ts(d_data, p->d_mask, p->d_bpwr, d_ts, nb, nt, nc, s));
    }
    inf.n_tsmask = 0;
    if (d_ts) {
      HCALL(hipMemcpyAsync(p->h_ts, d_ts, nbt * 4, hipMemcpyDeviceToHost, s), "D2H ts");
      HCALL(hipStreamSynchronize(s), "make_ts");
      memset(p->h_tsmask, 0, nbt * sizeof(int));
      dsaxfl_simple_tsflag(p->h_ts, nb, nt, c->tsthresh, p->h_tsmask);
      const int nts = dsaxfl_gather_mask(p->h_tsmask, (long)nbt, p->h_tsidx);
      inf.n_tsmask = nts;
      if (nts > 0) {
        HCALL(hipMemcpyAsync(p->d_tsidx, p->h_tsidx, (size_t)nts * 4, hipMemcpyHostToDevice, s), "H2D ts mask");
        FCALL(flk_flagts(d_data, p->d_tsidx, nts, p->d_bpwr, nb, nt, nc, s));
      }
    }
  }
  HCALL(hipStreamSynchronize(s), "flagger step");
  p->prev_tpwr = tpwr;
  if (info) *info = inf;
  return DSAXBF_OK;
}

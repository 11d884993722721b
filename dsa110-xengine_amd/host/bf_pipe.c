/*
 * bf_pipe.c -- streaming blocks through a plan with pinned, double-buffered
 * host->device copies (C, HIP runtime C API).
 *
 * The reference loops over 128 "ints" per block with pageable cudaMemcpyAsync
 * on 4 streams (dsaX_beamformer.cu:955-1003), so its copies serialise with the
 * kernels.  Here a block's upload runs on a copy stream into one of two device
 * slots while the previous block is beamformed on the plan's stream; the AGC
 * dependency (block k's beam power sets block k+1's scale, :1005-1066) stays on
 * the host between runs.
 *
 * Use:
 *   dsaxbf_pipe_create(&pp, plan);
 *   fill dsaxbf_pipe_input(pp, 0); dsaxbf_pipe_upload(pp, 0);
 *   for k: fill input(pp, (k+1)%2); upload(pp, (k+1)%2);      <- overlaps run k
 *          run(pp, k%2, scale, &out, &bp); agc(bp) -> scale;
 */
#include <hip/hip_runtime_api.h>
#include <stdlib.h>
#include <string.h>

#include "bf_internal.h"
#include "dsaxbf.h"

#define NSLOT 2

struct dsaxbf_pipe {
  dsaxbf_plan *plan;
  int device, nbeam;
  size_t nin, nout;
  hipStream_t copy;
  struct {
    uint8_t *h_in, *d_in, *h_out, *d_out;
    float *h_bp, *d_bp, *d_scale;
    hipEvent_t uploaded, done;
    int uploading;
  } s[NSLOT];
};

void dsaxbf_pipe_destroy(dsaxbf_pipe *pp) {
  if (!pp) return;
  hipSetDevice(pp->device);
  /* the pipe's own copies and each slot's last run (the plan is not touched:
     it may already be gone) */
  if (pp->copy) hipStreamSynchronize(pp->copy);
  for (int i = 0; i < NSLOT; i++)
    if (pp->s[i].done) hipEventSynchronize(pp->s[i].done);
  for (int i = 0; i < NSLOT; i++) {
    hipHostFree(pp->s[i].h_in);
    hipHostFree(pp->s[i].h_out);
    hipHostFree(pp->s[i].h_bp);
    hipFree(pp->s[i].d_in);
    hipFree(pp->s[i].d_out);
    hipFree(pp->s[i].d_bp);
    hipFree(pp->s[i].d_scale);
    if (pp->s[i].uploaded) hipEventDestroy(pp->s[i].uploaded);
    if (pp->s[i].done) hipEventDestroy(pp->s[i].done);
  }
  if (pp->copy) hipStreamDestroy(pp->copy);
  free(pp);
}

int dsaxbf_pipe_create(dsaxbf_pipe **out, dsaxbf_plan *plan) {
  dsaxbf_clear_error();
  if (!out || !plan) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  *out = NULL;
  dsaxbf_config cfg;
  int rc = dsaxbf_get_config(plan, &cfg);
  if (rc) return rc;
  if (hipSetDevice(cfg.device) != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "hipSetDevice");
  dsaxbf_pipe *pp = calloc(1, sizeof(*pp));
  if (!pp) return dsaxbf_fail(DSAXBF_ENOMEM, "pipe allocation failed");
  pp->plan = plan;
  pp->device = cfg.device;
  pp->nbeam = cfg.nbeam;
  pp->nin = dsaxbf_input_bytes(&cfg);
  pp->nout = dsaxbf_output_bytes(&cfg);
  hipError_t e = hipStreamCreateWithFlags(&pp->copy, hipStreamNonBlocking);
  for (int i = 0; i < NSLOT && e == hipSuccess; i++) {
    e = hipHostMalloc((void **)&pp->s[i].h_in, pp->nin, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&pp->s[i].h_out, pp->nout, hipHostMallocDefault);
    if (e == hipSuccess)
      e = hipHostMalloc((void **)&pp->s[i].h_bp, sizeof(float) * pp->nbeam, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc((void **)&pp->s[i].d_in, pp->nin);
    if (e == hipSuccess) e = hipMalloc((void **)&pp->s[i].d_out, pp->nout);
    if (e == hipSuccess) e = hipMalloc((void **)&pp->s[i].d_bp, sizeof(float) * pp->nbeam);
    if (e == hipSuccess) e = hipMalloc((void **)&pp->s[i].d_scale, sizeof(float) * pp->nbeam);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&pp->s[i].uploaded, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&pp->s[i].done, hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    dsaxbf_pipe_destroy(pp);
    return dsaxbf_fail(DSAXBF_ENOMEM, "pipe buffers: %s", hipGetErrorString(e));
  }
  *out = pp;
  return DSAXBF_OK;
}

uint8_t *dsaxbf_pipe_input(dsaxbf_pipe *pp, int slot) {
  return (pp && slot >= 0 && slot < NSLOT) ? pp->s[slot].h_in : NULL;
}

int dsaxbf_pipe_upload(dsaxbf_pipe *pp, int slot) {
  if (!pp || slot < 0 || slot >= NSLOT) return dsaxbf_fail(DSAXBF_EINVAL, "bad pipe slot");
  if (hipSetDevice(pp->device) != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "hipSetDevice");
  /* the slot's device input is free once its previous block has been run */
  hipError_t e = hipStreamWaitEvent(pp->copy, pp->s[slot].done, 0);
  if (e == hipSuccess)
    e = hipMemcpyAsync(pp->s[slot].d_in, pp->s[slot].h_in, pp->nin, hipMemcpyHostToDevice, pp->copy);
  if (e == hipSuccess) e = hipEventRecord(pp->s[slot].uploaded, pp->copy);
  if (e != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "pipe upload: %s", hipGetErrorString(e));
  pp->s[slot].uploading = 1;
  return DSAXBF_OK;
}

int dsaxbf_pipe_run(dsaxbf_pipe *pp, int slot, const float *h_scale, const uint8_t **h_out,
                    const float **h_beam_power) {
  if (!pp || slot < 0 || slot >= NSLOT || !h_scale)
    return dsaxbf_fail(DSAXBF_EINVAL, "bad pipe run arguments");
  if (!pp->s[slot].uploading) return dsaxbf_fail(DSAXBF_ESTATE, "pipe slot %d was not uploaded", slot);
  if (hipSetDevice(pp->device) != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "hipSetDevice");
  hipStream_t s = (hipStream_t)dsaxbf_get_stream(pp->plan);
  hipError_t e = hipStreamWaitEvent(s, pp->s[slot].uploaded, 0);
  if (e == hipSuccess)
    e = hipMemcpyAsync(pp->s[slot].d_scale, h_scale, sizeof(float) * pp->nbeam, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "pipe run: %s", hipGetErrorString(e));
  int rc = dsaxbf_execute(pp->plan, pp->s[slot].d_in, pp->s[slot].d_scale, pp->s[slot].d_out, NULL,
                          pp->s[slot].d_bp, s);
  if (rc) return rc;
  e = hipMemcpyAsync(pp->s[slot].h_out, pp->s[slot].d_out, pp->nout, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(pp->s[slot].h_bp, pp->s[slot].d_bp, sizeof(float) * pp->nbeam,
                       hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipEventRecord(pp->s[slot].done, s);
  if (e == hipSuccess) e = hipEventSynchronize(pp->s[slot].done);
  if (e != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "pipe run: %s", hipGetErrorString(e));
  pp->s[slot].uploading = 0;
  if (h_out) *h_out = pp->s[slot].h_out;
  if (h_beam_power) *h_beam_power = pp->s[slot].h_bp;
  return DSAXBF_OK;
}

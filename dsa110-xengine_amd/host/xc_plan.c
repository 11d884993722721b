/*
 * xc_plan.c -- the correlator X-engine plan (C, HIP runtime C API).
 *
 * Replaces the reference's per-block correlator call sequence
 * (dsaX_bfCorr.cu:472-569: copy, reorder, four batched GEMMs, two reorder
 * kernels, with cudaMalloc/cublasCreate per call) by a turn of the packed
 * block into (channel, pol) planes, one correlation launch over time slices
 * and the sum of the slices, on device buffers the plan owns.
 */
#include <hip/hip_runtime_api.h>
#include <stdlib.h>
#include <string.h>

#include "../csrc/bf_launch.h"
#include "bf_internal.h"
#include "dsaxbf.h"

size_t bfk_xcorr_vis_bytes(int nant, int nchan);
size_t bfk_xcorr_part_bytes(int nant, int nchan, int nslice);
int bfk_xcorr_default_slices(int nchan, int npkt);
int bfk_xcorr(int nant, int nchan, int npkt, int order, const uint8_t *d_in, uint8_t *d_turn, float *d_part,
              int nslice, float *d_vis, void *stream);

struct dsaxbf_xengine {
  int nant, nchan, npkt, order, device, nslice;
  hipStream_t stream;
  uint8_t *d_in, *d_turn;
  float *d_vis, *d_part;
};

size_t dsaxbf_xengine_vis_bytes(int nant, int nchan) {
  return nant < 1 || nchan < 1 ? 0 : bfk_xcorr_vis_bytes(nant, nchan);
}

void dsaxbf_xengine_destroy(dsaxbf_xengine *x) {
  if (!x) return;
  hipSetDevice(x->device);
  if (x->stream) hipStreamSynchronize(x->stream);
  hipFree(x->d_in);
  hipFree(x->d_turn);
  hipFree(x->d_vis);
  hipFree(x->d_part);
  if (x->stream) hipStreamDestroy(x->stream);
  free(x);
}

int dsaxbf_xengine_create(dsaxbf_xengine **out, int nant, int nchan, int npkt, int input_order,
                          int device) {
  dsaxbf_clear_error();
  if (!out) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  *out = NULL;
  if (nant < 1 || nant > 96) return dsaxbf_fail(DSAXBF_EINVAL, "nant=%d outside 1..96", nant);
  if (nchan < 8 || nchan % 8) return dsaxbf_fail(DSAXBF_EINVAL, "nchan=%d must be a positive multiple of 8", nchan);
  if (npkt < 32 || npkt % 32) return dsaxbf_fail(DSAXBF_EINVAL, "npkt=%d must be a positive multiple of 32", npkt);
  if (input_order != 0 && input_order != 1)
    return dsaxbf_fail(DSAXBF_EINVAL, "input_order=%d unknown", input_order);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return dsaxbf_fail(DSAXBF_ENODEV, "no HIP device %d (found %d)", device, ndev);
  if (hipSetDevice(device) != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "hipSetDevice");
  dsaxbf_xengine *x = calloc(1, sizeof(*x));
  if (!x) return dsaxbf_fail(DSAXBF_ENOMEM, "host allocation failed");
  x->nant = nant;
  x->nchan = nchan;
  x->npkt = npkt;
  x->order = input_order;
  x->device = device;
  x->nslice = bfk_xcorr_default_slices(nchan, npkt);
  const size_t nin = (size_t)npkt * nant * nchan * 4, npart = bfk_xcorr_part_bytes(nant, nchan, x->nslice);
  hipError_t e = hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void **)&x->d_in, nin);
  if (e == hipSuccess) e = hipMalloc((void **)&x->d_turn, nin);
  if (e == hipSuccess) e = hipMalloc((void **)&x->d_vis, bfk_xcorr_vis_bytes(nant, nchan));
  if (e == hipSuccess && npart) e = hipMalloc((void **)&x->d_part, npart);
  if (e != hipSuccess) {
    dsaxbf_xengine_destroy(x);
    return dsaxbf_fail(DSAXBF_ENOMEM, "device allocation failed: %s", hipGetErrorString(e));
  }
  *out = x;
  return DSAXBF_OK;
}

void *dsaxbf_xengine_stream(dsaxbf_xengine *x) { return x ? (void *)x->stream : NULL; }

int dsaxbf_xengine_execute(dsaxbf_xengine *x, const uint8_t *d_in, float *d_vis, void *stream) {
  if (!x || !d_in || !d_vis) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  int rc = bfk_xcorr(x->nant, x->nchan, x->npkt, x->order, d_in, x->d_turn, x->d_part, x->nslice, d_vis,
                     stream ? stream : (void *)x->stream);
  return rc ? dsaxbf_fail(rc, "%s", bfk_last_error()) : DSAXBF_OK;
}

int dsaxbf_xengine_set_slices(dsaxbf_xengine *x, int nslice) {
  dsaxbf_clear_error();
  if (!x) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  if (nslice == 0) nslice = bfk_xcorr_default_slices(x->nchan, x->npkt);
  if (nslice < 1 || nslice > x->npkt / 32)
    return dsaxbf_fail(DSAXBF_EINVAL, "nslice=%d outside 1..%d", nslice, x->npkt / 32);
  if (hipSetDevice(x->device) != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "hipSetDevice");
  hipError_t e = hipStreamSynchronize(x->stream);
  if (e != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "correlator stream: %s", hipGetErrorString(e));
  float *part = NULL;
  const size_t npart = bfk_xcorr_part_bytes(x->nant, x->nchan, nslice);
  if (npart && hipMalloc((void **)&part, npart) != hipSuccess)
    return dsaxbf_fail(DSAXBF_ENOMEM, "partial visibilities (%zu bytes)", npart);
  hipFree(x->d_part);
  x->d_part = part;
  x->nslice = nslice;
  return DSAXBF_OK;
}

int dsaxbf_xengine_get_slices(const dsaxbf_xengine *x) { return x ? x->nslice : -1; }

int dsaxbf_xengine_process_block(dsaxbf_xengine *x, const uint8_t *h_in, float *h_vis) {
  dsaxbf_clear_error();
  if (!x || !h_in || !h_vis) return dsaxbf_fail(DSAXBF_EINVAL, "NULL argument");
  if (hipSetDevice(x->device) != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "hipSetDevice");
  const size_t nin = (size_t)x->npkt * x->nant * x->nchan * 4, nv = bfk_xcorr_vis_bytes(x->nant, x->nchan);
  hipError_t e = hipMemcpyAsync(x->d_in, h_in, nin, hipMemcpyHostToDevice, x->stream);
  if (e != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "H2D block: %s", hipGetErrorString(e));
  int rc = dsaxbf_xengine_execute(x, x->d_in, x->d_vis, x->stream);
  if (rc) return rc;
  e = hipMemcpyAsync(h_vis, x->d_vis, nv, hipMemcpyDeviceToHost, x->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
  if (e != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "correlator: %s", hipGetErrorString(e));
  return DSAXBF_OK;
}

/*
 * bf_batch.c -- many blocks of different shapes beamformed by one HIP graph
 * launch (C, HIP runtime C API).
 *
 * The reference beamforms one compiled-in block shape per process
 * (dsaX_beamformer.cu:939-1090).  A batch binds n plans (each its own shape,
 * weights and device buffers) and replays, per launch, the beamforming
 * launches of bf_jobs.hip (one launch per precision and tavg for
 * the 64-antenna FP16 / FP8 blocks, whatever their tile shapes) and one launch
 * for every block's beam-power sums, all captured once in a graph.  The
 * autotuner times each block shape's variants on that block's own buffers,
 * then lets the whole graph choose between those picks, the plans' previous
 * variants and the shared schedule everywhere.
 */
#include <hip/hip_runtime_api.h>
#include <stdlib.h>
#include <string.h>

#include "../csrc/bf_launch.h"
#include "bf_internal.h"
#include "dsaxbf.h"

struct dsaxbf_batch {
  int n, device;
  dsaxbf_plan **plans;
  bfk_dims *dims;
  bfk_batch_bufs *bufs;
  float **d_bp;             /* per block, NULL = no beam-power sum */
  void *kern;               /* bfk batch tables */
  hipStream_t stream;
  hipGraph_t graph;
  hipGraphExec_t exec;
  int nkernels;
};

static void release_graph(dsaxbf_batch *b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = NULL;
  b->graph = NULL;
  bfk_batch_free(b->kern);
  b->kern = NULL;
}

/* (Re)build the job tables from the plans' current variants and capture the
 * launch sequence. */
static int build(dsaxbf_batch *b) {
  release_graph(b);
  for (int i = 0; i < b->n; i++) {
    int rc = dsaxbf_plan_internals(b->plans[i], &b->dims[i], (void **)&b->bufs[i].d_wfrag,
                                   (float **)&b->bufs[i].d_wscale, &b->bufs[i].d_bpart);
    if (rc) return rc;
    if (b->dims[i].stuff != DSAXBF_STUFF_NONE)
      return dsaxbf_fail(DSAXBF_EINVAL, "block %d: -s/-q modes are not batched", i);
    b->bufs[i].d_beam_power = b->d_bp[i];
  }
  if (bfk_batch_prepare(b->dims, b->bufs, b->n, &b->kern))
    return dsaxbf_fail(DSAXBF_EINVAL, "%s", bfk_last_error());
  hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "begin capture: %s", hipGetErrorString(e));
  int rc = bfk_batch_launch(b->kern, b->stream);
  int nk = bfk_batch_launches(b->kern);
  hipGraph_t g = NULL;
  e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return dsaxbf_fail(DSAXBF_EHIP, "batch capture: %s", bfk_last_error());
  }
  if (e != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "end capture: %s", hipGetErrorString(e));
  b->graph = g;
  e = hipGraphInstantiate(&b->exec, g, NULL, NULL, 0);
  if (e != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "graph instantiate: %s", hipGetErrorString(e));
  b->nkernels = nk;
  return DSAXBF_OK;
}

void dsaxbf_batch_destroy(dsaxbf_batch *b) {
  if (!b) return;
  hipSetDevice(b->device);
  if (b->stream) hipStreamSynchronize(b->stream);
  release_graph(b);
  if (b->stream) hipStreamDestroy(b->stream);
  free(b->plans);
  free(b->dims);
  free(b->bufs);
  free(b->d_bp);
  free(b);
}

int dsaxbf_batch_create(dsaxbf_batch **out, dsaxbf_plan *const *plans, int n,
                        const uint8_t *const *d_in, const float *const *d_scale,
                        uint8_t *const *d_out, float *const *d_beam_power) {
  dsaxbf_clear_error();
  if (!out || !plans || n < 1 || !d_in || !d_scale || !d_out)
    return dsaxbf_fail(DSAXBF_EINVAL, "bad batch arguments");
  *out = NULL;
  dsaxbf_config c0;
  int rc = dsaxbf_get_config(plans[0], &c0);
  if (rc) return rc;
  for (int i = 0; i < n; i++) {
    dsaxbf_config c;
    if (!plans[i] || !d_in[i] || !d_scale[i] || !d_out[i])
      return dsaxbf_fail(DSAXBF_EINVAL, "block %d: NULL plan or buffer", i);
    dsaxbf_get_config(plans[i], &c);
    if (c.device != c0.device) return dsaxbf_fail(DSAXBF_EINVAL, "block %d is on another device", i);
    /* a plan owns the per-unit beam-power scratch its block writes, so two
       blocks of one launch cannot share a plan */
    for (int k = 0; k < i; k++)
      if (plans[k] == plans[i])
        return dsaxbf_fail(DSAXBF_EINVAL, "block %d reuses the plan of block %d", i, k);
  }
  if (hipSetDevice(c0.device) != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "hipSetDevice");
  dsaxbf_batch *b = calloc(1, sizeof(*b));
  if (!b) return dsaxbf_fail(DSAXBF_ENOMEM, "batch allocation failed");
  b->n = n;
  b->device = c0.device;
  b->plans = calloc(n, sizeof(*b->plans));
  b->dims = calloc(n, sizeof(*b->dims));
  b->bufs = calloc(n, sizeof(*b->bufs));
  b->d_bp = calloc(n, sizeof(*b->d_bp));
  if (!b->plans || !b->dims || !b->bufs || !b->d_bp) {
    dsaxbf_batch_destroy(b);
    return dsaxbf_fail(DSAXBF_ENOMEM, "batch allocation failed");
  }
  for (int i = 0; i < n; i++) {
    b->plans[i] = plans[i];
    b->bufs[i].d_in = d_in[i];
    b->bufs[i].d_scale = d_scale[i];
    b->bufs[i].d_out = d_out[i];
    b->bufs[i].d_power = NULL;
    b->d_bp[i] = d_beam_power ? d_beam_power[i] : NULL;
  }
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    dsaxbf_batch_destroy(b);
    return dsaxbf_fail(DSAXBF_EHIP, "batch stream: %s", hipGetErrorString(e));
  }
  rc = build(b);
  if (rc) {
    dsaxbf_batch_destroy(b);
    return rc;
  }
  *out = b;
  return DSAXBF_OK;
}

int dsaxbf_batch_launch(dsaxbf_batch *b, void *stream) {
  if (!b || !b->exec) return dsaxbf_fail(DSAXBF_EINVAL, "NULL batch");
  if (hipSetDevice(b->device) != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "hipSetDevice");
  hipError_t e = hipGraphLaunch(b->exec, stream ? (hipStream_t)stream : b->stream);
  if (e != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "graph launch: %s", hipGetErrorString(e));
  if (!stream) {
    e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "batch: %s", hipGetErrorString(e));
  }
  return DSAXBF_OK;
}

int dsaxbf_batch_kernels(const dsaxbf_batch *b) { return b ? b->nkernels : -1; }

/* Same shape for tuning purposes: identical work per tile and tile count. */
static int same_shape(const bfk_dims *x, const bfk_dims *y) {
  return x->nant_pad == y->nant_pad && x->nchan == y->nchan && x->npkt == y->npkt &&
         x->nbeam == y->nbeam && x->tavg == y->tavg && x->prec == y->prec && x->wmode == y->wmode &&
         x->order == y->order;
}

/* Mean ms of one replay of the current graph over reps replays (after one). */
static int time_graph(dsaxbf_batch *b, int reps, float *out_ms) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "hipEventCreate");
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return dsaxbf_fail(DSAXBF_EHIP, "hipEventCreate");
  }
  hipError_t e = hipGraphLaunch(b->exec, b->stream);
  if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
  for (int r = 0; r < reps && e == hipSuccess; r++) e = hipGraphLaunch(b->exec, b->stream);
  if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (e != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "batch timing: %s", hipGetErrorString(e));
  *out_ms = t / (float)reps;
  return DSAXBF_OK;
}

/* Set the plans' variants to v[] and capture the graph again. */
static int apply(dsaxbf_batch *b, const int *v) {
  for (int i = 0; i < b->n; i++) {
    int rc = dsaxbf_set_variant(b->plans[i], v[i]);
    if (rc) return rc;
  }
  return build(b);
}

int dsaxbf_batch_autotune(dsaxbf_batch *b, int reps, float *ms) {
  dsaxbf_clear_error();
  if (!b) return dsaxbf_fail(DSAXBF_EINVAL, "NULL batch");
  if (reps < 1) return dsaxbf_fail(DSAXBF_EINVAL, "reps must be >= 1");
  if (hipSetDevice(b->device) != hipSuccess) return dsaxbf_fail(DSAXBF_EHIP, "hipSetDevice");
  if (ms)
    for (int i = 0; i < b->n * DSAXBF_NVARIANT; i++) ms[i] = -1.f;
  /* candidate assignments: [0] the variants the plans had, [1] each shape's
     fastest alone, [2] the shared schedule everywhere */
  int *cand = malloc(sizeof(int) * 3 * b->n);
  if (!cand) return dsaxbf_fail(DSAXBF_ENOMEM, "host allocation failed");
  int rc = DSAXBF_OK;
  for (int i = 0; i < b->n && rc == DSAXBF_OK; i++) {
    cand[i] = dsaxbf_get_variant(b->plans[i]);
    cand[2 * b->n + i] = DSAXBF_VARIANT_SHARED_256;
    int *best = &cand[b->n + i];
    *best = -1;
    for (int k = 0; k < i; k++)
      if (same_shape(&b->dims[k], &b->dims[i])) {
        *best = cand[b->n + k];
        break;
      }
    if (*best >= 0) continue;
    rc = dsaxbf_time_variants(b->plans[i], b->bufs[i].d_in, b->bufs[i].d_scale, b->bufs[i].d_out, reps,
                              ms ? ms + (size_t)i * DSAXBF_NVARIANT : NULL, best);
  }
  /* the whole graph decides between the candidates */
  int pick = 0;
  float best_ms = -1.f;
  for (int c = 0; c < 3 && rc == DSAXBF_OK; c++) {
    float t = 0.f;
    rc = apply(b, cand + c * b->n);
    if (rc == DSAXBF_OK) rc = time_graph(b, reps, &t);
    if (rc == DSAXBF_OK && (best_ms < 0.f || t < best_ms)) {
      best_ms = t;
      pick = c;
    }
  }
  if (rc == DSAXBF_OK) rc = apply(b, cand + pick * b->n);
  free(cand);
  return rc;
}

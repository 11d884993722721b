// bf_jobs.hip -- a batch of blocks in as few launches as their kernels allow.
//
// The reference beamforms one fixed-shape block at a time (five launches per
// 32-sample int, dsaX_beamformer.cu:955-1003).  A batch here holds blocks of
// different channel, beam and time counts, each with its own weights, buffers
// and tile shape (variant); every job keeps its own BfArgs, so shapes mix
// freely in a launch whose workgroups walk a device table of jobs.
//   - 64-antenna FP16 and FP8-split blocks: one launch per
//     (precision, tavg) for all of them, classic and shared tiles together
//     (bf_mixed.hip);
//   - FP6 blocks (shared tiles only) and blocks of more than 64 antennas: one
//     launch per kernel instantiation (schedule, precision, tavg, antenna
//     chunks, beam tiles per wave, sample split);
//   - then one launch for the per-beam power sums of every block that asks.
// bfk_batch_launch allocates nothing and never synchronises with the host, so
// a caller can capture it in a HIP graph.
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "bf_common.h"

extern "C" int bfk_classic_args(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag,
                                const float *d_wscale, const float *d_scale, uint8_t *d_out,
                                float *d_power, float *d_bpart, BfArgs *a, int *nb);
extern "C" int bfk_beamform_classic_batch(int prec, int tavg, int nkc, int nb, int split,
                                          const void *d_jobs, const int *d_tile_start, int njobs,
                                          int ntiles, void *stream);
extern "C" int bfk_shared_args(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag,
                               const float *d_wscale, const float *d_scale, uint8_t *d_out,
                               float *d_power, float *d_bpart, BfArgs *a);
extern "C" int bfk_beamform_shared_batch(int prec, int tavg, const void *d_jobs, const int *d_tile_start,
                                         int njobs, int ntiles, void *stream);
extern "C" int bfk_shared8_args(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag,
                                const float *d_wscale, const float *d_scale, uint8_t *d_out,
                                float *d_power, float *d_bpart, BfArgs *a);
extern "C" int bfk_beamform_shared8_batch(int prec, int tavg, int nkc, const void *d_jobs,
                                          const int *d_tile_start, int njobs, int ntiles, void *stream);
extern "C" int bfk_mix_kind(int variant);
extern "C" int bfk_beamform_mixed_batch(int prec, int tavg, const void *d_jobs, const int *d_tile_start,
                                        int ntiles, void *stream);
extern "C" int bfk_beam_power_batch(const void *d_jobs, const int *d_beam_start, int nbeams_total,
                                    void *stream);

namespace {
struct Group {
  int mixed, shared, prec, tavg, nkc, nb, split;
  int njobs, ntiles;
  BfArgs *d_jobs;
  int *d_tile_start;     // njobs + 1 prefix sums
};
struct Batch {
  std::vector<Group> groups;
  bfk_bp_job *d_bp = nullptr;   // beam-power jobs, and their beam prefix sums
  int *d_beam_start = nullptr;
  int nbeams = 0;
};
}

extern "C" void bfk_batch_free(void *handle) {
  Batch *b = (Batch *)handle;
  if (!b) return;
  for (Group &g : b->groups) {
    (void)hipFree(g.d_jobs);
    (void)hipFree(g.d_tile_start);
  }
  (void)hipFree(b->d_bp);
  (void)hipFree(b->d_beam_start);
  delete b;
}

extern "C" int bfk_batch_prepare(const bfk_dims *dims, const bfk_batch_bufs *bufs, int n, void **handle) {
  *handle = nullptr;
  if (n < 1) {
    bfk_set_error("bfk_batch_prepare: empty batch");
    return -1;
  }
  struct Job {
    int mixed, shared, prec, tavg, nkc, nb;
    BfArgs a;
  };
  std::vector<Job> jobs(n);
  for (int i = 0; i < n; i++) {
    const bfk_dims *d = &dims[i];
    // buffer offsets are 32-bit, as for single blocks
    if ((size_t)d->npkt * d->nant * d->nchan * 4 > 0x7fffffffu || bfk_wfrag_bytes(d) > 0x7fffffffu) {
      bfk_set_error("bfk_batch_prepare: block %d or its weights above 2 GiB", i);
      return -1;
    }
    if (!bfk_variant_supported(d, d->variant)) {
      bfk_set_error("bfk_batch_prepare: block %d: variant %d cannot run its shape", i, d->variant);
      return -1;
    }
    Job &j = jobs[i];
    j.shared = d->variant == BFK_VAR_SHARED_256;
    j.nb = 0;
    const bfk_batch_bufs &u = bufs[i];
    const int rc = !j.shared ? bfk_classic_args(d, u.d_in, u.d_wfrag, u.d_wscale, u.d_scale, u.d_out,
                                                u.d_power, u.d_bpart, &j.a, &j.nb)
                   : d->nkc == 1 ? bfk_shared_args(d, u.d_in, u.d_wfrag, u.d_wscale, u.d_scale, u.d_out,
                                                   u.d_power, u.d_bpart, &j.a)
                                 : bfk_shared8_args(d, u.d_in, u.d_wfrag, u.d_wscale, u.d_scale, u.d_out,
                                                    u.d_power, u.d_bpart, &j.a);
    if (rc) return -1;
    j.prec = d->prec;
    j.tavg = d->tavg;
    j.nkc = d->nkc;
    j.mixed = d->nkc == 1 && d->prec <= 1;      // classic tiles exist: one launch for all shapes
    j.a.kind = bfk_mix_kind(d->variant);
  }
  Batch *b = new Batch;
  std::vector<char> done(n, 0);
  for (int i = 0; i < n; i++) {
    if (done[i]) continue;
    Group g{jobs[i].mixed, jobs[i].shared, jobs[i].prec, jobs[i].tavg, jobs[i].nkc, jobs[i].nb,
            jobs[i].a.split, 0, 0, nullptr, nullptr};
    std::vector<BfArgs> tab;
    std::vector<int> start{0};
    for (int k = i; k < n; k++) {
      const Job &j = jobs[k];
      if (done[k] || j.mixed != g.mixed || j.prec != g.prec || j.tavg != g.tavg) continue;
      if (!g.mixed && (j.shared != g.shared || j.nkc != g.nkc || j.nb != g.nb || j.a.split != g.split))
        continue;
      done[k] = 1;
      tab.push_back(j.a);
      start.push_back(start.back() + j.a.ntiles);
    }
    g.njobs = (int)tab.size();
    g.ntiles = start.back();
    hipError_t e = hipMalloc((void **)&g.d_jobs, sizeof(BfArgs) * tab.size());
    if (e == hipSuccess) e = hipMalloc((void **)&g.d_tile_start, sizeof(int) * start.size());
    if (e == hipSuccess)
      e = hipMemcpy(g.d_jobs, tab.data(), sizeof(BfArgs) * tab.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(g.d_tile_start, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice);
    b->groups.push_back(g);
    if (e != hipSuccess) {
      bfk_batch_free(b);
      return bfk_check(e, "bfk_batch_prepare");
    }
  }
  // the beam-power sums of the blocks that ask for them, in one table
  std::vector<bfk_bp_job> bp;
  std::vector<int> bstart{0};
  for (int i = 0; i < n; i++)
    if (bufs[i].d_beam_power) {
      bp.push_back(bfk_bp_job{bufs[i].d_bpart, bufs[i].d_beam_power, dims[i].nbeam,
                              (dims[i].nsamp / 32) * dims[i].ngroup});
      bstart.push_back(bstart.back() + dims[i].nbeam);
    }
  if (!bp.empty()) {
    hipError_t e = hipMalloc((void **)&b->d_bp, sizeof(bfk_bp_job) * bp.size());
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_beam_start, sizeof(int) * bstart.size());
    if (e == hipSuccess) e = hipMemcpy(b->d_bp, bp.data(), sizeof(bfk_bp_job) * bp.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(b->d_beam_start, bstart.data(), sizeof(int) * bstart.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      bfk_batch_free(b);
      return bfk_check(e, "bfk_batch_prepare");
    }
    b->nbeams = bstart.back();
  }
  *handle = b;
  return 0;
}

extern "C" int bfk_batch_launch(void *handle, void *stream) {
  Batch *b = (Batch *)handle;
  for (const Group &g : b->groups) {
    const int rc = g.mixed ? bfk_beamform_mixed_batch(g.prec, g.tavg, g.d_jobs, g.d_tile_start, g.ntiles, stream)
                   : !g.shared ? bfk_beamform_classic_batch(g.prec, g.tavg, g.nkc, g.nb, g.split, g.d_jobs,
                                                          g.d_tile_start, g.njobs, g.ntiles, stream)
                   : g.nkc == 1 ? bfk_beamform_shared_batch(g.prec, g.tavg, g.d_jobs, g.d_tile_start, g.njobs,
                                                            g.ntiles, stream)
                                : bfk_beamform_shared8_batch(g.prec, g.tavg, g.nkc, g.d_jobs, g.d_tile_start,
                                                             g.njobs, g.ntiles, stream);
    if (rc) return rc;
  }
  return b->nbeams ? bfk_beam_power_batch(b->d_bp, b->d_beam_start, b->nbeams, stream) : 0;
}

extern "C" int bfk_batch_launches(void *handle) {
  if (!handle) return 0;
  const Batch *b = (const Batch *)handle;
  return (int)b->groups.size() + (b->nbeams ? 1 : 0);
}

// bf_jobs.hip -- many blocks in one launch per kernel instantiation.
//
// The reference beamforms one fixed-shape block at a time (five launches per
// 32-sample int, dsaX_beamformer.cu:955-1003).  A batch here holds blocks of
// different channel, beam and time counts, each with its own weights, buffers
// and tile shape (variant).  Blocks that share a classic kernel instantiation
// (schedule, precision, tavg, antenna chunks, beam tiles per wave, sample
// split) go into one launch whose workgroups walk a device table of jobs;
// every job keeps its own BfArgs, so shapes mix freely.  The schedules are
// the classic one (every tile shape) and the shared-unpack one (256-beam
// tiles: 4 waves to 64 antennas, 8 waves above).  bfk_batch_launch allocates nothing and never
// synchronises with the host, so a caller can capture it in a HIP graph.
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

namespace {
struct Group {
  int shared, prec, tavg, nkc, nb, split;
  int njobs, ntiles;
  BfArgs *d_jobs;
  int *d_tile_start;     // njobs + 1 prefix sums
};
struct Batch {
  std::vector<Group> groups;
};
}

extern "C" void bfk_batch_free(void *handle) {
  Batch *b = (Batch *)handle;
  if (!b) return;
  for (Group &g : b->groups) {
    (void)hipFree(g.d_jobs);
    (void)hipFree(g.d_tile_start);
  }
  delete b;
}

extern "C" int bfk_batch_prepare(const bfk_dims *dims, const bfk_batch_bufs *bufs, int n, void **handle) {
  *handle = nullptr;
  if (n < 1) {
    bfk_set_error("bfk_batch_prepare: empty batch");
    return -1;
  }
  struct Job {
    int shared, prec, tavg, nkc, nb;
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
  }
  Batch *b = new Batch;
  std::vector<char> done(n, 0);
  for (int i = 0; i < n; i++) {
    if (done[i]) continue;
    Group g{jobs[i].shared, jobs[i].prec, jobs[i].tavg, jobs[i].nkc, jobs[i].nb, jobs[i].a.split,
            0, 0, nullptr, nullptr};
    std::vector<BfArgs> tab;
    std::vector<int> start{0};
    for (int k = i; k < n; k++) {
      const Job &j = jobs[k];
      if (done[k] || j.shared != g.shared || j.prec != g.prec || j.tavg != g.tavg || j.nkc != g.nkc ||
          j.nb != g.nb || j.a.split != g.split)
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
  *handle = b;
  return 0;
}

extern "C" int bfk_batch_launch(void *handle, void *stream) {
  Batch *b = (Batch *)handle;
  for (const Group &g : b->groups) {
    const int rc = !g.shared ? bfk_beamform_classic_batch(g.prec, g.tavg, g.nkc, g.nb, g.split, g.d_jobs,
                                                          g.d_tile_start, g.njobs, g.ntiles, stream)
                   : g.nkc == 1 ? bfk_beamform_shared_batch(g.prec, g.tavg, g.d_jobs, g.d_tile_start, g.njobs,
                                                            g.ntiles, stream)
                                : bfk_beamform_shared8_batch(g.prec, g.tavg, g.nkc, g.d_jobs, g.d_tile_start,
                                                             g.njobs, g.ntiles, stream);
    if (rc) return rc;
  }
  return 0;
}

extern "C" int bfk_batch_launches(void *handle) {
  return handle ? (int)((Batch *)handle)->groups.size() : 0;
}

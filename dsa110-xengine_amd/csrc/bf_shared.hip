// bf_shared.hip -- the "shared" schedule of the fused beamformer: the waves
// of a tile unpack each (channel, pol) once, into LDS MFMA operands that all
// of them read, instead of each wave unpacking the same voltages.
//
// Reference path: src/dsaX_beamformer.cu promoter (:170-193), beamformer
// (:267-416), badder (:78-112) and adder (:120-158), fused into one launch.
//
// Why.  On gfx950 a 32x32 MFMA hides only a few VALU instructions issued
// beside it (about 4 FP32 ops behind a 32-cycle FP16 MFMA, 8 behind a 64-cycle
// FP8 one); beyond that VALU time adds to MFMA time.  In the classic schedule
// each of the 4 waves of a tile turns the same packed voltages into its own
// operands, and unpack + |B|^2 detect come to 6-12 VALU instructions per MFMA.
// Here the 4 waves of a tile (one 64-beam group each, both 32-sample units,
// as in the classic schedule) share the unpack: per (channel, pol) each thread
// converts a quarter's worth of packed bytes into A_re operands in LDS, and
// every wave derives A_im from A_re with XORs and moves (A_im = [Di, Dr] is a
// swap-and-negate of A_re = [Dr, -Di]).  That brings the VALU work to ~4 (FP16)
// and ~8 (FP8) instructions per MFMA.
//
// FP6-split (E2M3 voltage codes v/8 and E2M3 hi + lo weights, each part
// boosted per lane by a power of two that the MFMA's E8M0 operand scale takes
// back) runs the FP8
// chains at twice the matrix rate: a 32x32x64 MFMA on 6-bit operands takes 32
// cycles instead of 64.
//
// LDS stays within 80 KB, so two workgroups share a CU and cover each other's
// staging and barriers: the packed stage uses 64-byte rows whose 8-byte chunks
// are XOR-swizzled by row (conflict-free unpack reads), and the A_re buffer
// is double-buffered (FP16 per 32-sample unit, FP8 per (channel, pol); 2 x 8
// KB either way).  Results are bit-identical to the classic schedule.
#include "bf_shared_tile.h"

using namespace shared;

template <int PREC, int TAVG, int GROUP>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
void bf_shared_kernel(BfArgs a) {
  __shared__ __attribute__((aligned(16))) uint8_t stage[SSTAGE];
  __shared__ __attribute__((aligned(16))) uint8_t abuf[abuf_bytes<PREC>()];
  // XCD-aware order as in the classic schedule (tile_id explains it)
  const int P = blockIdx.x;
  const int L = (P & 7) * (gridDim.x >> 3) + (P >> 3);
  if (L >= a.ntiles) return;
  shared_tile<PREC, TAVG, GROUP>(a, L, stage, abuf);
}

// Batched: the tiles of njobs blocks back to back; job j owns global tiles
// [tile_start[j], tile_start[j+1]).
template <int PREC, int TAVG>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
void bf_shared_batch_kernel(const BfArgs *__restrict__ jobs, const int *__restrict__ tile_start, int njobs) {
  __shared__ __attribute__((aligned(16))) uint8_t stage[SSTAGE];
  __shared__ __attribute__((aligned(16))) uint8_t abuf[abuf_bytes<PREC>()];
  const int P = blockIdx.x;
  const int L = (P & 7) * (gridDim.x >> 3) + (P >> 3);
  if (L >= tile_start[njobs]) return;
  int j = 0;
  while (L >= tile_start[j + 1]) j++;
  shared_tile<PREC, TAVG>(jobs[j], L - tile_start[j], stage, abuf);
}

template <int PREC, int GROUP>
static void launch_g(int tavg, dim3 grid, hipStream_t s, const BfArgs &a) {
  switch (tavg) {
    case 8:  hipLaunchKernelGGL((bf_shared_kernel<PREC, 8, GROUP>), grid, dim3(NT), 0, s, a); break;
    case 16: hipLaunchKernelGGL((bf_shared_kernel<PREC, 16, GROUP>), grid, dim3(NT), 0, s, a); break;
    case 32: hipLaunchKernelGGL((bf_shared_kernel<PREC, 32, GROUP>), grid, dim3(NT), 0, s, a); break;
    default: hipLaunchKernelGGL((bf_shared_kernel<PREC, 64, GROUP>), grid, dim3(NT), 0, s, a); break;
  }
}

// Group weights (one row per pol for a whole tile): the FP8 and FP6 hi+lo
// splits at tavg 32 and 64 keep both rows in registers (245-252 VGPRs);
// FP16, the three-part FP6 split and tavg 8 or 16 (more bin sums) would
// spill with both, so they hold one row at a time, pol-outer.
template <int PREC>
static void launch(int tavg, dim3 grid, hipStream_t s, const BfArgs &a) {
  constexpr int G = PREC == 1 || PREC == 2 ? 1 : 2;
  if (a.wmode != 0) {
    launch_g<PREC, 0>(tavg, grid, s, a);
  } else if (G == 1 && tavg >= 32) {
    if (tavg == 32) hipLaunchKernelGGL((bf_shared_kernel<PREC, 32, G>), grid, dim3(NT), 0, s, a);
    else hipLaunchKernelGGL((bf_shared_kernel<PREC, 64, G>), grid, dim3(NT), 0, s, a);
  } else {
    launch_g<PREC, 2>(tavg, grid, s, a);
  }
}

template <int PREC>
static void launch_batch(int tavg, dim3 grid, hipStream_t s, const BfArgs *jobs, const int *ts, int nj) {
  switch (tavg) {
    case 8:  hipLaunchKernelGGL((bf_shared_batch_kernel<PREC, 8>), grid, dim3(NT), 0, s, jobs, ts, nj); break;
    case 16: hipLaunchKernelGGL((bf_shared_batch_kernel<PREC, 16>), grid, dim3(NT), 0, s, jobs, ts, nj); break;
    case 32: hipLaunchKernelGGL((bf_shared_batch_kernel<PREC, 32>), grid, dim3(NT), 0, s, jobs, ts, nj); break;
    default: hipLaunchKernelGGL((bf_shared_batch_kernel<PREC, 64>), grid, dim3(NT), 0, s, jobs, ts, nj); break;
  }
}

static int shared_shape_ok(const bfk_dims *d) {
  if (d->nkc != 1 || d->nant_pad != 64 || d->nsamp % TCU != 0) {
    bfk_set_error("bfk_beamform: shared schedule needs nant <= 64 (nant_pad=%d)", d->nant_pad);
    return 0;
  }
  return 1;
}

extern "C" int bfk_beamform_shared(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag,
                                   const float *d_wscale, const float *d_scale, uint8_t *d_out,
                                   float *d_power, float *d_bpart, void *stream) {
  if (!shared_shape_ok(d)) return -1;
  const BfArgs a = bf_make_args(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart, BSET);
  const dim3 grid((a.ntiles + 7) / 8 * 8);
  if (d->prec == 0) launch<0>(d->tavg, grid, (hipStream_t)stream, a);
  else if (d->prec == 1) launch<1>(d->tavg, grid, (hipStream_t)stream, a);
  else if (d->prec == 2) launch<2>(d->tavg, grid, (hipStream_t)stream, a);
  else launch<3>(d->tavg, grid, (hipStream_t)stream, a);
  return bfk_check(hipGetLastError(), "bf_shared_kernel");
}

// The kernel arguments of one block for a batch job table.
extern "C" int bfk_shared_args(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag, const float *d_wscale,
                               const float *d_scale, uint8_t *d_out, float *d_power, float *d_bpart, BfArgs *a) {
  if (!shared_shape_ok(d)) return -1;
  *a = bf_make_args(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart, BSET);
  return 0;
}

extern "C" int bfk_beamform_shared_batch(int prec, int tavg, const void *d_jobs, const int *d_tile_start,
                                         int njobs, int ntiles, void *stream) {
  const dim3 grid((ntiles + 7) / 8 * 8);
  if (prec == 0) launch_batch<0>(tavg, grid, (hipStream_t)stream, (const BfArgs *)d_jobs, d_tile_start, njobs);
  else if (prec == 1) launch_batch<1>(tavg, grid, (hipStream_t)stream, (const BfArgs *)d_jobs, d_tile_start, njobs);
  else if (prec == 2) launch_batch<2>(tavg, grid, (hipStream_t)stream, (const BfArgs *)d_jobs, d_tile_start, njobs);
  else launch_batch<3>(tavg, grid, (hipStream_t)stream, (const BfArgs *)d_jobs, d_tile_start, njobs);
  return bfk_check(hipGetLastError(), "bf_shared_batch_kernel");
}

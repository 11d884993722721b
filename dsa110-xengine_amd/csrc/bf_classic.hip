// bf_classic.hip -- the "classic" schedule of the fused beamformer: one tile
// per workgroup, two workgroups resident per CU (at most 80 KB of LDS and 256
// registers per wave each), so that one workgroup's staging and epilogue
// overlap the other's MFMA work.  Also the batched form: the tiles of many
// blocks (each with its own shape, weights and buffers) in one launch.
//
// Reference path: src/dsaX_beamformer.cu promoter (:170-193), beamformer
// (:267-416), badder (:78-112) and adder (:120-158), fused into one launch.
// Per tile (8-channel group x 64 samples x a set of beams):
//   (a) each thread loads the chunks (channels x 2 t x 2 pols) of 4 antennas of
//       one packet, transposes them 4x4 and writes one dword into each staged
//       [channel][pol][sample] row they feed;
//   (b) per (channel, pol) the complex GEMM in the real 4M form on MFMA,
//       FP32 accumulation, the wave's weights in registers;
//   (c) |B|^2, the sums over pols, channels and tavg samples, scaling and the
//       uint8 requantisation in registers.
//
// Tile shapes.  A wave owns NB 32-beam tiles.  Without a split the 4 waves take
// 4 x NB x 32 beams and both 32-sample units of the tile; with a split, wave
// pairs share NB x 32 beams and take one unit each, so a tile has half the
// beams (256, 128 or 64 beams per tile over NB = 2, 1 and the split).  Every
// shape runs the same MFMA chains and sums in the same order, so all give the
// same bits; which is fastest depends on the block (few beams favour small
// tiles).
//
// Arrays of more than 64 antennas (NKC = 2 or 4 chunks of 64) keep the LDS
// stage at the same size by staging the group's channels in NKC sub-steps of
// 8/NKC channels, each row holding all antennas; a wave then owns one 32-beam
// tile (NB = 1) so that its weights (all NKC chunks) still fit in registers.
#include "bf_classic_tile.h"

using namespace classic;

template <int PREC, int TAVG, int NKC, int NB, int SPLIT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
void bf_classic_kernel(BfArgs a) {
  __shared__ __attribute__((aligned(16))) uint8_t stage[Shape<NKC>::STAGE];
  // XCD-aware order: workgroup P runs on XCD P % 8; consecutive tiles L go
  // to one XCD (tile_id explains the order).
  const int P = blockIdx.x;
  const int L = (P & 7) * (gridDim.x >> 3) + (P >> 3);
  if (L >= a.ntiles) return;
  classic_tile<PREC, TAVG, NKC, NB, SPLIT>(a, L, stage);
}

// Batched: the tiles of njobs blocks back to back; job j owns global tiles
// [tile_start[j], tile_start[j+1]).
template <int PREC, int TAVG, int NKC, int NB, int SPLIT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
void bf_classic_batch_kernel(const BfArgs *__restrict__ jobs, const int *__restrict__ tile_start,
                             int njobs) {
  __shared__ __attribute__((aligned(16))) uint8_t stage[Shape<NKC>::STAGE];
  const int P = blockIdx.x;
  const int L = (P & 7) * (gridDim.x >> 3) + (P >> 3);
  if (L >= tile_start[njobs]) return;
  int j = 0;
  while (L >= tile_start[j + 1]) j++;
  classic_tile<PREC, TAVG, NKC, NB, SPLIT>(jobs[j], L - tile_start[j], stage);
}

// ---- host side --------------------------------------------------------------

// NB and split of a classic variant; 0 if the variant is not a classic one.
static int classic_shape(int variant, int nkc, int *nb, int *split) {
  switch (variant) {
    case BFK_VAR_CLASSIC_256: *nb = 2; *split = 0; return nkc == 1;
    case BFK_VAR_CLASSIC_128S: *nb = 2; *split = 1; return nkc == 1;
    case BFK_VAR_CLASSIC_128: *nb = 1; *split = 0; return 1;
    case BFK_VAR_CLASSIC_64: *nb = 1; *split = 1; return 1;
    default: return 0;
  }
}

extern "C" int bfk_classic_beams(int variant, int nkc) {
  int nb, split;
  if (!classic_shape(variant, nkc, &nb, &split)) return 0;
  return (NW >> split) * nb * 32;
}

template <int PREC, int NKC, int NB, int SPLIT>
static void launch(int tavg, dim3 grid, hipStream_t s, const BfArgs *a, const BfArgs *jobs,
                   const int *tile_start, int njobs) {
#define BF_LAUNCH(T)                                                                                  \
  do {                                                                                                \
    if (a) hipLaunchKernelGGL((bf_classic_kernel<PREC, T, NKC, NB, SPLIT>), grid, dim3(NT), 0, s, *a); \
    else hipLaunchKernelGGL((bf_classic_batch_kernel<PREC, T, NKC, NB, SPLIT>), grid, dim3(NT), 0, s,  \
                            jobs, tile_start, njobs);                                                 \
  } while (0)
  switch (tavg) {
    case 8: BF_LAUNCH(8); break;
    case 16: BF_LAUNCH(16); break;
    case 32: BF_LAUNCH(32); break;
    default: BF_LAUNCH(64); break;
  }
#undef BF_LAUNCH
}

// Launch either one job (a != NULL) or a batch (device tables) of one
// instantiation (prec, tavg, nkc, nb).
static int dispatch(int prec, int tavg, int nkc, int nb, int split, dim3 grid, hipStream_t s,
                    const BfArgs *a, const BfArgs *jobs, const int *tile_start, int njobs) {
  const int key = (nkc << 3) | (nb << 2) | (split << 1) | prec;
#define BF_CASE(P, K, B, S)                                                     \
  case (K << 3) | (B << 2) | (S << 1) | P:                                      \
    launch<P, K, B, S>(tavg, grid, s, a, jobs, tile_start, njobs);              \
    break
  switch (key) {
    BF_CASE(0, 1, 2, 0); BF_CASE(1, 1, 2, 0); BF_CASE(0, 1, 2, 1); BF_CASE(1, 1, 2, 1);
    BF_CASE(0, 1, 1, 0); BF_CASE(1, 1, 1, 0); BF_CASE(0, 1, 1, 1); BF_CASE(1, 1, 1, 1);
    BF_CASE(0, 2, 1, 0); BF_CASE(1, 2, 1, 0); BF_CASE(0, 2, 1, 1); BF_CASE(1, 2, 1, 1);
    BF_CASE(0, 4, 1, 0); BF_CASE(1, 4, 1, 0); BF_CASE(0, 4, 1, 1); BF_CASE(1, 4, 1, 1);
    default:
      bfk_set_error("bfk_beamform: no classic kernel for %d antenna chunks x %d beam tiles", nkc, nb);
      return -1;
  }
#undef BF_CASE
  return bfk_check(hipGetLastError(), "bf_classic_kernel");
}

extern "C" int bfk_classic_args(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag,
                                const float *d_wscale, const float *d_scale, uint8_t *d_out,
                                float *d_power, float *d_bpart, BfArgs *a, int *nb) {
  int split;
  if (d->nant_pad != 64 * d->nkc || (d->nkc != 1 && d->nkc != 2 && d->nkc != 4) ||
      !classic_shape(d->variant, d->nkc, nb, &split) || d->nsamp % TCU != 0) {
    bfk_set_error("bfk_beamform: variant %d cannot run nant_pad=%d nsamp=%d", d->variant, d->nant_pad,
                  d->nsamp);
    return -1;
  }
  *a = bf_make_args(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart,
                    (NW >> split) * *nb * 32);
  a->split = split;
  return 0;
}

extern "C" int bfk_beamform_classic(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag,
                                    const float *d_wscale, const float *d_scale, uint8_t *d_out,
                                    float *d_power, float *d_bpart, void *stream) {
  BfArgs a;
  int nb;
  if (bfk_classic_args(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart, &a, &nb)) return -1;
  const dim3 grid((a.ntiles + 7) / 8 * 8);
  return dispatch(d->prec, d->tavg, d->nkc, nb, a.split, grid, (hipStream_t)stream, &a, NULL, NULL, 0);
}

extern "C" int bfk_beamform_classic_batch(int prec, int tavg, int nkc, int nb, int split,
                                          const void *d_jobs, const int *d_tile_start, int njobs,
                                          int ntiles, void *stream) {
  const dim3 grid((ntiles + 7) / 8 * 8);
  return dispatch(prec, tavg, nkc, nb, split, grid, (hipStream_t)stream, NULL, (const BfArgs *)d_jobs,
                  d_tile_start, njobs);
}

// bf_mixed.hip -- one launch for all the 64-antenna blocks of a batch that
// share a precision (FP16 or FP8-split) and an averaging length, whatever
// their channel, beam and time counts and tile shapes.
//
// Reference: the per-int launch loop of dsaX_beamformer.cu:955-1003 (five
// launches per 32-sample int of one fixed-shape block).  Here the job table
// lists each block's arguments and its tile schedule (a classic tile shape
// or the shared-unpack tile, as its plan's variant says); the tiles of all
// jobs are numbered back to back, one workgroup each.  Workgroup P runs on
// XCD P % 8 and takes tile (P % 8) * (grid / 8) + P / 8, so consecutive tiles
// (the channel groups that share input lines, the tiles that share weights)
// run together on one L2.  Each tile runs the same device function as the
// block's own kernel, so the bits equal a per-block execute.
//
// A persistent form (two workgroups per CU walking the tile list) held the
// loop state beside the largest tile's registers and spilled; one tile per
// workgroup keeps every schedule at its own register count (180-208).
#include "bf_classic_tile.h"
#include "bf_shared_tile.h"

namespace {
constexpr int MIX_NT = 256;                 // both schedules: 4 waves
constexpr int MIX_LDS = 80 * 1024;          // classic stage, or shared stage + operands
static_assert(classic::NT == MIX_NT && shared::NT == MIX_NT, "one workgroup shape");
static_assert(classic::Shape<1>::STAGE <= MIX_LDS, "classic stage fits");
static_assert(shared::SSTAGE + shared::abuf_bytes<0>() <= MIX_LDS &&
              shared::SSTAGE + shared::abuf_bytes<1>() <= MIX_LDS, "shared stage fits");
}  // namespace

// BfArgs::kind of a mixed batch job
enum { MIX_CLASSIC_256 = 0, MIX_CLASSIC_128S = 1, MIX_CLASSIC_128 = 2, MIX_CLASSIC_64 = 3, MIX_SHARED_256 = 4 };

template <int PREC, int TAVG>
__global__ __launch_bounds__(MIX_NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
void bf_mixed_batch_kernel(const BfArgs *__restrict__ jobs, const int *__restrict__ tile_start, int ntiles) {
  __shared__ __attribute__((aligned(16))) uint8_t lds[MIX_LDS];
  const int P = blockIdx.x;
  const int L = (P & 7) * (gridDim.x >> 3) + (P >> 3);
  if (L >= ntiles) return;
  int j = 0;
  while (L >= tile_start[j + 1]) j++;
  const BfArgs &a = jobs[j];
  const int tl = L - tile_start[j];
  switch (a.kind) {
    case MIX_CLASSIC_256: classic::classic_tile<PREC, TAVG, 1, 2, 0>(a, tl, lds); break;
    case MIX_CLASSIC_128S: classic::classic_tile<PREC, TAVG, 1, 2, 1>(a, tl, lds); break;
    case MIX_CLASSIC_128: classic::classic_tile<PREC, TAVG, 1, 1, 0>(a, tl, lds); break;
    case MIX_CLASSIC_64: classic::classic_tile<PREC, TAVG, 1, 1, 1>(a, tl, lds); break;
    default: shared::shared_tile<PREC, TAVG>(a, tl, lds, lds + shared::SSTAGE); break;
  }
}

extern "C" int bfk_mix_kind(int variant) {
  switch (variant) {
    case BFK_VAR_CLASSIC_256: return MIX_CLASSIC_256;
    case BFK_VAR_CLASSIC_128S: return MIX_CLASSIC_128S;
    case BFK_VAR_CLASSIC_128: return MIX_CLASSIC_128;
    case BFK_VAR_CLASSIC_64: return MIX_CLASSIC_64;
    default: return MIX_SHARED_256;
  }
}

extern "C" int bfk_beamform_mixed_batch(int prec, int tavg, const void *d_jobs, const int *d_tile_start,
                                        int ntiles, void *stream) {
  // one workgroup per tile, the grid a multiple of the 8 XCDs
  const dim3 grid((ntiles + 7) / 8 * 8), block(MIX_NT);
  hipStream_t s = (hipStream_t)stream;
  const BfArgs *jobs = (const BfArgs *)d_jobs;
#define MIX_LAUNCH(P, T) hipLaunchKernelGGL((bf_mixed_batch_kernel<P, T>), grid, block, 0, s, jobs, d_tile_start, ntiles)
  if (prec == 0) {
    switch (tavg) {
      case 8: MIX_LAUNCH(0, 8); break;
      case 16: MIX_LAUNCH(0, 16); break;
      case 32: MIX_LAUNCH(0, 32); break;
      default: MIX_LAUNCH(0, 64); break;
    }
  } else if (prec == 1) {
    switch (tavg) {
      case 8: MIX_LAUNCH(1, 8); break;
      case 16: MIX_LAUNCH(1, 16); break;
      case 32: MIX_LAUNCH(1, 32); break;
      default: MIX_LAUNCH(1, 64); break;
    }
  } else {
    bfk_set_error("bfk_beamform_mixed_batch: precision %d has no classic tiles", prec);
    return -1;
  }
#undef MIX_LAUNCH
  return bfk_check(hipGetLastError(), "bf_mixed_batch_kernel");
}

// bf_classic.hip -- the "classic" schedule of the fused beamformer: one tile
// per workgroup, two workgroups resident per CU (80 KB of LDS and at most
// 256 registers per wave each), so that one workgroup's staging and epilogue
// overlap the other's MFMA work.
//
// Reference path: src/dsaX_beamformer.cu promoter (:170-193), beamformer
// (:267-416), badder (:78-112) and adder (:120-158), fused into one launch.
// Per tile (8-channel group x 64 samples x 256 beams):
//   (a) each thread loads the 32-byte chunks (8 channels x 2 t x 2 pols) of 4
//       antennas of one packet, transposes them 4x4 and writes one dword into
//       each of the 32 [channel][pol][sample] rows they feed;
//   (b) per (channel, pol) the complex GEMM in the real 4M form on MFMA,
//       FP32 accumulation, weights of the wave's 64 beams in registers;
//   (c) |B|^2, the sums over pols, channels and tavg samples, scaling and the
//       uint8 requantisation in registers.
#include "bf_common.h"

namespace {
constexpr int NW = 4, NT = NW * 64;       // waves, threads per workgroup
constexpr int BSET = NW * BTW * 32;       // beams per tile
}

template <int PREC, int TAVG>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
void bf_classic_kernel(BfArgs a) {
  __shared__ __attribute__((aligned(16))) uint8_t stage[STAGE_BYTES];
  constexpr int NP = TAVG <= 32 ? 32 / TAVG : 1;
  constexpr int NTT = TAVG == 64 ? 1 : 2;

  // XCD-aware order: workgroup P runs on XCD P % 8; consecutive tiles L go
  // to one XCD, so its resident workgroups share one group's weights in L2.
  const int P = blockIdx.x;
  const int L = (P & 7) * (gridDim.x >> 3) + (P >> 3);
  if (L >= a.ntiles) return;
  const TileId t = tile_id(a, L);
  const int tid = threadIdx.x;

  // (a) stage
  for (int task = tid; task < PCU * 16; task += NT) {
    const int pl = task >> 4, aq = task & 15;
    const long p = (long)t.tblk * PCU + pl;
    uint32_t x[4][8];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int ant = aq * 4 + j;
      if (ant < a.nant) {
        const uint4 *src = (const uint4 *)(a.in + chunk_offset(a.order, p, ant, t.g, a.nant, a.nchan,
                                                               a.ngroup));
        const uint4 v0 = src[0], v1 = src[1];
        x[j][0] = v0.x; x[j][1] = v0.y; x[j][2] = v0.z; x[j][3] = v0.w;
        x[j][4] = v1.x; x[j][5] = v1.y; x[j][6] = v1.z; x[j][7] = v1.w;
      } else {
#pragma unroll
        for (int w = 0; w < 8; w++) x[j][w] = 0u;
      }
    }
    stage_store(stage, x, pl, aq);
  }
  __syncthreads();

  const int lane = tid & 63, wave = tid >> 6;
  const int n = lane & 31, h = lane >> 5;
  const int btw = t.bset * (BSET / 32) + wave * BTW;   // first beam tile of this wave

  float part[NTT][BTW][NP];
#pragma unroll
  for (int i = 0; i < NTT; i++)
#pragma unroll
    for (int bt = 0; bt < BTW; bt++)
#pragma unroll
      for (int j = 0; j < NP; j++) part[i][bt][j] = 0.f;

  // (b) 16 (channel, pol) GEMMs, two 32-sample units each
  for (int cq = 0; cq < 16; cq++) {
    const long wr = weight_row(a, t.g, cq);
    WFrag<PREC> W;
    W.load(a.wfrag, a.wscale, wr * a.nbt + btw, wr * a.nbeam_pad + btw * 32, lane, n);
#pragma unroll
    for (int tt = 0; tt < 2; tt++) {
      Acc A;
      gemm_unit<PREC>(A, W, stage + (cq * TCU + tt * 32) * ROWB, n, h);
      detect<TAVG>(A, W.sc, part[NTT == 1 ? 0 : tt]);
    }
  }

  // (c)
  tile_output<TAVG>(a, t, btw, n, h, part);
}

template <int PREC>
static void launch(int tavg, dim3 grid, hipStream_t s, const BfArgs &a) {
  switch (tavg) {
    case 8:  hipLaunchKernelGGL((bf_classic_kernel<PREC, 8>), grid, dim3(NT), 0, s, a); break;
    case 16: hipLaunchKernelGGL((bf_classic_kernel<PREC, 16>), grid, dim3(NT), 0, s, a); break;
    case 32: hipLaunchKernelGGL((bf_classic_kernel<PREC, 32>), grid, dim3(NT), 0, s, a); break;
    default: hipLaunchKernelGGL((bf_classic_kernel<PREC, 64>), grid, dim3(NT), 0, s, a); break;
  }
}

extern "C" int bfk_beamform_classic(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag,
                                    const float *d_wscale, const float *d_scale, uint8_t *d_out,
                                    float *d_power, float *d_bpart, void *stream) {
  if (d->nant_pad != 64 || d->nbeam_pad % BSET != 0 || d->nsamp % TCU != 0) {
    bfk_set_error("bfk_beamform: unsupported shape nant_pad=%d nbeam_pad=%d nsamp=%d", d->nant_pad,
                  d->nbeam_pad, d->nsamp);
    return -1;
  }
  const BfArgs a = bf_make_args(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart, BSET);
  const dim3 grid((a.ntiles + 7) / 8 * 8);
  if (d->prec == 0) launch<0>(d->tavg, grid, (hipStream_t)stream, a);
  else launch<1>(d->tavg, grid, (hipStream_t)stream, a);
  return bfk_check(hipGetLastError(), "bf_classic_kernel");
}

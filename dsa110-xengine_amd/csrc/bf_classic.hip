// bf_classic.hip -- the "classic" schedule of the fused beamformer: one tile
// per workgroup, two workgroups resident per CU (at most 80 KB of LDS and 256
// registers per wave each), so that one workgroup's staging and epilogue
// overlap the other's MFMA work.
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
// Arrays of more than 64 antennas (NKC = 2 or 4 chunks of 64) keep the LDS
// stage at the same size by staging the group's channels in NKC sub-steps of
// 8/NKC channels, each row holding all antennas; a wave then owns one 32-beam
// tile so that its weights (all NKC chunks) still fit in registers.
#include "bf_common.h"

namespace {
constexpr int NW = 4, NT = NW * 64;       // waves, threads per workgroup

template <int NKC> struct Shape {
  static constexpr int NB = NKC == 1 ? BTW : 1;               // beam tiles per wave
  static constexpr int BSET = NW * NB * 32;                   // beams per tile
  static constexpr int ROWB = NKC == 1 ? 80 : 64 * NKC + 16;  // staged row stride
  static constexpr int NCI = 8 / NKC;                         // channels per sub-step
  static constexpr int CQS = 2 * NCI;                         // (channel, pol) per sub-step
  static constexpr int STAGE = CQS * TCU * ROWB;
  static_assert(STAGE <= 80 * 1024, "two workgroups per CU");
};
}

template <int PREC, int TAVG, int NKC>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
void bf_classic_kernel(BfArgs a) {
  using S = Shape<NKC>;
  constexpr int NB = S::NB, ROWB_ = S::ROWB, NCI = S::NCI, CQS = S::CQS;
  __shared__ __attribute__((aligned(16))) uint8_t stage[S::STAGE];
  constexpr int NP = TAVG <= 32 ? 32 / TAVG : 1;
  constexpr int NTT = TAVG == 64 ? 1 : 2;

  // XCD-aware order: workgroup P runs on XCD P % 8; consecutive tiles L go
  // to one XCD (tile_id explains the order).
  const int P = blockIdx.x;
  const int L = (P & 7) * (gridDim.x >> 3) + (P >> 3);
  if (L >= a.ntiles) return;
  const TileId t = tile_id(a, L);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int n = lane & 31, h = lane >> 5;
  const int btw = t.bset * (S::BSET / 32) + wave * NB;   // first beam tile of this wave

  float part[NTT][NB][NP];
#pragma unroll
  for (int i = 0; i < NTT; i++)
#pragma unroll
    for (int bt = 0; bt < NB; bt++)
#pragma unroll
      for (int j = 0; j < NP; j++) part[i][bt][j] = 0.f;

  for (int sub = 0; sub < NKC; sub++) {
    // (a) stage channels sub*NCI .. sub*NCI + NCI - 1 of the group
    if (sub) __syncthreads();          // the previous sub-step's rows are read
    for (int task = tid; task < PCU * 16 * NKC; task += NT) {
      const int pl = task / (16 * NKC), aq = task % (16 * NKC);
      const long p = (long)t.tblk * PCU + pl;
      uint32_t x[4][NCI];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int ant = aq * 4 + j;
        if (ant < a.nant) {
          const uint8_t *src = a.in + chunk_offset(a.order, p, ant, t.g, a.nant, a.nchan, a.ngroup) +
                               sub * NCI * 4;
          if constexpr (NCI == 8) {
            const uint4 v0 = ((const uint4 *)src)[0], v1 = ((const uint4 *)src)[1];
            x[j][0] = v0.x; x[j][1] = v0.y; x[j][2] = v0.z; x[j][3] = v0.w;
            x[j][4] = v1.x; x[j][5] = v1.y; x[j][6] = v1.z; x[j][7] = v1.w;
          } else if constexpr (NCI == 4) {
            const uint4 v0 = *(const uint4 *)src;
            x[j][0] = v0.x; x[j][1] = v0.y; x[j][2] = v0.z; x[j][3] = v0.w;
          } else {
            const uint2 v0 = *(const uint2 *)src;
            x[j][0] = v0.x; x[j][1] = v0.y;
          }
        } else {
#pragma unroll
          for (int w = 0; w < NCI; w++) x[j][w] = 0u;
        }
      }
      stage_store<NCI>(stage, x, pl, aq, ROWB_);
    }
    __syncthreads();

    // (b) CQS (channel, pol) GEMMs, two 32-sample units each
    for (int cql = 0; cql < CQS; cql++) {
      const long wr = weight_row(a, t.g, sub * CQS + cql);
      WFrag<PREC, NB, NKC> W;
      W.load(a.wfrag, a.wscale, wr * a.nbt + btw, wr * a.nbeam_pad + btw * 32, lane, n);
#pragma unroll
      for (int tt = 0; tt < 2; tt++) {
        Acc<NB> A;
        gemm_unit(A, W, stage + (cql * TCU + tt * 32) * ROWB_, ROWB_, n, h);
        detect<TAVG>(A, W.sc, part[NTT == 1 ? 0 : tt]);
      }
    }
  }

  // (c)
  tile_output<TAVG>(a, t, btw, n, h, part);
}

template <int PREC, int NKC>
static void launch(int tavg, dim3 grid, hipStream_t s, const BfArgs &a) {
  switch (tavg) {
    case 8:  hipLaunchKernelGGL((bf_classic_kernel<PREC, 8, NKC>), grid, dim3(NT), 0, s, a); break;
    case 16: hipLaunchKernelGGL((bf_classic_kernel<PREC, 16, NKC>), grid, dim3(NT), 0, s, a); break;
    case 32: hipLaunchKernelGGL((bf_classic_kernel<PREC, 32, NKC>), grid, dim3(NT), 0, s, a); break;
    default: hipLaunchKernelGGL((bf_classic_kernel<PREC, 64, NKC>), grid, dim3(NT), 0, s, a); break;
  }
}

template <int NKC>
static int beamform(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag, const float *d_wscale,
                    const float *d_scale, uint8_t *d_out, float *d_power, float *d_bpart,
                    hipStream_t s) {
  if (d->nbeam_pad % Shape<NKC>::BSET != 0 || d->nsamp % TCU != 0) {
    bfk_set_error("bfk_beamform: unsupported shape nbeam_pad=%d nsamp=%d", d->nbeam_pad, d->nsamp);
    return -1;
  }
  const BfArgs a = bf_make_args(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart,
                                Shape<NKC>::BSET);
  const dim3 grid((a.ntiles + 7) / 8 * 8);
  if (d->prec == 0) launch<0, NKC>(d->tavg, grid, s, a);
  else launch<1, NKC>(d->tavg, grid, s, a);
  return bfk_check(hipGetLastError(), "bf_classic_kernel");
}

extern "C" int bfk_beamform_classic(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag,
                                    const float *d_wscale, const float *d_scale, uint8_t *d_out,
                                    float *d_power, float *d_bpart, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  if (d->nant_pad != 64 * d->nkc) {
    bfk_set_error("bfk_beamform: nant_pad=%d is not 64*nkc", d->nant_pad);
    return -1;
  }
  switch (d->nkc) {
    case 1: return beamform<1>(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart, s);
    case 2: return beamform<2>(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart, s);
    case 4: return beamform<4>(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart, s);
    default:
      bfk_set_error("bfk_beamform: %d antenna chunks unsupported (1, 2 or 4)", d->nkc);
      return -1;
  }
}

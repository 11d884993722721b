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
#include "bf_common.h"

namespace {
constexpr int NW = 4, NT = NW * 64;       // waves, threads per workgroup

template <int NKC> struct Shape {
  static constexpr int ROWB = NKC == 1 ? 80 : 64 * NKC + 16;  // staged row stride
  static constexpr int NCI = 8 / NKC;                         // channels per sub-step
  static constexpr int CQS = 2 * NCI;                         // (channel, pol) per sub-step
  static constexpr int STAGE = CQS * TCU * ROWB;
  static_assert(STAGE <= 80 * 1024, "two workgroups per CU");
};
}

// One tile L of job a (all threads of the workgroup).
template <int PREC, int TAVG, int NKC, int NB, int SPLIT>
__device__ __forceinline__ void classic_tile(const BfArgs &a, int L, uint8_t *stage) {
  using S = Shape<NKC>;
  constexpr int ROWB_ = S::ROWB, NCI = S::NCI, CQS = S::CQS;
  constexpr int NP = TAVG <= 32 ? 32 / TAVG : 1;
  const TileId t = tile_id(a, L);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int n = lane & 31, h = lane >> 5;
  constexpr int nbq = NW >> SPLIT;                       // beam groups per tile
  constexpr int ntt = SPLIT ? 1 : 2;                     // units per wave
  const int bq = wave % nbq, th = wave / nbq;            // th: unit of a split pair
  const int btw = (t.bset * nbq + bq) * NB;              // first beam tile of this wave
  const int tt_first = SPLIT ? th : 0;

  float part[2][NB][NP];             // per unit this wave runs
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int bt = 0; bt < NB; bt++)
#pragma unroll
      for (int j = 0; j < NP; j++) part[i][bt][j] = 0.f;

  // Weights of (channel, pol) 0 first, so that their latency overlaps the
  // staging; the last unit of each (channel, pol) prefetches the next one's in
  // place (the last one reloads itself).  Not with 256 antennas: there the
  // weights take 128 registers, and holding them through a staging sub-step
  // spills, so each (channel, pol) loads its own.
  constexpr bool PFW = NKC < 4;
  WFrag<PREC, NB, PFW ? NKC : 1> W;    // (unused with 256 antennas)
  if constexpr (PFW) {
    const long wr = weight_row(a, t.g, 0);
    W.load(a.wfrag, a.wscale, wr * a.nbt + btw, wr * a.nbeam_pad + btw * 32, lane, n);
  }

  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  typedef uint32_t u2 __attribute__((ext_vector_type(2)));
  const __amdgpu_buffer_rsrc_t rin = buffer_rsrc(a.in, a.in_bytes);
  constexpr int TPT = PCU * 16 * NKC / NT;     // staging tasks per thread
  static_assert(TPT * NT == PCU * 16 * NKC, "whole staging tasks per thread");
  // tasks whose loads are in flight together: all of them with 64 antennas,
  // half at a time above, where the weights hold 64 or 128 registers
  constexpr int TB = NKC >= 2 ? TPT / 2 : TPT;
  constexpr int TUNROLL = NKC >= 2 ? 1 : TPT / TB;   // batches unrolled
  for (int sub = 0; sub < NKC; sub++) {
    // (a) stage channels sub*NCI .. sub*NCI + NCI - 1 of the group, TB tasks'
    // loads issued before their first store
    if (sub) __syncthreads();          // the previous sub-step's rows are read
#pragma unroll TUNROLL
    for (int k0 = 0; k0 < TPT; k0 += TB) {
      uint32_t x[TB][4][NCI];
#pragma unroll
      for (int k = 0; k < TB; k++) {
        const int task = tid + (k0 + k) * NT;
        const int pl = task / (16 * NKC), aq = task % (16 * NKC);
        const long p = (long)t.tblk * PCU + pl;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          // branch-free buffer loads: padding antennas read past the block's
          // end and get zeros (a load inside a branch would be waited for at
          // the join, one round trip per antenna)
          const int ant = aq * 4 + j;
          const uint32_t off =
              ant < a.nant ? (uint32_t)chunk_offset(a.order, p, ant, t.g, a.nant, a.nchan, a.ngroup) + sub * NCI * 4
                           : 0x80000000u;
          if constexpr (NCI == 8) {
            const u4 v0 = __builtin_amdgcn_raw_buffer_load_b128(rin, off, 0, 0);
            const u4 v1 = __builtin_amdgcn_raw_buffer_load_b128(rin, off, 16, 0);
            x[k][j][0] = v0.x; x[k][j][1] = v0.y; x[k][j][2] = v0.z; x[k][j][3] = v0.w;
            x[k][j][4] = v1.x; x[k][j][5] = v1.y; x[k][j][6] = v1.z; x[k][j][7] = v1.w;
          } else if constexpr (NCI == 4) {
            const u4 v0 = __builtin_amdgcn_raw_buffer_load_b128(rin, off, 0, 0);
            x[k][j][0] = v0.x; x[k][j][1] = v0.y; x[k][j][2] = v0.z; x[k][j][3] = v0.w;
          } else {
            const u2 v0 = __builtin_amdgcn_raw_buffer_load_b64(rin, off, 0, 0);
            x[k][j][0] = v0.x; x[k][j][1] = v0.y;
          }
        }
      }
#pragma unroll
      for (int k = 0; k < TB; k++) {
        const int task = tid + (k0 + k) * NT;
        stage_store<NCI>(stage, x[k], task / (16 * NKC), task % (16 * NKC), ROWB_);
      }
    }
    __syncthreads();

    // (b) CQS (channel, pol) GEMMs, one or two 32-sample units each
    for (int cql = 0; cql < CQS; cql++) {
      const int cq = sub * CQS + cql;
      if constexpr (PFW) {
        const long wn = weight_row(a, t.g, cq < 15 ? cq + 1 : cq);
#pragma unroll
        for (int k = 0; k < ntt; k++) {
          Acc<NB> A;
          const uint8_t *rows = stage + (cql * TCU + (tt_first + k) * 32) * ROWB_;
          if (k == ntt - 1) gemm_unit<true>(A, W, rows, ROWB_, n, h, a.wfrag, wn * a.nbt + btw);
          else gemm_unit(A, W, rows, ROWB_, n, h);
          detect<TAVG>(A, W.sc, part[k]);
        }
        W.load_scale(a.wscale, wn * a.nbeam_pad + btw * 32, n);
      } else {
        const long wr = weight_row(a, t.g, cq);
        WFrag<PREC, NB, NKC> Wc;
        Wc.load(a.wfrag, a.wscale, wr * a.nbt + btw, wr * a.nbeam_pad + btw * 32, lane, n);
#pragma unroll
        for (int k = 0; k < ntt; k++) {
          Acc<NB> A;
          gemm_unit(A, Wc, stage + (cql * TCU + (tt_first + k) * 32) * ROWB_, ROWB_, n, h);
          detect<TAVG>(A, Wc.sc, part[k]);
        }
      }
    }
  }

  // (c) a split TAVG-64 bin needs both units: the second wave of each pair
  // hands its half to the first through LDS
  if (TAVG == 64 && SPLIT) {
    float *xch = (float *)stage;
    __syncthreads();                   // every wave is done with the stage
    if (th == 1)
#pragma unroll
      for (int bt = 0; bt < NB; bt++) xch[(bq * NB + bt) * 64 + lane] = part[0][bt][0];
    __syncthreads();
    if (th == 0) {
#pragma unroll
      for (int bt = 0; bt < NB; bt++) part[1][bt][0] = xch[(bq * NB + bt) * 64 + lane];
      tile_output<TAVG>(a, t, btw, n, h, part, 0, 2);
    }
    return;
  }
  tile_output<TAVG>(a, t, btw, n, h, part, tt_first, ntt);
}

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

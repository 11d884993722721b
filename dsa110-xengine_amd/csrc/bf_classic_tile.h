// bf_classic_tile.h -- one tile of the "classic" schedule of the fused
// beamformer (the schedule is described in bf_classic.hip), as a device
// function shared by the per-block kernels there and the mixed batch kernel
// (bf_mixed.hip).
#ifndef BF_CLASSIC_TILE_H
#define BF_CLASSIC_TILE_H
#include "bf_common.h"

namespace classic {
constexpr int NW = 4, NT = NW * 64;       // waves, threads per workgroup

template <int NKC> struct Shape {
  static constexpr int ROWB = NKC == 1 ? 80 : 64 * NKC + 16;  // staged row stride
  static constexpr int NCI = 8 / NKC;                         // channels per sub-step
  static constexpr int CQS = 2 * NCI;                         // (channel, pol) per sub-step
  static constexpr int STAGE = CQS * TCU * ROWB;
  static_assert(STAGE <= 80 * 1024, "two workgroups per CU");
};

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
}  // namespace classic
#endif

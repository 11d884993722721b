// bf_fused.hip -- the fused beamformer kernel for gfx950.
//
// Reference path: src/dsaX_beamformer.cu promoter (:170-193), beamformer
// (:267-416), badder (:78-112) and adder (:120-158).  The reference runs five
// launches per 32-sample "int" (640 per block) and keeps unpacked FP16 planes
// and FP32 partial powers in HBM between them.  Here one persistent kernel
// does, per tile (one 8-channel output group x 64 samples x 256 beams):
//   (a) coalesced 32-byte loads of the packed 4-bit block, transposed through
//       LDS into [channel][pol][sample][antenna] rows (MFMA K-contiguous);
//   (b) per (channel, pol) a complex GEMM in the real 4M form on
//       v_mfma_f32_32x32x16_f16 (FP16) or v_mfma_f32_32x32x64_f8f6f4
//       (FP8 E4M3 voltages x hi+lo FP8 weights), FP32 accumulation;
//   (c) |B|^2, the sums over 2 pols x 8 channels x tavg samples, scaling and
//       the uint8 requantisation in registers, with no HBM round trip.
//
// GEMM orientation: rows (M) = samples, columns (N) = beams, K = antennas x
// {re, im}.  Each lane of a 32x32 accumulator holds one beam column and 16
// sample rows, so the time average is an in-lane sum plus one cross-lane add.
// The complex structure sits on the voltage side: with B-operand K-slots
// [Wr, Wi], the real part uses A = [Dr, -Di] and the imaginary part
// A = [Di, Dr]; weights are stored once.
//
// Schedule.  One workgroup of 4 waves per CU (one wave per SIMD) walks a list
// of tiles.  Everything that would otherwise stall the MFMA pipe is overlapped
// by construction:
//   - the next tile's packed voltages are loaded into registers while the
//     current tile computes, then transposed into the other half of a
//     double-buffered LDS stage (2 x 80 KB);
//   - a tile is 32 GEMM units (16 (channel, pol) x two 32-sample row tiles);
//     unit u accumulates into one of two accumulator sets while the |B|^2
//     detect of unit u-1 reads the other;
//   - weights of the next (channel, pol) load into the second of two weight
//     register sets while the current one is in use.
#include "bf_common.h"

namespace {
constexpr int NW = 4, NT = NW * 64;       // waves, threads per workgroup
constexpr int BTW = 2;                    // 32-beam tiles per wave
constexpr int BSET = NW * BTW * 32;       // beams per tile
constexpr int TCU = 64;                   // samples per tile
constexpr int PCU = TCU / 2;              // packets per tile
constexpr int ROWB = 80;                  // LDS bytes per (chan, pol, sample) row
constexpr int STAGE_BYTES = 16 * TCU * ROWB;
constexpr int TASKS = PCU * 16 / NT;      // (packet, antenna-quad) tasks per thread
}

struct BfArgs {
  const uint8_t *in;
  const uint8_t *wfrag;
  const float *wscale;
  const float *scale;
  uint8_t *out;
  float *power;   // optional
  float *bpart;   // [nbeam][nsamp/64][ngroup] per-tile beam power sums
  int nant, nchan, ngroup, nbeam, nbeam_pad, nbt, wmode, order;
  int ntblk, ntiles, tiles_per_xcd;
  uint32_t pstride, astride, gstride;   // input bytes per packet, antenna, channel group
};

struct TileId {
  int tblk, g, bset;
};
// Time blocks are the fastest index so the tiles in flight on one XCD share a
// channel group's weights in its L2.
__device__ __forceinline__ TileId tile_id(const BfArgs &a, int L) {
  TileId t;
  t.tblk = L % a.ntblk;
  const int rest = L / a.ntblk;
  t.g = rest % a.ngroup;
  t.bset = rest / a.ngroup;
  return t;
}

// ---- staging ---------------------------------------------------------------
// The next tile's packed chunks are copied global -> LDS by the memory system
// (global_load_lds_dwordx4: no VGPRs, completion tracked by vmcnt) into a raw
// buffer laid out [packet][antenna][2 x 16 B]; at the tile boundary they are
// transposed LDS -> LDS into the [channel][pol][sample][antenna] stage.
constexpr int RAW_BYTES = PCU * 64 * 32;             // 64 antenna slots per packet
constexpr int RAW_SLOTS = RAW_BYTES / 16;            // 16-byte slots
constexpr int RAW_LOADS = RAW_SLOTS / NT;            // direct loads per thread

__device__ __forceinline__ void raw_load(uint8_t *raw, const BfArgs &a, const TileId &t, int wave,
                                         int lane) {
  // chunk offset = packet*pstride + antenna*astride + group*gstride (both
  // input orders), all below 4 GiB
  const uint8_t *tile = a.in + ((size_t)t.tblk * PCU * a.pstride + (size_t)t.g * a.gstride);
#pragma unroll
  for (int i = 0; i < RAW_LOADS; i++) {
    const int base = (wave * RAW_LOADS + i) * 64;     // first slot of this wave instruction
    const int slot = base + lane;
    const int pl = slot >> 7, ant = (slot >> 1) & 63, half = slot & 1;
    // antennas past nant read antenna nant-1 (in bounds) and are zeroed by
    // the transpose
    const uint32_t src_ant = (uint32_t)min(ant, a.nant - 1);
    const uint32_t off = (uint32_t)pl * a.pstride + src_ant * a.astride + (uint32_t)half * 16u;
    __builtin_amdgcn_global_load_lds((const void *)(tile + off),
                                     (__attribute__((address_space(3))) void *)(raw + base * 16), 16, 0, 0);
  }
}

// 4x4 byte transposes: antenna-major raw chunks -> one dword (4 antennas) per
// [channel][pol][sample] row; task = (packet, quad of 4 antennas).
__device__ __forceinline__ void raw_transpose(const uint8_t *raw, uint8_t *stage, int nant, int tid) {
#pragma unroll
  for (int k = 0; k < TASKS; k++) {
    const int task = tid + k * NT;
    const int pl = task >> 4, aq = task & 15;
    uint32_t v[4][8];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int ant = aq * 4 + j;
      const uint4 c0 = *(const uint4 *)(raw + ((pl * 64 + ant) * 2 + 0) * 16);
      const uint4 c1 = *(const uint4 *)(raw + ((pl * 64 + ant) * 2 + 1) * 16);
      const bool live = ant < nant;
      v[j][0] = live ? c0.x : 0u; v[j][1] = live ? c0.y : 0u;
      v[j][2] = live ? c0.z : 0u; v[j][3] = live ? c0.w : 0u;
      v[j][4] = live ? c1.x : 0u; v[j][5] = live ? c1.y : 0u;
      v[j][6] = live ? c1.z : 0u; v[j][7] = live ? c1.w : 0u;
    }
    uint8_t *base = stage + aq * 4;
#pragma unroll
    for (int ci = 0; ci < 8; ci++) {
      const uint32_t x0 = v[0][ci], x1 = v[1][ci], x2 = v[2][ci], x3 = v[3][ci];
      const uint32_t l01 = __builtin_amdgcn_perm(x1, x0, 0x05010400u);
      const uint32_t h01 = __builtin_amdgcn_perm(x1, x0, 0x07030602u);
      const uint32_t l23 = __builtin_amdgcn_perm(x3, x2, 0x05010400u);
      const uint32_t h23 = __builtin_amdgcn_perm(x3, x2, 0x07030602u);
      *(uint32_t *)(base + ((ci * 2 + 0) * TCU + 2 * pl) * ROWB) = __builtin_amdgcn_perm(l23, l01, 0x05040100u);
      *(uint32_t *)(base + ((ci * 2 + 1) * TCU + 2 * pl) * ROWB) = __builtin_amdgcn_perm(l23, l01, 0x07060302u);
      *(uint32_t *)(base + ((ci * 2 + 0) * TCU + 2 * pl + 1) * ROWB) = __builtin_amdgcn_perm(h23, h01, 0x05040100u);
      *(uint32_t *)(base + ((ci * 2 + 1) * TCU + 2 * pl + 1) * ROWB) = __builtin_amdgcn_perm(h23, h01, 0x07060302u);
    }
  }
}

// ---- weights -----------------------------------------------------------------
template <int PREC> struct WFrag;
template <> struct WFrag<0> {
  h8v w[BTW][8];
  float sc[BTW];
  __device__ __forceinline__ void load(const uint8_t *wfrag, const float *wscale, long cqrow,
                                       long scrow, int lane, int n) {
#pragma unroll
    for (int bt = 0; bt < BTW; bt++)
#pragma unroll
      for (int s = 0; s < 8; s++)
        w[bt][s] = *(const h8v *)(wfrag + (((cqrow + bt) * 8 + s) * 64 + lane) * 16);
#pragma unroll
    for (int bt = 0; bt < BTW; bt++) sc[bt] = wscale[scrow + bt * 32 + n];
  }
};
template <> struct WFrag<1> {
  i8v hi[BTW][2], lo[BTW][2];
  float sc[BTW];
  __device__ __forceinline__ void load(const uint8_t *wfrag, const float *wscale, long cqrow,
                                       long scrow, int lane, int n) {
#pragma unroll
    for (int bt = 0; bt < BTW; bt++)
#pragma unroll
      for (int s2 = 0; s2 < 2; s2++) {
        const long rec = (((cqrow + bt) * 2 + s2) * 2) * 64 + lane;
        hi[bt][s2] = *(const i8v *)(wfrag + rec * 32);
        lo[bt][s2] = *(const i8v *)(wfrag + (rec + 64) * 32);
      }
#pragma unroll
    for (int bt = 0; bt < BTW; bt++) sc[bt] = wscale[scrow + bt * 32 + n];
  }
};

// ---- one GEMM unit: 32 sample rows of one (channel, pol) -----------------------
struct Acc {
  f16v r[BTW], i[BTW];
};

template <int PREC>
__device__ __forceinline__ void gemm_unit(Acc &A, const WFrag<PREC> &W, const uint8_t *rows, int n,
                                          int h) {
  const uint8_t *src = rows + n * ROWB + h * 32;
  const uint4 d0 = *(const uint4 *)src, d1 = *(const uint4 *)(src + 16);
  const uint32_t dw[8] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
#pragma unroll
  for (int bt = 0; bt < BTW; bt++) { A.r[bt] = (f16v){}; A.i[bt] = (f16v){}; }
  if constexpr (PREC == 0) {
#pragma unroll
    for (int s = 0; s < 8; s++) {
      h8v are, aim;
      unpack_f16(dw[s], are, aim);
#pragma unroll
      for (int bt = 0; bt < BTW; bt++) {
        A.r[bt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(are, W.w[bt][s], A.r[bt], 0, 0, 0);
        A.i[bt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(aim, W.w[bt][s], A.i[bt], 0, 0, 0);
      }
    }
  } else {
#pragma unroll
    for (int s2 = 0; s2 < 2; s2++) {
      i8v are, aim;
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const uint32_t x = dw[4 * s2 + u];
        const uint32_t dr = nib_e4m3(x), di = nib_e4m3(x >> 4);
        are[2 * u] = (int)dr;
        are[2 * u + 1] = (int)(di ^ 0x80808080u);
        aim[2 * u] = (int)di;
        aim[2 * u + 1] = (int)dr;
      }
#pragma unroll
      for (int bt = 0; bt < BTW; bt++) {
        A.r[bt] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(are, W.hi[bt][s2], A.r[bt], 0, 0, 0, 0, 0, 0);
        A.i[bt] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(aim, W.hi[bt][s2], A.i[bt], 0, 0, 0, 0, 0, 0);
        A.r[bt] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(are, W.lo[bt][s2], A.r[bt], 0, 0, 0, 0, 0, 0);
        A.i[bt] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(aim, W.lo[bt][s2], A.i[bt], 0, 0, 0, 0, 0, 0);
      }
    }
  }
}

// |B|^2 summed over the accumulator rows of each bin (four independent FMA
// chains per bin), scaled into the partial power.
template <int TAVG>
__device__ __forceinline__ void detect(const Acc &A, const float (&sc)[BTW],
                                       float (&part)[BTW][TAVG <= 32 ? 32 / TAVG : 1]) {
  constexpr int NP = TAVG <= 32 ? 32 / TAVG : 1;
  constexpr int VPB = 16 / NP;
#pragma unroll
  for (int bt = 0; bt < BTW; bt++)
#pragma unroll
    for (int j = 0; j < NP; j++) {
      float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
#pragma unroll
      for (int v = j * VPB; v < (j + 1) * VPB; v += 2) {
        s0 = __builtin_fmaf(A.r[bt][v], A.r[bt][v], s0);
        s1 = __builtin_fmaf(A.i[bt][v], A.i[bt][v], s1);
        s2 = __builtin_fmaf(A.r[bt][v + 1], A.r[bt][v + 1], s2);
        s3 = __builtin_fmaf(A.i[bt][v + 1], A.i[bt][v + 1], s3);
      }
      part[bt][j] = __builtin_fmaf((s0 + s1) + (s2 + s3), sc[bt], part[bt][j]);
    }
}

// ---- the kernel -----------------------------------------------------------------
template <int PREC, int TAVG>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(1, 1)))
void bf_fused_kernel(BfArgs a) {
  __shared__ __attribute__((aligned(16))) uint8_t lds[STAGE_BYTES + RAW_BYTES];
  constexpr int NP = TAVG <= 32 ? 32 / TAVG : 1;   // bins per 32-row unit
  constexpr int NTT = TAVG == 64 ? 1 : 2;          // row units with their own bins

  // Persistent, XCD-aware: workgroup P runs on XCD P % 8 and walks that XCD's
  // contiguous share of the tile list with stride = workgroups per XCD.
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3, nslot = gridDim.x >> 3;
  const int tbeg = xcd * a.tiles_per_xcd;
  const int tend = min(tbeg + a.tiles_per_xcd, a.ntiles);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int n = lane & 31, h = lane >> 5;

  int L = tbeg + slot;
  if (L >= tend) return;
  uint8_t *const stage = lds;
  uint8_t *const raw = lds + STAGE_BYTES;
  raw_load(raw, a, tile_id(a, L), wave, lane);

  while (L < tend) {
    const TileId t = tile_id(a, L);
    const int Lnext = L + nslot;
    // the raw copy of this tile has landed (own loads, then the barrier for
    // everyone's); the stage was last read by the previous tile
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    raw_transpose(raw, stage, a.nant, tid);
    __syncthreads();
    if (Lnext < tend) raw_load(raw, a, tile_id(a, Lnext), wave, lane);   // lands during compute

    const int btw = t.bset * (BSET / 32) + wave * BTW;
    auto cqrow = [&](int cq) -> long {
      return (long)((a.wmode ? t.g * 8 + (cq >> 1) : t.g) * 2 + (cq & 1));
    };

    float part[NTT][BTW][NP];
#pragma unroll
    for (int i = 0; i < NTT; i++)
#pragma unroll
      for (int bt = 0; bt < BTW; bt++)
#pragma unroll
        for (int j = 0; j < NP; j++) part[i][bt][j] = 0.f;

    // 32 units u = 2*cq + tt: unit u fills accumulator set u&1 while the
    // detect of unit u-1 drains the other; weights of cq+1 load at unit 2cq.
    WFrag<PREC> W0, W1;
    Acc A0, A1;
    W0.load(a.wfrag, a.wscale, cqrow(0) * a.nbt + btw, cqrow(0) * a.nbeam_pad + btw * 32, lane, n);
    gemm_unit<PREC>(A0, W0, stage, n, h);                                   // unit 0
    // sched_barrier keeps each (gemm u, detect u-1) pair in its own scheduling
    // region so the scheduler interleaves within it instead of hoisting later
    // units (which would need more than the 512 registers a wave has).
    for (int cq = 0; cq < 16; cq += 2) {
      W1.load(a.wfrag, a.wscale, cqrow(cq + 1) * a.nbt + btw, cqrow(cq + 1) * a.nbeam_pad + btw * 32,
              lane, n);
      gemm_unit<PREC>(A1, W0, stage + (cq * TCU + 32) * ROWB, n, h);          // unit 2cq+1
      detect<TAVG>(A0, W0.sc, part[0]);                                        // unit 2cq
      __builtin_amdgcn_sched_barrier(0);
      gemm_unit<PREC>(A0, W1, stage + ((cq + 1) * TCU) * ROWB, n, h);         // unit 2cq+2
      detect<TAVG>(A1, W0.sc, part[NTT - 1]);                                  // unit 2cq+1
      __builtin_amdgcn_sched_barrier(0);
      if (cq + 2 < 16)
        W0.load(a.wfrag, a.wscale, cqrow(cq + 2) * a.nbt + btw,
                cqrow(cq + 2) * a.nbeam_pad + btw * 32, lane, n);
      gemm_unit<PREC>(A1, W1, stage + ((cq + 1) * TCU + 32) * ROWB, n, h);    // unit 2cq+3
      detect<TAVG>(A0, W1.sc, part[0]);                                        // unit 2cq+2
      __builtin_amdgcn_sched_barrier(0);
      if (cq + 2 < 16) gemm_unit<PREC>(A0, W0, stage + ((cq + 2) * TCU) * ROWB, n, h);  // 2cq+4
      detect<TAVG>(A1, W1.sc, part[NTT - 1]);                                  // unit 2cq+3
      __builtin_amdgcn_sched_barrier(0);
    }

    // lanes h=0 and h=1 hold the two row halves of each bin.  Besides the
    // output, the tile's sum per beam (badder in the reference) goes to bpart.
#pragma unroll
    for (int bt = 0; bt < BTW; bt++) {
      const int b = (btw + bt) * 32 + n;
      float tile_sum = 0.f;
#pragma unroll
      for (int tt = 0; tt < NTT; tt++)
#pragma unroll
        for (int j = 0; j < NP; j++) {
          float x = part[tt][bt][j];
          x += __shfl_xor(x, 32, 64);
          tile_sum += x;
          if (h == 0 && b < a.nbeam) {
            const int o = TAVG == 64 ? t.tblk : t.tblk * (TCU / TAVG) + tt * (32 / TAVG) + j;
            const long idx = ((long)o * a.nbeam + b) * a.ngroup + t.g;
            if (a.power) a.power[idx] = x;
            a.out[idx] = requant_u8(x, a.scale[b]);
          }
        }
      if (h == 0 && b < a.nbeam) a.bpart[((long)b * a.ntblk + t.tblk) * a.ngroup + t.g] = tile_sum;
    }

    L = Lnext;
  }
}

template <int PREC>
static void launch_fused(int tavg, dim3 grid, hipStream_t s, const BfArgs &a) {
  switch (tavg) {
    case 8:  hipLaunchKernelGGL((bf_fused_kernel<PREC, 8>), grid, dim3(NT), 0, s, a); break;
    case 16: hipLaunchKernelGGL((bf_fused_kernel<PREC, 16>), grid, dim3(NT), 0, s, a); break;
    case 32: hipLaunchKernelGGL((bf_fused_kernel<PREC, 32>), grid, dim3(NT), 0, s, a); break;
    default: hipLaunchKernelGGL((bf_fused_kernel<PREC, 64>), grid, dim3(NT), 0, s, a); break;
  }
}

extern "C" int bfk_beamform(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag,
                            const float *d_wscale, const float *d_scale, uint8_t *d_out,
                            float *d_power, float *d_bpart, void *stream) {
  if (d->nant_pad != 64 || d->nbeam_pad % BSET != 0 || d->nsamp % TCU != 0) {
    bfk_set_error("bfk_beamform: unsupported shape nant_pad=%d nbeam_pad=%d nsamp=%d", d->nant_pad,
                  d->nbeam_pad, d->nsamp);
    return -1;
  }
  static thread_local int ncu = 0;
  if (ncu == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 8)
      ncu = 256;
  }
  BfArgs a;
  a.in = d_in;
  a.wfrag = (const uint8_t *)d_wfrag;
  a.wscale = d_wscale;
  a.scale = d_scale;
  a.out = d_out;
  a.power = d_power;
  a.bpart = d_bpart;
  a.nant = d->nant;
  a.nchan = d->nchan;
  a.ngroup = d->ngroup;
  a.nbeam = d->nbeam;
  a.nbeam_pad = d->nbeam_pad;
  a.nbt = d->nbt;
  a.wmode = d->wmode;
  a.order = d->order;
  a.ntblk = d->nsamp / TCU;
  a.pstride = (uint32_t)d->nant * d->nchan * 4;
  a.astride = d->order == 0 ? (uint32_t)d->nchan * 4 : 32u;
  a.gstride = d->order == 0 ? 32u : (uint32_t)d->nant * 32;
  a.ntiles = (d->nbeam_pad / BSET) * d->ngroup * a.ntblk;
  a.tiles_per_xcd = (a.ntiles + 7) / 8;
  // one workgroup per CU (160 KB of LDS each), spread evenly over 8 XCDs
  int per_xcd = ncu / 8;
  if (per_xcd > a.tiles_per_xcd) per_xcd = a.tiles_per_xcd;
  const dim3 grid(8 * per_xcd);
  hipStream_t s = (hipStream_t)stream;
  if (d->prec == 0) launch_fused<0>(d->tavg, grid, s, a);
  else launch_fused<1>(d->tavg, grid, s, a);
  return bfk_check(hipGetLastError(), "bf_fused_kernel");
}

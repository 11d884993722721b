// bf_shared8.hip -- the shared-unpack schedule for 128 and 256 antennas: one
// 8-wave workgroup per CU unpacks each 32-sample unit once into LDS MFMA
// operands that all eight waves (one 32-beam tile each, 256 beams per tile)
// read.
//
// Reference path: src/dsaX_beamformer.cu promoter (:170-193), beamformer
// (:267-416), badder (:78-112) and adder (:120-158), fused into one launch.
//
// Why.  With more than 64 antennas a wave's weights for all K chunks hold
// 64 (128 antennas) or 128 (256 antennas) registers, so a wave owns a single
// 32-beam tile, and in the classic schedule every unpacked voltage operand
// then feeds only two MFMAs: the unpack VALU work exceeds what the matrix
// pipe hides.  Here the unpack is done once per tile for eight waves, and
// each wave only reads A_re from LDS (one 16-byte read per K step) and
// derives A_im from it, about 2.6 (FP16) and 3 (FP8) VALU instructions per
// MFMA.
//
// Layout.  The group's 8 channels are staged in NKC sub-steps of 8/NKC
// channels (64 KB: [(channel, pol)][sample][antenna] rows of 64*NKC bytes
// whose 16-byte chunks are XOR-swizzled by row, so the unpack reads are
// conflict-free).  Where registers allow (128 antennas, FP8), the next
// sub-step's voltages are loaded while the current one is computed.  A_re operands: FP16 per unit, FP8
// per (channel, pol) (both units), double-buffered.  The MFMA chains and the
// detect follow gemm_unit / detect with one beam tile per wave, so results are
// bit-identical to the classic 128- and 64-beam tile schedules.
#include "bf_common.h"

namespace {
constexpr int NW8 = 8, NT8 = NW8 * 64;     // waves, threads per workgroup
constexpr int BSET8 = NW8 * 32;            // beams per tile
constexpr int STAGE8 = 64 * 1024;          // one channel sub-step, any NKC

template <int NKC> struct S8 {
  static constexpr int NCI = 8 / NKC;           // channels per sub-step
  static constexpr int CQS = 2 * NCI;           // (channel, pol) per sub-step
  static constexpr int ROW = 64 * NKC;          // staged row bytes (antennas)
  static constexpr int NCH = ROW / 16;          // 16-byte chunks per row
  static constexpr int TPT = PCU * 16 * NKC / NT8;   // staging tasks per thread (= NKC)
  static constexpr int UNIT16 = NKC * 8 * 64 * 16;    // FP16 A_re bytes per unit
  static constexpr int UNIT8 = NKC * 2 * 64 * 32;     // FP8 A_re bytes per unit
};

// Byte offset of 16-byte chunk c of staged row r (chunks XORed with the row).
template <int NKC>
__device__ __forceinline__ int srow8(int r, int c) {
  return r * S8<NKC>::ROW + ((c ^ (r & (S8<NKC>::NCH - 1))) << 4);
}
}  // namespace

// One sub-step's voltages in flight: per task (packet, antenna quad) the NCI
// channel dwords of its 4 antennas.
template <int NKC> struct Sub8Regs {
  uint32_t v[S8<NKC>::TPT][4][S8<NKC>::NCI];
};

template <int NKC>
__device__ __forceinline__ void sub8_load(Sub8Regs<NKC> &R, const BfArgs &a, __amdgpu_buffer_rsrc_t rin,
                                          const TileId &t, int sub, int tid) {
  using S = S8<NKC>;
  const uint32_t astep = a.order == 0 ? (uint32_t)a.nchan * 4u : 32u;   // bytes between antennas
#pragma unroll
  for (int k = 0; k < S::TPT; k++) {
    const int task = tid + k * NT8, pl = task / (16 * NKC), aq = task % (16 * NKC);
    const long p = (long)t.tblk * PCU + pl;
    const uint32_t base =
        (uint32_t)chunk_offset(a.order, p, aq * 4, t.g, a.nant, a.nchan, a.ngroup) + sub * S::NCI * 4;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const bool in = aq * 4 + j < a.nant;
      if constexpr (S::NCI == 2) {
        typedef uint32_t u2 __attribute__((ext_vector_type(2)));
        const u2 x = __builtin_amdgcn_raw_buffer_load_b64(rin, base, j * astep, 0);
        R.v[k][j][0] = in ? x[0] : 0u;
        R.v[k][j][1] = in ? x[1] : 0u;
      } else {
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
        const u4 x = __builtin_amdgcn_raw_buffer_load_b128(rin, base, j * astep, 0);
#pragma unroll
        for (int c = 0; c < 4; c++) R.v[k][j][c] = in ? x[c] : 0u;
      }
    }
  }
}

// 4x4 byte transposes: the channel dwords (2 times x 2 pols) of 4 antennas
// -> one dword (4 antennas) in each of the 4 rows they feed.
template <int NKC>
__device__ __forceinline__ void sub8_store(uint8_t *stage, const Sub8Regs<NKC> &R, int tid) {
  using S = S8<NKC>;
#pragma unroll
  for (int k = 0; k < S::TPT; k++) {
    const int task = tid + k * NT8, pl = task / (16 * NKC), aq = task % (16 * NKC);
    const int c = aq >> 2, w = (aq & 3) * 4;
#pragma unroll
    for (int ci = 0; ci < S::NCI; ci++) {
      const uint32_t x0 = R.v[k][0][ci], x1 = R.v[k][1][ci], x2 = R.v[k][2][ci], x3 = R.v[k][3][ci];
      const uint32_t l01 = __builtin_amdgcn_perm(x1, x0, 0x05010400u);
      const uint32_t h01 = __builtin_amdgcn_perm(x1, x0, 0x07030602u);
      const uint32_t l23 = __builtin_amdgcn_perm(x3, x2, 0x05010400u);
      const uint32_t h23 = __builtin_amdgcn_perm(x3, x2, 0x07030602u);
      const int r0 = (ci * 2 + 0) * TCU + 2 * pl, r1 = (ci * 2 + 1) * TCU + 2 * pl;
      *(uint32_t *)(stage + srow8<NKC>(r0, c) + w) = __builtin_amdgcn_perm(l23, l01, 0x05040100u);
      *(uint32_t *)(stage + srow8<NKC>(r1, c) + w) = __builtin_amdgcn_perm(l23, l01, 0x07060302u);
      *(uint32_t *)(stage + srow8<NKC>(r0 + 1, c) + w) = __builtin_amdgcn_perm(h23, h01, 0x05040100u);
      *(uint32_t *)(stage + srow8<NKC>(r1 + 1, c) + w) = __builtin_amdgcn_perm(h23, h01, 0x07060302u);
    }
  }
}

// FP16 A_re of unit u (= 2*cql + tt) of the sub-step: K step S = kc*8 + s of
// lane l at (S*64 + l)*16.  Thread t covers lane t & 63 and the NKC steps
// (t >> 6)*NKC ...: 4*NKC packed bytes of one row, read in one piece.
template <int NKC>
__device__ __forceinline__ void unpack8_f16(const uint8_t *stage, uint8_t *abuf, int u, int tid) {
  const int li = tid & 63, n = li & 31, h = li >> 5, q = tid >> 6;
  const int r = (u >> 1) * TCU + (u & 1) * 32 + n;
  const int S0 = q * NKC, kc = S0 >> 3, s0 = S0 & 7;
  const int byte = kc * 64 + h * 32 + 4 * s0;           // first antenna of the piece
  const uint8_t *src = stage + srow8<NKC>(r, byte >> 4) + (byte & 15);
  uint32_t x[NKC];
  if constexpr (NKC == 4) {
    const uint4 v = *(const uint4 *)src;
    x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
  } else {
    const uint2 v = *(const uint2 *)src;
    x[0] = v.x; x[1] = v.y;
  }
#pragma unroll
  for (int i = 0; i < NKC; i++) {
    h8v re, im;
    unpack_f16(x[i], re, im);
    *(h8v *)(abuf + ((S0 + i) * 64 + li) * 16) = re;
  }
}

// FP8 A_re of both units of sub-step (channel, pol) cql: unit tt, chunk kc,
// step s2, lane l: record C = (tt*NKC + kc)*2 + s2, its 32 bytes in two
// 16-byte planes, half hl at ((C*2 + hl)*64 + l)*16 (16-byte lane stride:
// conflict-free).  Thread t covers lane t & 63 and NKC/2 of the records.
template <int NKC>
__device__ __forceinline__ void unpack8_f8(const uint8_t *stage, uint8_t *abuf, int cql, int tid) {
  const int li = tid & 63, n = li & 31, h = li >> 5, q = tid >> 6;
#pragma unroll
  for (int i = 0; i < NKC / 2; i++) {
    const int C = q * (NKC / 2) + i, tt = C / (2 * NKC), kc = (C >> 1) % NKC, s2 = C & 1;
    const int r = cql * TCU + tt * 32 + n;
    const uint4 v = *(const uint4 *)(stage + srow8<NKC>(r, kc * 4 + h * 2 + s2));
    const uint32_t x[4] = {v.x, v.y, v.z, v.w};
    uint32_t o[8];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      o[2 * k] = nib_e4m3(x[k]);
      o[2 * k + 1] = nib_e4m3(x[k] >> 4) ^ 0x80808080u;
    }
    uint8_t *p = abuf + ((C * 2) * 64 + li) * 16;
    *(uint4 *)p = make_uint4(o[0], o[1], o[2], o[3]);
    *(uint4 *)(p + 64 * 16) = make_uint4(o[4], o[5], o[6], o[7]);
  }
}

// FP6 A_re, the record layout of unpack8_f8: six dwords per lane, dwords 0..3
// in plane 0 and 4..5 in the first half of plane 1's 16-byte lane slot.
template <int NKC>
__device__ __forceinline__ void unpack8_f6(const uint8_t *stage, uint8_t *abuf, int cql, int tid) {
  const int li = tid & 63, n = li & 31, h = li >> 5, q = tid >> 6;
#pragma unroll
  for (int i = 0; i < NKC / 2; i++) {
    const int C = q * (NKC / 2) + i, tt = C / (2 * NKC), kc = (C >> 1) % NKC, s2 = C & 1;
    const int r = cql * TCU + tt * 32 + n;
    const uint4 v = *(const uint4 *)(stage + srow8<NKC>(r, kc * 4 + h * 2 + s2));
    const uint32_t x[4] = {v.x, v.y, v.z, v.w};
    uint32_t o[6];
    f6_operand(x, o);
    uint8_t *p = abuf + ((C * 2) * 64 + li) * 16;
    *(uint4 *)p = make_uint4(o[0], o[1], o[2], o[3]);
    *(uint2 *)(p + 64 * 16) = make_uint2(o[4], o[5]);
  }
}

// One unit from LDS A_re: gemm_unit's MFMA chains (one beam tile), A_im
// derived.  With PF, each K step of W is reloaded from row wnext right after
// its last MFMA has issued (rolling prefetch, pinned by scheduling barriers
// that VALU, SALU and LDS work may still cross).
template <int PREC, int NKC, bool PF>
__device__ __forceinline__ void unit8w(Acc<1> &A, WFrag<PREC, 1, NKC> &W, const uint8_t *unit, int lane,
                                       const uint8_t *wfrag, long wnext) {
  __builtin_amdgcn_s_setprio(1);
  A.r[0] = (f16v){};
  A.i[0] = (f16v){};
#pragma unroll
  for (int kc = 0; kc < NKC; kc++) {
    if constexpr (PREC == 0) {
#pragma unroll
      for (int s = 0; s < 8; s++) {
        const uint4 r = *(const uint4 *)(unit + ((kc * 8 + s) * 64 + lane) * 16);
        // A_re = [Dr0,Dr1,-Di0,-Di1,Dr2,Dr3,-Di2,-Di3] -> A_im = [Di0,Di1,Dr0,Dr1,...]
        const uint4 m = make_uint4(r.y ^ 0x80008000u, r.x, r.w ^ 0x80008000u, r.z);
        A.r[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8v, r), W.w[0][kc][s], A.r[0], 0, 0, 0);
        A.i[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8v, m), W.w[0][kc][s], A.i[0], 0, 0, 0);
        if constexpr (PF) {
          __builtin_amdgcn_sched_barrier(0x386);
          W.load_k(wfrag, wnext, 0, kc, s, lane);
          __builtin_amdgcn_sched_barrier(0x386);
        }
      }
    } else {
#pragma unroll
      for (int s2 = 0; s2 < 2; s2++) {
        const uint8_t *src = unit + (((kc * 2 + s2) * 2) * 64 + lane) * 16;
        i8v are, aim;
        if constexpr (PREC == 1) {
          const uint4 r0 = *(const uint4 *)src, r1 = *(const uint4 *)(src + 64 * 16);
          // A_re dwords [Dr, -Di] per 4 antennas -> A_im [Di, Dr]
          are = (i8v){(int)r0.x, (int)r0.y, (int)r0.z, (int)r0.w, (int)r1.x, (int)r1.y, (int)r1.z, (int)r1.w};
          aim = (i8v){(int)(r0.y ^ 0x80808080u), (int)r0.x, (int)(r0.w ^ 0x80808080u), (int)r0.z,
                      (int)(r1.y ^ 0x80808080u), (int)r1.x, (int)(r1.w ^ 0x80808080u), (int)r1.z};
        } else {
          f6_re_im(*(const uint4 *)src, *(const uint2 *)(src + 64 * 16), are, aim);
        }
        A.r[0] = mfma64<PREC>(are, W.hi[0][kc][s2], A.r[0], W.shi(0, kc, s2));
        A.i[0] = mfma64<PREC>(aim, W.hi[0][kc][s2], A.i[0], W.shi(0, kc, s2));
        if constexpr (PREC == 3) {
          A.r[0] = mfma64<PREC>(are, W.mid[0][kc][s2], A.r[0], W.smid(0, kc, s2));
          A.i[0] = mfma64<PREC>(aim, W.mid[0][kc][s2], A.i[0], W.smid(0, kc, s2));
        }
        A.r[0] = mfma64<PREC>(are, W.lo[0][kc][s2], A.r[0], W.slo(0, kc, s2));
        A.i[0] = mfma64<PREC>(aim, W.lo[0][kc][s2], A.i[0], W.slo(0, kc, s2));
        if constexpr (PF) {
          __builtin_amdgcn_sched_barrier(0x386);
          W.load_k(wfrag, wnext, 0, kc, s2, lane);
          __builtin_amdgcn_sched_barrier(0x386);
        }
      }
    }
  }
  __builtin_amdgcn_s_setprio(0);
}

template <int PREC, int NKC> constexpr int abuf8_bytes() {
  return PREC == 0 ? 2 * S8<NKC>::UNIT16 : 4 * S8<NKC>::UNIT8;   // double buffers
}

template <int PREC, int TAVG, int NKC>
__device__ __forceinline__ void shared8_tile(const BfArgs &a, int L, uint8_t *stage, uint8_t *abuf) {
  using S = S8<NKC>;
  constexpr int NP = TAVG <= 32 ? 32 / TAVG : 1;
  const TileId t = tile_id(a, L);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int n = lane & 31, h = lane >> 5;
  const int btw = t.bset * NW8 + wave;       // this wave's 32-beam tile
  // Next sub-step's voltages held in registers during the current one: only
  // where that fits beside the weights (128 antennas, FP8 and FP6); elsewhere
  // each sub-step loads its own after the previous one.
  constexpr bool PFS = (PREC == 1 || PREC == 2) && NKC == 2;

  // sub-step 0's voltages and the first weights; each later sub-step's
  // voltages are loaded while the previous one is computed
  const __amdgpu_buffer_rsrc_t rin = buffer_rsrc(a.in, a.in_bytes);
  Sub8Regs<NKC> R;
  sub8_load<NKC>(R, a, rin, t, 0, tid);
  WFrag<PREC, 1, NKC> W;
  {
    const long wr = weight_row(a, t.g, 0);
    W.load(a.wfrag, a.wscale, wr * a.nbt + btw, wr * a.nbeam_pad + btw * 32, lane, n);
  }
  sub8_store<NKC>(stage, R, tid);
  if (PFS) sub8_load<NKC>(R, a, rin, t, 1, tid);
  __syncthreads();

  float part[2][1][NP];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < NP; j++) part[i][0][j] = 0.f;

  for (int sub = 0; sub < NKC; sub++) {
    // this sub-step's stage is in LDS; its first A_re operands:
    if constexpr (PREC == 0) unpack8_f16<NKC>(stage, abuf, 0, tid);
    else if constexpr (PREC == 1) unpack8_f8<NKC>(stage, abuf, 0, tid);
    else unpack8_f6<NKC>(stage, abuf, 0, tid);
    __syncthreads();
    for (int cql = 0; cql < S::CQS; cql++) {
      const int cq = sub * S::CQS + cql;
      const long wn = weight_row(a, t.g, cq < 15 ? cq + 1 : cq);
      const bool last = cql == S::CQS - 1;
      if constexpr (PREC == 0) {
        Acc<1> A;
        unpack8_f16<NKC>(stage, abuf + S::UNIT16, 2 * cql + 1, tid);
        unit8w<PREC, NKC, false>(A, W, abuf, lane, a.wfrag, 0);
        detect<TAVG>(A, W.sc, part[0]);
        __syncthreads();
        if (!last) unpack8_f16<NKC>(stage, abuf, 2 * cql + 2, tid);
        unit8w<PREC, NKC, true>(A, W, abuf + S::UNIT16, lane, a.wfrag, wn * a.nbt + btw);
        detect<TAVG>(A, W.sc, part[1]);
      } else {
        const uint8_t *ab = abuf + (cql & 1) * 2 * S::UNIT8;
        if (!last) {
          uint8_t *nb = abuf + ((cql + 1) & 1) * 2 * S::UNIT8;
          if constexpr (PREC == 1) unpack8_f8<NKC>(stage, nb, cql + 1, tid);
          else unpack8_f6<NKC>(stage, nb, cql + 1, tid);
        }
        Acc<1> A;
        unit8w<PREC, NKC, false>(A, W, ab, lane, a.wfrag, 0);
        detect<TAVG>(A, W.sc, part[0]);
        unit8w<PREC, NKC, true>(A, W, ab + S::UNIT8, lane, a.wfrag, wn * a.nbt + btw);
        detect<TAVG>(A, W.sc, part[1]);
      }
      W.load_scale(a.wscale, wn * a.nbeam_pad + btw * 32, n);
      __syncthreads();
    }
    // the next sub-step: its voltages (in registers) replace the stage, whose
    // last reader finished before the barrier above; then its loads go out
    if (sub + 1 < NKC) {
      if (!PFS) sub8_load<NKC>(R, a, rin, t, sub + 1, tid);
      sub8_store<NKC>(stage, R, tid);
      if (PFS && sub + 2 < NKC) sub8_load<NKC>(R, a, rin, t, sub + 2, tid);
      __syncthreads();
    }
  }

  tile_output<TAVG>(a, t, btw, n, h, part, 0, 2);
}

template <int PREC, int TAVG, int NKC>
__global__ __launch_bounds__(NT8) __attribute__((amdgpu_waves_per_eu(2, 2)))
void bf_shared8_kernel(BfArgs a) {
  __shared__ __attribute__((aligned(16))) uint8_t stage[STAGE8];
  __shared__ __attribute__((aligned(16))) uint8_t abuf[abuf8_bytes<PREC, NKC>()];
  // XCD-aware order as in the classic schedule (tile_id explains it)
  const int P = blockIdx.x;
  const int L = (P & 7) * (gridDim.x >> 3) + (P >> 3);
  if (L >= a.ntiles) return;
  shared8_tile<PREC, TAVG, NKC>(a, L, stage, abuf);
}

// Batched: the tiles of njobs blocks back to back; job j owns global tiles
// [tile_start[j], tile_start[j+1]).
template <int PREC, int TAVG, int NKC>
__global__ __launch_bounds__(NT8) __attribute__((amdgpu_waves_per_eu(2, 2)))
void bf_shared8_batch_kernel(const BfArgs *__restrict__ jobs, const int *__restrict__ tile_start, int njobs) {
  __shared__ __attribute__((aligned(16))) uint8_t stage[STAGE8];
  __shared__ __attribute__((aligned(16))) uint8_t abuf[abuf8_bytes<PREC, NKC>()];
  const int P = blockIdx.x;
  const int L = (P & 7) * (gridDim.x >> 3) + (P >> 3);
  if (L >= tile_start[njobs]) return;
  int j = 0;
  while (L >= tile_start[j + 1]) j++;
  shared8_tile<PREC, TAVG, NKC>(jobs[j], L - tile_start[j], stage, abuf);
}

// ---- host side --------------------------------------------------------------

template <int PREC, int NKC>
static void launch8(int tavg, dim3 grid, hipStream_t s, const BfArgs &a) {
  switch (tavg) {
    case 8:  hipLaunchKernelGGL((bf_shared8_kernel<PREC, 8, NKC>), grid, dim3(NT8), 0, s, a); break;
    case 16: hipLaunchKernelGGL((bf_shared8_kernel<PREC, 16, NKC>), grid, dim3(NT8), 0, s, a); break;
    case 32: hipLaunchKernelGGL((bf_shared8_kernel<PREC, 32, NKC>), grid, dim3(NT8), 0, s, a); break;
    default: hipLaunchKernelGGL((bf_shared8_kernel<PREC, 64, NKC>), grid, dim3(NT8), 0, s, a); break;
  }
}

template <int PREC, int NKC>
static void launch8_batch(int tavg, dim3 grid, hipStream_t s, const BfArgs *jobs, const int *ts, int nj) {
  switch (tavg) {
    case 8:  hipLaunchKernelGGL((bf_shared8_batch_kernel<PREC, 8, NKC>), grid, dim3(NT8), 0, s, jobs, ts, nj); break;
    case 16: hipLaunchKernelGGL((bf_shared8_batch_kernel<PREC, 16, NKC>), grid, dim3(NT8), 0, s, jobs, ts, nj); break;
    case 32: hipLaunchKernelGGL((bf_shared8_batch_kernel<PREC, 32, NKC>), grid, dim3(NT8), 0, s, jobs, ts, nj); break;
    default: hipLaunchKernelGGL((bf_shared8_batch_kernel<PREC, 64, NKC>), grid, dim3(NT8), 0, s, jobs, ts, nj); break;
  }
}

static int shared8_shape_ok(const bfk_dims *d) {
  if ((d->nkc != 2 && d->nkc != 4) || d->nant_pad != 64 * d->nkc || d->nsamp % TCU != 0 ||
      d->nbeam_pad % BSET8 != 0) {
    bfk_set_error("bfk_beamform: 8-wave shared schedule needs 65-256 antennas (nant_pad=%d)", d->nant_pad);
    return 0;
  }
  if (d->prec == 3 && d->nkc != 2) {
    // three weight parts of four antenna chunks would take 192 registers
    bfk_set_error("bfk_beamform: the three-part FP6 split runs up to 128 antennas (nant_pad=%d)",
                  d->nant_pad);
    return 0;
  }
  return 1;
}

extern "C" int bfk_beamform_shared8(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag,
                                    const float *d_wscale, const float *d_scale, uint8_t *d_out,
                                    float *d_power, float *d_bpart, void *stream) {
  if (!shared8_shape_ok(d)) return -1;
  const BfArgs a = bf_make_args(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart, BSET8);
  const dim3 grid((a.ntiles + 7) / 8 * 8);
  hipStream_t s = (hipStream_t)stream;
  if (d->nkc == 2) {
    if (d->prec == 0) launch8<0, 2>(d->tavg, grid, s, a);
    else if (d->prec == 1) launch8<1, 2>(d->tavg, grid, s, a);
    else if (d->prec == 2) launch8<2, 2>(d->tavg, grid, s, a);
    else launch8<3, 2>(d->tavg, grid, s, a);
  } else {
    if (d->prec == 0) launch8<0, 4>(d->tavg, grid, s, a);
    else if (d->prec == 1) launch8<1, 4>(d->tavg, grid, s, a);
    else launch8<2, 4>(d->tavg, grid, s, a);
  }
  return bfk_check(hipGetLastError(), "bf_shared8_kernel");
}

extern "C" int bfk_shared8_args(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag, const float *d_wscale,
                                const float *d_scale, uint8_t *d_out, float *d_power, float *d_bpart,
                                BfArgs *a) {
  if (!shared8_shape_ok(d)) return -1;
  *a = bf_make_args(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart, BSET8);
  return 0;
}

extern "C" int bfk_beamform_shared8_batch(int prec, int tavg, int nkc, const void *d_jobs,
                                          const int *d_tile_start, int njobs, int ntiles, void *stream) {
  const dim3 grid((ntiles + 7) / 8 * 8);
  hipStream_t s = (hipStream_t)stream;
  const BfArgs *jobs = (const BfArgs *)d_jobs;
  if (nkc == 2) {
    if (prec == 0) launch8_batch<0, 2>(tavg, grid, s, jobs, d_tile_start, njobs);
    else if (prec == 1) launch8_batch<1, 2>(tavg, grid, s, jobs, d_tile_start, njobs);
    else if (prec == 2) launch8_batch<2, 2>(tavg, grid, s, jobs, d_tile_start, njobs);
    else launch8_batch<3, 2>(tavg, grid, s, jobs, d_tile_start, njobs);
  } else {
    if (prec == 0) launch8_batch<0, 4>(tavg, grid, s, jobs, d_tile_start, njobs);
    else if (prec == 1) launch8_batch<1, 4>(tavg, grid, s, jobs, d_tile_start, njobs);
    else launch8_batch<2, 4>(tavg, grid, s, jobs, d_tile_start, njobs);
  }
  return bfk_check(hipGetLastError(), "bf_shared8_batch_kernel");
}

// bf_shared_tile.h -- one tile of the "shared" schedule of the fused
// beamformer (the schedule is described in bf_shared.hip), as a device
// function shared by the kernels there and the mixed batch kernel
// (bf_mixed.hip).
#ifndef BF_SHARED_TILE_H
#define BF_SHARED_TILE_H
#include <type_traits>

#include "bf_common.h"

namespace shared {
constexpr int NW = 4, NT = NW * 64;        // waves, threads per workgroup
constexpr int BSET = NW * BTW * 32;        // beams per tile
constexpr int SROW = 64;                   // packed stage row bytes (swizzled)
constexpr int SSTAGE = 16 * TCU * SROW;    // 64 KB
constexpr int AB16 = 2 * 8 * 64 * 16;      // FP16 A_re per (channel, pol): 16 KB
constexpr int AB8 = 2 * 2 * 64 * 32;       // FP8 A_re per (channel, pol): 8 KB

// Byte offset of 8-byte chunk c (0..7) of staged row r: chunks are XORed with
// bits 2..4 of the row so that the 32 rows a unpack instruction reads fall on
// distinct banks.
__device__ __forceinline__ int srow(int r, int c) { return r * SROW + ((c ^ ((r >> 2) & 7)) << 3); }

// 4x4 transposes of 4 antennas' 32-byte chunks of packet pl into the 32 rows
// they feed (swizzled stage; dword aq*4 of a row is in chunk aq >> 1).
__device__ __forceinline__ void shared_stage_store(uint8_t *stage, const uint32_t (&v)[4][8], int pl, int aq) {
  const int c = aq >> 1, w = (aq & 1) * 4;
#pragma unroll
  for (int ci = 0; ci < 8; ci++) {
    const uint32_t x0 = v[0][ci], x1 = v[1][ci], x2 = v[2][ci], x3 = v[3][ci];
    const uint32_t l01 = __builtin_amdgcn_perm(x1, x0, 0x05010400u);
    const uint32_t h01 = __builtin_amdgcn_perm(x1, x0, 0x07030602u);
    const uint32_t l23 = __builtin_amdgcn_perm(x3, x2, 0x05010400u);
    const uint32_t h23 = __builtin_amdgcn_perm(x3, x2, 0x07030602u);
    const int r0 = (ci * 2 + 0) * TCU + 2 * pl, r1 = (ci * 2 + 1) * TCU + 2 * pl;
    *(uint32_t *)(stage + srow(r0, c) + w) = __builtin_amdgcn_perm(l23, l01, 0x05040100u);
    *(uint32_t *)(stage + srow(r1, c) + w) = __builtin_amdgcn_perm(l23, l01, 0x07060302u);
    *(uint32_t *)(stage + srow(r0 + 1, c) + w) = __builtin_amdgcn_perm(h23, h01, 0x05040100u);
    *(uint32_t *)(stage + srow(r1 + 1, c) + w) = __builtin_amdgcn_perm(h23, h01, 0x07060302u);
  }
}

// A_re operands into LDS.
//  FP16, one unit u = 2*cq + tt at a time: step s, lane l at (s*64 + l)*16
//        (8 halves); thread t covers lane t & 63 and steps 2*(t>>6), +1.
//  FP8, both units of (channel, pol) cq: unit tt, step s2, lane l: the 32
//        bytes in two 16-byte planes, half hl at (((tt*2 + s2)*2 + hl)*64 + l)*16;
//        thread t covers lane t & 63 of unit (t>>6)&1 and the 16 antennas of
//        step t>>7.
// Lanes are the fastest index with a 16-byte stride, so the 16-byte writes
// and reads are contiguous and free of bank conflicts.
__device__ __forceinline__ void unpack_unit_f16(const uint8_t *stage, uint8_t *abuf, int u, int tid) {
  const int li = tid & 63, n = li & 31, h = li >> 5, sp = tid >> 6;
  const int r = (u >> 1) * TCU + (u & 1) * 32 + n;
  const uint2 x = *(const uint2 *)(stage + srow(r, h * 4 + sp));
  h8v re0, im0, re1, im1;
  unpack_f16(x.x, re0, im0);
  unpack_f16(x.y, re1, im1);
  *(h8v *)(abuf + ((2 * sp) * 64 + li) * 16) = re0;
  *(h8v *)(abuf + ((2 * sp + 1) * 64 + li) * 16) = re1;
}

__device__ __forceinline__ void unpack_cq_f8(const uint8_t *stage, uint8_t *abuf, int cq, int tid) {
  const int li = tid & 63, n = li & 31, h = li >> 5;
  const int tt = (tid >> 6) & 1, q = tid >> 7;
  const int r = cq * TCU + tt * 32 + n;
  const uint2 x0 = *(const uint2 *)(stage + srow(r, h * 4 + 2 * q));
  const uint2 x1 = *(const uint2 *)(stage + srow(r, h * 4 + 2 * q + 1));
  const uint32_t x[4] = {x0.x, x0.y, x1.x, x1.y};
  uint32_t o[8];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    o[2 * k] = nib_e4m3(x[k]);
    o[2 * k + 1] = nib_e4m3(x[k] >> 4) ^ 0x80808080u;
  }
  uint8_t *p = abuf + (((tt * 2 + q) * 2) * 64 + li) * 16;
  *(uint4 *)p = make_uint4(o[0], o[1], o[2], o[3]);
  *(uint4 *)(p + 64 * 16) = make_uint4(o[4], o[5], o[6], o[7]);
}

// FP6, both units of (channel, pol) cq, the same thread mapping as FP8: the
// 16 antennas of a lane's step are slots 0..15 (Dr) and 16..31 (-Di), six
// dwords per lane in two planes, unit tt and step s2 at (tt*4 + s2*1.5) KB:
// dwords 0..3 at lane*16, dwords 4..5 at 1024 + lane*8 (both conflict-free).
__device__ __forceinline__ void unpack_cq_f6(const uint8_t *stage, uint8_t *abuf, int cq, int tid) {
  const int li = tid & 63, n = li & 31, h = li >> 5;
  const int tt = (tid >> 6) & 1, q = tid >> 7;
  const int r = cq * TCU + tt * 32 + n;
  const uint2 x0 = *(const uint2 *)(stage + srow(r, h * 4 + 2 * q));
  const uint2 x1 = *(const uint2 *)(stage + srow(r, h * 4 + 2 * q + 1));
  const uint32_t x[4] = {x0.x, x0.y, x1.x, x1.y};
  uint32_t o[6];
  f6_operand(x, o);
  uint8_t *p = abuf + tt * 4096 + q * 1536;
  *(uint4 *)(p + li * 16) = make_uint4(o[0], o[1], o[2], o[3]);
  *(uint2 *)(p + 1024 + li * 8) = make_uint2(o[4], o[5]);
}

// The detect of a finished chain (its accumulators and beam scale) is
// deferred: it runs two |B|^2 terms at a time between the next chain's MFMAs,
// where the VALU work hides behind the matrix pipe, instead of as a block of
// VALU after the chain that the MFMAs wait for.
struct Pending {
  f16v r, i;
  float sc;
};

// Element v of the pending chain into the running bin sums, in the order and
// with the partial sums of detect1 (even v -> sums 0/1, odd v -> 2/3), so the
// result is bit-identical.
template <int NP>
__device__ __forceinline__ void det_elem(const Pending &pd, int v, float (&sd)[NP][4]) {
  constexpr int VPB = 16 / NP;
  const int j = v / VPB, o = 2 * (v & 1);
  sd[j][o] = __builtin_fmaf(pd.r[v], pd.r[v], sd[j][o]);
  sd[j][o + 1] = __builtin_fmaf(pd.i[v], pd.i[v], sd[j][o + 1]);
  // an empty use, ordered with the scheduling barriers around it, so that the
  // compiler keeps the two FMAs at this point of the MFMA sequence
  asm volatile("" ::"v"(sd[j][o]), "v"(sd[j][o + 1]));
}

template <int NP>
__device__ __forceinline__ void det_final(const float (&sd)[NP][4], float sc, float (&part)[NP]) {
#pragma unroll
  for (int j = 0; j < NP; j++)
    part[j] = __builtin_fmaf((sd[j][0] + sd[j][1]) + (sd[j][2] + sd[j][3]), sc, part[j]);
}

// FP16 unit from LDS A_re: the MFMA chains of gemm_unit in the same order
// (both beam tiles per K step), so results are bit-identical to the classic
// schedule, A_im derived from A_re.  With PF, each K step of W is reloaded
// from row wnext right after its last MFMA has issued (a rolling prefetch of
// the next (channel, pol)'s weights that needs no extra registers), pinned
// there by scheduling barriers that VALU, SALU and LDS work may still cross.
// The prefetches are buffer loads whose row offset (wave-uniform) is a
// scalar: no 64-bit address arithmetic per load.
// (For FP16 this block order measured faster than the deferred-detect chains
// that FP8 uses: those need A_im derived once per beam tile.)
template <bool PF>
__device__ __forceinline__ void unit16(Acc<BTW> &A, WFrag<0, BTW, 1> &W, const uint8_t *unit, int lane,
                                       const WSrc &src, uint32_t wnext) {
  __builtin_amdgcn_s_setprio(1);
#pragma unroll
  for (int bt = 0; bt < BTW; bt++) { A.r[bt] = (f16v){}; A.i[bt] = (f16v){}; }
#pragma unroll
  for (int s = 0; s < 8; s++) {
    const uint4 r = *(const uint4 *)(unit + (s * 64 + lane) * 16);
    // A_re = [Dr0,Dr1,-Di0,-Di1,Dr2,Dr3,-Di2,-Di3] -> A_im = [Di0,Di1,Dr0,Dr1,...]
    const uint4 m = make_uint4(r.y ^ 0x80008000u, r.x, r.w ^ 0x80008000u, r.z);
    const h8v are = __builtin_bit_cast(h8v, r), aim = __builtin_bit_cast(h8v, m);
#pragma unroll
    for (int bt = 0; bt < BTW; bt++) {
      A.r[bt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(are, W.w[bt][0][s], A.r[bt], 0, 0, 0);
      A.i[bt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(aim, W.w[bt][0][s], A.i[bt], 0, 0, 0);
    }
    if constexpr (PF) {
      __builtin_amdgcn_sched_barrier(0x386);
#pragma unroll
      for (int bt = 0; bt < BTW; bt++) W.load_one_rb(src, wnext, bt, s);
      __builtin_amdgcn_sched_barrier(0x386);
    }
  }
  __builtin_amdgcn_s_setprio(0);
}

// FP8 / FP6 chain: beam tile BT of one 32-sample unit, the unit's A_re /
// A_im for both K steps given, accumulated over K in the order of gemm_unit
// (per step the weight parts hi, then lo; or hi, mid, lo for the three-part
// FP6 split), so FP8 results are bit-identical to the classic schedule.
// Between its MFMAs it detects the pending chain into `into`, two elements
// after each MFMA from the second on, in element order (so the sums are those
// of detect1).  With PF, each K step of W[BT] is reloaded from row wnext
// right after its last MFMA (the rolling prefetch of unit16).  Scheduling
// barriers keep this interleave as written.
template <int PREC, int BT, bool PF, int NP>
__device__ __forceinline__ void chain8(f16v &cr, f16v &ci, WFrag<PREC, BTW, 1> &W, const i8v (&are)[2],
                                       const i8v (&aim)[2], const WSrc &src, uint32_t wnext,
                                       const Pending &pd, float (&into)[NP]) {
  constexpr int NPART = PREC == 3 ? 3 : 2;
  float sd[NP][4];
#pragma unroll
  for (int j = 0; j < NP; j++) sd[j][0] = sd[j][1] = sd[j][2] = sd[j][3] = 0.f;
  cr = (f16v){};
  ci = (f16v){};
#pragma unroll
  for (int s2 = 0; s2 < 2; s2++) {
#pragma unroll
    for (int pt = 0; pt < NPART; pt++) {
      const int m = (s2 * NPART + pt) * 2;   // index of this pair's first MFMA in the chain
      const i8v *w;
      int sb;
      if constexpr (PREC == 3) {
        w = pt == 0 ? &W.hi[BT][0][s2] : pt == 1 ? &W.mid[BT][0][s2] : &W.lo[BT][0][s2];
        sb = pt == 0 ? W.shi(BT, 0, s2) : pt == 1 ? W.smid(BT, 0, s2) : W.slo(BT, 0, s2);
      } else {
        w = pt == 0 ? &W.hi[BT][0][s2] : &W.lo[BT][0][s2];
        sb = pt == 0 ? W.shi(BT, 0, s2) : W.slo(BT, 0, s2);
      }
      cr = mfma64<PREC>(are[s2], *w, cr, sb);
      if (m > 0 && 2 * m - 1 < 16) { det_elem<NP>(pd, 2 * m - 2, sd); det_elem<NP>(pd, 2 * m - 1, sd); }
      __builtin_amdgcn_sched_barrier(0);
      ci = mfma64<PREC>(aim[s2], *w, ci, sb);
      if (2 * m + 1 < 16) { det_elem<NP>(pd, 2 * m, sd); det_elem<NP>(pd, 2 * m + 1, sd); }
      if (PF && pt == NPART - 1) W.load_one_rb(src, wnext, BT, s2);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  constexpr int DONE = 4 * NPART * 2 - 2;   // elements detected between the MFMAs
#pragma unroll
  for (int v = DONE; v < 16; v++) det_elem<NP>(pd, v, sd);
  det_final<NP>(sd, pd.sc, into);
}

// One FP8 or FP6 unit, both beam tiles: chain 0 detects the pending chain
// into into0, chain 1 detects chain 0 into into1, and chain 1 is left pending.
template <int PREC, bool PF, int NP>
__device__ __forceinline__ void unit8(Pending &pd, WFrag<PREC, BTW, 1> &W, const uint8_t *unit, int lane,
                                      const WSrc &src, uint32_t wnext, float (&into0)[NP], float (&into1)[NP]) {
  f16v cr, ci;
  __builtin_amdgcn_s_setprio(1);
  i8v are[2], aim[2];
#pragma unroll
  for (int s2 = 0; s2 < 2; s2++) {
    if constexpr (PREC == 1) {
      const uint4 r0 = *(const uint4 *)(unit + ((s2 * 2) * 64 + lane) * 16);
      const uint4 r1 = *(const uint4 *)(unit + ((s2 * 2 + 1) * 64 + lane) * 16);
      // A_re dwords [Dr, -Di] per 4 antennas -> A_im [Di, Dr]
      are[s2] = (i8v){(int)r0.x, (int)r0.y, (int)r0.z, (int)r0.w, (int)r1.x, (int)r1.y, (int)r1.z, (int)r1.w};
      aim[s2] = (i8v){(int)(r0.y ^ 0x80808080u), (int)r0.x, (int)(r0.w ^ 0x80808080u), (int)r0.z,
                      (int)(r1.y ^ 0x80808080u), (int)r1.x, (int)(r1.w ^ 0x80808080u), (int)r1.z};
    } else {
      const uint4 r0 = *(const uint4 *)(unit + s2 * 1536 + lane * 16);
      const uint2 r1 = *(const uint2 *)(unit + s2 * 1536 + 1024 + lane * 8);
      f6_re_im(r0, r1, are[s2], aim[s2]);
    }
  }
  chain8<PREC, 0, PF, NP>(cr, ci, W, are, aim, src, wnext, pd, into0);
  pd.r = cr; pd.i = ci; pd.sc = W.sc[0];
  chain8<PREC, 1, PF, NP>(cr, ci, W, are, aim, src, wnext, pd, into1);
  __builtin_amdgcn_s_setprio(0);
  pd.r = cr; pd.i = ci; pd.sc = W.sc[1];
}

template <int PREC> constexpr int abuf_bytes() { return PREC == 0 ? AB16 : 2 * AB8; }   // double buffers
static_assert(2 * 4096 <= AB8, "FP6 operands of one (channel, pol) fit the FP8 half-buffer");

// Tile L of the block described by a; stage and abuf are the workgroup's LDS.
// GROUP: weights per 8-channel group, so the tile's (channel, pol) steps use
// two weight rows only, one per pol.
//   0: the rolling weights of fine mode (each step loads the next step's row).
//   1 (FP8 / FP6 hi+lo): both rows stay in registers for the whole tile and
//      the steps load no weights.  The same MFMAs, detects and barriers as the
//      rolling path, in the same order: the results are identical.
//   2: the steps run pol-outer, (channel 0..7, pol 0) then (channel 0..7,
//      pol 1), so one row is held at a time and the second is loaded in place
//      once, during step 7's second unit, with the rolling path's registers.
//      The powers are summed in another order than the rolling path's.
template <int PREC, int TAVG, int GROUP = 0>
__device__ __forceinline__ void shared_tile(const BfArgs &a, int L, uint8_t *stage, uint8_t *abuf) {
  static_assert(GROUP != 1 || PREC == 1 || PREC == 2, "FP16 and three-part FP6 cannot hold both rows");
  constexpr int UNIT = PREC == 0 ? AB16 / 2 : AB8 / 2;    // operand bytes per unit
  constexpr int NP = TAVG <= 32 ? 32 / TAVG : 1;
  const TileId t = tile_id(a, L);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform: weight rows stay scalar
  WSrc src;
  src.rw = buffer_rsrc(a.wfrag, a.wfrag_bytes);
  src.rs = buffer_rsrc(a.wscale, a.wscale_bytes);
  src.vl = (uint32_t)lane * (PREC == 0 ? 16u : 32u);
  src.vn4 = (uint32_t)(lane & 31) * 4u;
  const int n = lane & 31, h = lane >> 5;
  const int btw = (t.bset * NW + wave) * BTW;

  // Weights of (channel, pol) 0 are loaded right after the staging loads:
  // their latency overlaps the staging, and the stage stores (which wait for
  // every older load) do not wait for them.  The second unit of each (channel,
  // pol) prefetches the next one's in place (the last one reloads itself).
  WFrag<PREC, BTW, 1> W;
  WFrag<PREC, BTW, 1> W1;                 // pol 1's weights (GROUP 1)

  // (a) stage: the two (packet, antenna quad) tasks of each thread, every
  // load issued before the first store.  Buffer loads without branches: a
  // load inside a branch would be waited for at the branch's join, one round
  // trip per antenna.
  static_assert(PCU * 16 == 2 * NT, "two staging tasks per thread");
  {
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    const __amdgpu_buffer_rsrc_t rin = buffer_rsrc(a.in, a.in_bytes);
    uint32_t x[2][4][8];
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const int task = tid + k * NT, pl = task >> 4, aq = task & 15;
      const long p = (long)t.tblk * PCU + pl;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        // padding antennas read past the block's end: the buffer returns zeros
        const int ant = aq * 4 + j;
        const uint32_t off = ant < a.nant
                                 ? (uint32_t)chunk_offset(a.order, p, ant, t.g, a.nant, a.nchan, a.ngroup)
                                 : 0x80000000u;
        const u4 v0 = __builtin_amdgcn_raw_buffer_load_b128(rin, off, 0, 0);
        const u4 v1 = __builtin_amdgcn_raw_buffer_load_b128(rin, off, 16, 0);
        x[k][j][0] = v0.x; x[k][j][1] = v0.y; x[k][j][2] = v0.z; x[k][j][3] = v0.w;
        x[k][j][4] = v1.x; x[k][j][5] = v1.y; x[k][j][6] = v1.z; x[k][j][7] = v1.w;
      }
    }
    {
      const long wr = weight_row(a, t.g, 0);
      __builtin_amdgcn_sched_barrier(0);
      W.load(a.wfrag, a.wscale, wr * a.nbt + btw, wr * a.nbeam_pad + btw * 32, lane, n);
      if constexpr (GROUP == 1) {
        const long wr1 = weight_row(a, t.g, 1);
        W1.load(a.wfrag, a.wscale, wr1 * a.nbt + btw, wr1 * a.nbeam_pad + btw * 32, lane, n);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const int task = tid + k * NT;
      shared_stage_store(stage, x[k], task >> 4, task & 15);
    }
  }
  __syncthreads();

  float part[2][BTW][NP];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int bt = 0; bt < BTW; bt++)
#pragma unroll
      for (int j = 0; j < NP; j++) part[i][bt][j] = 0.f;

  // (b) operands to LDS one step ahead in a double buffer (FP16: per unit,
  // 8 KB halves; FP8: per (channel, pol), 8 KB halves), so each wave's unpack
  // of the next step interleaves with its MFMAs of this one and one barrier
  // per step publishes the next half and frees this one
  if constexpr (PREC == 0 && GROUP == 2) {
    // step k takes (channel, pol) 2 (k % 8) + k / 8; steps 0-6 and 8-15 in
    // loops that load no weights, step 7 between them loads pol 1's row (one
    // loop with a branch on k == 7 would need copies of the fragments at the
    // join and spill)
    const uint32_t w1 = weight_row(a, t.g, 1);
    auto cqof = [](int k) { return 2 * (k & 7) + (k >> 3); };
    auto step = [&](int k, auto prefetch) {
      const int cq = cqof(k);
      Acc<BTW> A;
      unpack_unit_f16(stage, abuf + UNIT, 2 * cq + 1, tid);
      unit16<false>(A, W, abuf, lane, src, 0);
      detect<TAVG>(A, W.sc, part[0]);
      __syncthreads();
      if (k < 15) unpack_unit_f16(stage, abuf, 2 * cqof(k + 1), tid);
      unit16<decltype(prefetch)::value>(A, W, abuf + UNIT, lane, src, w1 * a.nbt + btw);
      detect<TAVG>(A, W.sc, part[1]);
      if constexpr (decltype(prefetch)::value) W.load_scale_rb(src, w1 * a.nbeam_pad + btw * 32);
      __syncthreads();
    };
    unpack_unit_f16(stage, abuf, 0, tid);
    __syncthreads();
#pragma unroll 1
    for (int k = 0; k < 7; k++) step(k, std::false_type{});
    step(7, std::true_type{});
#pragma unroll 1
    for (int k = 8; k < 16; k++) step(k, std::false_type{});
  } else if constexpr (PREC == 0) {
    unpack_unit_f16(stage, abuf, 0, tid);
    __syncthreads();
    for (int cq = 0; cq < 16; cq++) {
      const uint32_t wn = weight_row(a, t.g, cq < 15 ? cq + 1 : cq);
      Acc<BTW> A;
      unpack_unit_f16(stage, abuf + UNIT, 2 * cq + 1, tid);
      unit16<false>(A, W, abuf, lane, src, 0);
      detect<TAVG>(A, W.sc, part[0]);
      __syncthreads();
      if (cq < 15) unpack_unit_f16(stage, abuf, 2 * cq + 2, tid);
      unit16<true>(A, W, abuf + UNIT, lane, src, wn * a.nbt + btw);
      detect<TAVG>(A, W.sc, part[1]);
      W.load_scale_rb(src, wn * a.nbeam_pad + btw * 32);
      __syncthreads();
    }
  } else {
    // Chains in order (cq, unit, beam tile); each one's detect is deferred to
    // the next chain.  The first pending chain is all zero, which adds exactly
    // nothing; the last one is detected after the loop.
    Pending pd;
    pd.r = (f16v){};
    pd.i = (f16v){};
    pd.sc = 0.f;
    auto unpack_cq = [&](uint8_t *dst, int c) {
      if constexpr (PREC == 1) unpack_cq_f8(stage, dst, c, tid);
      else unpack_cq_f6(stage, dst, c, tid);
    };
    unpack_cq(abuf, 0);
    __syncthreads();
    if constexpr (GROUP == 1) {
      auto step = [&](int cq, WFrag<PREC, BTW, 1> &Wc) {
        if (cq < 15) unpack_cq(abuf + ((cq + 1) & 1) * AB8, cq + 1);
        const uint8_t *ab = abuf + (cq & 1) * AB8;
        unit8<PREC, false, NP>(pd, Wc, ab, lane, src, 0, part[1][1], part[0][0]);
        unit8<PREC, false, NP>(pd, Wc, ab + UNIT, lane, src, 0, part[0][1], part[1][0]);
        __syncthreads();
      };
      for (int cq = 0; cq < 16; cq += 2) {
        step(cq, W);
        step(cq + 1, W1);
      }
    } else if constexpr (GROUP == 2) {
      // step k takes (channel, pol) 2 (k % 8) + k / 8, laid out as in the
      // FP16 branch
      const uint32_t w1 = weight_row(a, t.g, 1);
      auto step = [&](int k, auto prefetch) {
        if (k < 15) unpack_cq(abuf + ((k + 1) & 1) * AB8, 2 * ((k + 1) & 7) + ((k + 1) >> 3));
        const uint8_t *ab = abuf + (k & 1) * AB8;
        unit8<PREC, false, NP>(pd, W, ab, lane, src, 0, part[1][1], part[0][0]);
        unit8<PREC, decltype(prefetch)::value, NP>(pd, W, ab + UNIT, lane, src, w1 * a.nbt + btw, part[0][1],
                                                   part[1][0]);
        if constexpr (decltype(prefetch)::value) W.load_scale_rb(src, w1 * a.nbeam_pad + btw * 32);
        __syncthreads();
      };
#pragma unroll 1
      for (int k = 0; k < 7; k++) step(k, std::false_type{});
      step(7, std::true_type{});
#pragma unroll 1
      for (int k = 8; k < 16; k++) step(k, std::false_type{});
    } else {
      for (int cq = 0; cq < 16; cq++) {
        const uint32_t wn = weight_row(a, t.g, cq < 15 ? cq + 1 : cq);
        if (cq < 15) unpack_cq(abuf + ((cq + 1) & 1) * AB8, cq + 1);
        const uint8_t *ab = abuf + (cq & 1) * AB8;
        unit8<PREC, false, NP>(pd, W, ab, lane, src, 0, part[1][1], part[0][0]);
        unit8<PREC, true, NP>(pd, W, ab + UNIT, lane, src, wn * a.nbt + btw, part[0][1], part[1][0]);
        W.load_scale_rb(src, wn * a.nbeam_pad + btw * 32);
        __syncthreads();
      }
    }
    float sd[NP][4];
#pragma unroll
    for (int j = 0; j < NP; j++) sd[j][0] = sd[j][1] = sd[j][2] = sd[j][3] = 0.f;
#pragma unroll
    for (int v = 0; v < 16; v++) det_elem<NP>(pd, v, sd);
    det_final<NP>(sd, pd.sc, part[1][1]);
  }

  // (c)
  tile_output<TAVG>(a, t, btw, n, h, part, 0, 2);
}
}  // namespace shared
#endif

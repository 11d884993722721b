// bf_common.h -- device helpers shared by the gfx950 beamformer kernels:
// vector types, the weight-fragment layout, FP8 E4M3 conversion, the 4-bit
// voltage unpack into MFMA operand slots and the uint8 requantisation.
#ifndef BF_COMMON_H
#define BF_COMMON_H
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include "bf_launch.h"

typedef _Float16 h2v __attribute__((ext_vector_type(2)));
typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef int i8v __attribute__((ext_vector_type(8)));

// Last launcher error (one per host thread), set by bfk_set_error.
void bfk_set_error(const char *fmt, ...);
int bfk_check(hipError_t e, const char *what);

// ---------------------------------------------------------------------------
// Weight fragment layouts (written by bf_weights_kernel, read by the fused
// kernel).  Antenna a of 64-antenna chunk kc maps to lane half h = (a%64)/32.
// FP16 (v_mfma_f32_32x32x16_f16, 8 K-slots per lane per step):
//   step s = (a%32)/4, slots [Wr a0, Wr a1, Wi a0, Wi a1, Wr a2, Wr a3, Wi a2,
//   Wi a3] for a_i = 64kc + 32h + 4s + i; one 16-byte record per lane,
//   index (((cq*nbt + bt)*nkc + kc)*8 + s)*64 + lane, cq = wc*2 + pol.
// FP8 (v_mfma_f32_32x32x64_f8f6f4, 32 K-slots per lane per step):
//   step s2 = (a%32)/16, dword 2u = Wr of ants 4u..4u+3 of the 16,
//   dword 2u+1 = Wi; hi and lo parts are separate 32-byte records,
//   index ((((cq*nbt + bt)*nkc + kc)*2 + s2)*2 + hl)*64 + lane.
// FP6 (the same MFMA with E2M3 operands, 6-bit slot j at bits 6j..6j+5 of six
//   dwords): step s2 as for FP8, slots 0..15 = Wr of the 16 antennas, slots
//   16..31 = Wi; hi and lo records at the FP8 indices, the operand in bytes
//   0..23; the hi record's dwords 6 and 7 hold the lane's E8M0 B-operand
//   scales of the hi and lo parts.
// FP6 three-part (hi + mid + lo E2M3 weights): hi, mid and lo records per
//   step s2 at index ((((cq*nbt + bt)*nkc + kc)*2 + s2)*3 + part)*64 + lane;
//   the hi record's dwords 6 and 7 hold the hi and mid E8M0 scales, the mid
//   record's dword 6 the lo scale.
// All formats but the three-part one use 128 bytes per lane per (cq, beam
// tile, antenna chunk); it uses 192.
// The matching A (voltage) slots are built by unpack_f16 / nib_e4m3 / the
// FP6 converter (unpack_cq_f6).
// ---------------------------------------------------------------------------
constexpr int WFRAG_LANE_BYTES = 128;

// OCP FP8 E4M3 (gfx950 "fp8"): bias 7, max 448, subnormal step 2^-9.
// Round to nearest even, saturating.  Byte k in 0..15 encodes k * 2^-9.
__device__ static inline uint32_t e4m3_encode(double v) {
  uint32_t sgn = v < 0 ? 0x80u : 0u;
  double a = fabs(v);
  if (a >= 448.0) return sgn | 0x7Eu;
  if (a < 0.015625) {                     // below 2^-6: subnormal grid 2^-9
    double q = rint(a * 512.0);           // 0..8 (8 = smallest normal)
    return sgn | (uint32_t)q;
  }
  int e;
  double m = frexp(a, &e);                // a = m * 2^e, m in [0.5, 1)
  double mq = rint((m * 2.0 - 1.0) * 8.0);
  int ex = e - 1;
  if (mq >= 8.0) { mq = 0.0; ex += 1; }
  if (ex > 8 || (ex == 8 && mq > 6.0)) return sgn | 0x7Eu;
  return sgn | ((uint32_t)(ex + 7) << 3) | (uint32_t)mq;
}
__device__ static inline double e4m3_decode(uint32_t b) {
  double s = (b & 0x80u) ? -1.0 : 1.0;
  uint32_t e = (b >> 3) & 15u, m = b & 7u;
  if (e == 0) return s * (double)m * (1.0 / 512.0);
  return s * (1.0 + m / 8.0) * ldexp(1.0, (int)e - 7);
}

// FP16 unpack: 4 packed samples (4 antennas) -> A_re = [Dr0,Dr1,-Di0,-Di1,
// Dr2,Dr3,-Di2,-Di3] and A_im = [Di0,Di1,Dr0,Dr1,Di2,Di3,Dr2,Dr3] against
// weight slots [Wr,Wr,Wi,Wi,...].  A nibble n is formed exactly as
// half(0x6400 | (n ^ 8)) - 1032 = n.
__device__ __forceinline__ void unpack_f16(uint32_t x, h8v &are, h8v &aim) {
  const uint32_t x8 = x ^ 0x88888888u;
  const uint32_t lo = x8 & 0x0F0F0F0Fu;
  const uint32_t hi = (x8 >> 4) & 0x0F0F0F0Fu;
  const h2v off = {(_Float16)-1032.0f, (_Float16)-1032.0f};
  h2v r01 = __builtin_bit_cast(h2v, __builtin_amdgcn_perm(0x64646464u, lo, 0x04010400u)) + off;
  h2v r23 = __builtin_bit_cast(h2v, __builtin_amdgcn_perm(0x64646464u, lo, 0x04030402u)) + off;
  h2v i01 = __builtin_bit_cast(h2v, __builtin_amdgcn_perm(0x64646464u, hi, 0x04010400u)) + off;
  h2v i23 = __builtin_bit_cast(h2v, __builtin_amdgcn_perm(0x64646464u, hi, 0x04030402u)) + off;
  h2v n01 = __builtin_bit_cast(h2v, __builtin_bit_cast(uint32_t, i01) ^ 0x80008000u);
  h2v n23 = __builtin_bit_cast(h2v, __builtin_bit_cast(uint32_t, i23) ^ 0x80008000u);
  are = (h8v){r01[0], r01[1], n01[0], n01[1], r23[0], r23[1], n23[0], n23[1]};
  aim = (h8v){i01[0], i01[1], r01[0], r01[1], i23[0], i23[1], r23[0], r23[1]};
}

// FP8 unpack: E4M3 byte k (k <= 15, sign in bit 7) is exactly k * 2^-9, so a
// nibble only needs its two's complement turned into sign-magnitude:
// |n| = (s ^ smear) + sign.  Applied to x gives Dr, to x >> 4 gives Di.
__device__ __forceinline__ uint32_t nib_e4m3(uint32_t xs) {
  const uint32_t s = xs & 0x0F0F0F0Fu;
  const uint32_t m = xs & 0x08080808u;     // sign bit of each nibble
  const uint32_t t = m >> 3;               // 1 per negative byte
  const uint32_t sm = (m << 1) - t;        // 0x0F per negative byte
  return ((s ^ sm) + t) | (m << 4);
}

// OCP FP6 E2M3 (gfx950 "fp6"): bias 1, max 7.5, subnormal step 1/8.  Round
// to nearest even, saturating.  Code c in 0..8 encodes c/8, so a 4-bit
// voltage v is the sign-magnitude code (v<0)<<5 | |v| worth v/8, exactly.
// Slot j of a packed operand is at bits 6j..6j+5 of its six dwords.
__device__ static inline uint32_t e2m3_encode(double v) {
  uint32_t sgn = v < 0 ? 0x20u : 0u;
  double a = fabs(v);
  if (a >= 7.5) return sgn | 0x1Fu;
  if (a < 1.0) return sgn | (uint32_t)rint(a * 8.0);   // 0..8 (8 = 1.0, the smallest normal)
  int e;
  double m = frexp(a, &e);                // a = m * 2^e, m in [0.5, 1), e in 1..3
  double mq = rint((m * 2.0 - 1.0) * 8.0);
  int ex = e - 1;
  if (mq >= 8.0) { mq = 0.0; ex += 1; }
  if (ex > 2) return sgn | 0x1Fu;
  return sgn | ((uint32_t)(ex + 1) << 3) | (uint32_t)mq;
}
__device__ static inline double e2m3_decode(uint32_t c) {
  double s = (c & 0x20u) ? -1.0 : 1.0;
  uint32_t e = (c >> 3) & 3u, m = c & 7u;
  if (e == 0) return s * (double)m * 0.125;
  return s * (1.0 + m / 8.0) * ldexp(1.0, (int)e - 1);
}

// Sign bits of 16 packed E2M3 slots (bits 6j+5) in each of the 3 dwords.
constexpr uint32_t E2M3_SIGN0 = 0x20820820u, E2M3_SIGN1 = 0x08208208u, E2M3_SIGN2 = 0x82082082u;

// (unsigned char)(__float2int_rn(x*scale)/2), dsaX_beamformer.cu:150-156:
// round half to even, saturate to int, C division, wrap modulo 256.
__device__ __forceinline__ uint8_t requant_u8(float x, float scale) {
  float v = x * scale;
  float r = __builtin_rintf(v);
  int iv;
  if (!(r == r)) iv = 0;
  else if (r >= 2147483648.0f) iv = 2147483647;
  else if (r <= -2147483648.0f) iv = (-2147483647 - 1);
  else iv = (int)r;
  return (uint8_t)(unsigned)(iv / 2);
}

// Byte offset of the 32-byte chunk (8 channels x 2 times x 2 pols) of
// packet p, antenna a, channel group g in an input block.
__device__ __forceinline__ long chunk_offset(int order, long p, int a, int g, int nant, int nchan,
                                             int ngroup) {
  return order == 0 ? ((p * nant + a) * nchan + g * 8) * 4 : ((p * ngroup + g) * nant + a) * 32;
}

// ---------------------------------------------------------------------------
// The fused beamformer's tile.  A tile is one 8-channel output group g x 64
// samples x a set of beams.  Its voltages are staged in LDS as rows
// [channel*2 + pol][sample][antenna] of 64 packed bytes (4-bit re | 4-bit im
// << 4), padded to ROWB bytes; row r of (channel, pol) cq sits at
// (cq*TCU + r)*ROWB.  A wave owns BTW 32-beam tiles; for each (channel, pol)
// and each 32-sample half of the tile (a "unit") it runs the complex GEMM on
// MFMA with its weights in registers, then detects |B|^2 into per-bin sums.
// ---------------------------------------------------------------------------
constexpr int BTW = 2;                    // 32-beam tiles per wave
constexpr int TCU = 64;                   // samples per tile
constexpr int PCU = TCU / 2;              // packets per tile
constexpr int ROWB = 80;                  // LDS bytes per staged row
constexpr int STAGE_BYTES = 16 * TCU * ROWB;

struct BfArgs {
  const uint8_t *in;
  const uint8_t *wfrag;
  const float *wscale;
  const float *scale;
  uint8_t *out;
  float *power;   // optional
  float *bpart;   // [nbeam][nsamp/32][ngroup] per-unit beam power sums
  int split;      // classic: wave pairs split a tile's two 32-sample units
  int nant, nchan, ngroup, nbeam, nbeam_pad, nbt, wmode, order;
  int ntblk, nbsets, ntiles, tiles_per_xcd;
  int layout, nout;                       // output layout (0 [t][beam][g], 1 [beam][t][g]), bins
  int kind;                               // tile schedule of a mixed batch job (MIX_*)
  uint32_t pstride, astride, gstride;     // input bytes per packet, antenna, channel group
  uint32_t in_bytes, wfrag_bytes, wscale_bytes;
};

// Output (and power) index of bin o, beam b, group g: [o][beam][g], the
// BF_BLOCK_KEY layout (dsaX_beamformer.cu:150-156), or beam-major
// [beam][o][g], in which beams 64s..64s+63 are byte for byte the buffer that
// dsaX_dbnic's transmit thread s builds by its transpose (dsaX_dbnic.c:111-116).
__device__ __forceinline__ long out_index(int layout, int nout, int nbeam, int ngroup, int o, int b, int g) {
  return layout == 0 ? ((long)o * nbeam + b) * ngroup + g : ((long)b * nout + o) * ngroup + g;
}

// Tile L -> (time block, group, beam set).  Consecutive tiles run at the same
// time on one XCD.  When ngroup is a multiple of 4 the order is (group mod 4)
// fastest, then time block, then group / 4: the four tiles that read the four
// 32-byte quarters of each 128-byte input line run together, so one XCD's L2
// fetches each line once, and the following tiles still share those groups'
// weights in L2.  Otherwise time blocks are the fastest index.
struct TileId {
  int tblk, g, bset;
};
__device__ __forceinline__ TileId tile_id(const BfArgs &a, int L) {
  TileId t;
  if ((a.ngroup & 3) == 0) {
    const int gl = L & 3;
    int rest = L >> 2;
    t.tblk = rest % a.ntblk;
    rest /= a.ntblk;
    const int gq = a.ngroup >> 2;
    t.g = (rest % gq) * 4 + gl;
    t.bset = rest / gq;
    return t;
  }
  t.tblk = L % a.ntblk;
  const int rest = L / a.ntblk;
  t.g = rest % a.ngroup;
  t.bset = rest / a.ngroup;
  return t;
}

// Weight row of (channel, pol) cq of group g: fine mode has one per channel.
__device__ __forceinline__ int weight_row(const BfArgs &a, int g, int cq) {
  return (a.wmode ? g * 8 + (cq >> 1) : g) * 2 + (cq & 1);
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void *p, uint32_t bytes) {
  // raw (stride 0) buffer, gfx9 dword-3 format bits; reads past `bytes`
  // return zero instead of touching memory
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}

// Weight fragments of one (channel, pol) for the wave's NB beam tiles and NK
// 64-antenna chunks, and the per-beam power factor that undoes the weight (and
// FP8 voltage) scaling.  wrow_bt = weight row * nbt + first beam tile and
// sc_off = weight row * nbeam_pad + first beam are wave-uniform.
// load(): flat loads; load_buf(): buffer loads whose offsets beyond the lane
// part are wave-uniform (SGPRs), so no per-load VGPR address is needed.
// Buffer sources of the weight fragments and scales for rolling prefetches
// whose row offsets are wave-uniform: the row part of each address goes in
// the scalar offset and the lane part is one register, where a pointer would
// take a 64-bit address (and its arithmetic) per load.
struct WSrc {
  __amdgpu_buffer_rsrc_t rw, rs;
  uint32_t vl;    // lane byte offset in a fragment record: 16 * lane (FP16), 32 * lane (FP8 / FP6)
  uint32_t vn4;   // 4 * beam column n in a scale row
};
__device__ __forceinline__ i8v rb_ld32(const WSrc &src, uint32_t vo, uint32_t so) {
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  const u4 a = __builtin_amdgcn_raw_buffer_load_b128(src.rw, vo, so, 0);
  const u4 b = __builtin_amdgcn_raw_buffer_load_b128(src.rw, vo + 16, so, 0);
  return (i8v){(int)a.x, (int)a.y, (int)a.z, (int)a.w, (int)b.x, (int)b.y, (int)b.z, (int)b.w};
}
// 24 bytes (the six operand dwords of an FP6 fragment; dwords 6, 7 undefined)
__device__ __forceinline__ i8v rb_ld24(const WSrc &src, uint32_t vo, uint32_t so) {
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  typedef uint32_t u2 __attribute__((ext_vector_type(2)));
  const u4 a = __builtin_amdgcn_raw_buffer_load_b128(src.rw, vo, so, 0);
  const u2 b = __builtin_amdgcn_raw_buffer_load_b64(src.rw, vo + 16, so, 0);
  typedef int i4v __attribute__((ext_vector_type(4)));
  typedef int i2v __attribute__((ext_vector_type(2)));
  typedef int i6v __attribute__((ext_vector_type(6)));
  const i4v ai = __builtin_bit_cast(i4v, a);
  const i2v bi = __builtin_bit_cast(i2v, b);
  const i6v v = __builtin_shufflevector(ai, __builtin_shufflevector(bi, bi, 0, 1, -1, -1), 0, 1, 2, 3, 4, 5);
  return __builtin_shufflevector(v, v, 0, 1, 2, 3, 4, 5, -1, -1);
}
template <int NB>
__device__ __forceinline__ void rb_scale(float (&sc)[NB], const WSrc &src, uint32_t sc_off) {
#pragma unroll
  for (int bt = 0; bt < NB; bt++)
    sc[bt] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(src.rs, src.vn4, (sc_off + bt * 32u) * 4u, 0));
}

template <int PREC, int NB = BTW, int NK = 1> struct WFrag;
template <int NB, int NK> struct WFrag<0, NB, NK> {
  h8v w[NB][NK][8];
  float sc[NB];
  __device__ __forceinline__ void load(const uint8_t *wfrag, const float *wscale, long wrow_bt,
                                       long sc_off, int lane, int n) {
#pragma unroll
    for (int bt = 0; bt < NB; bt++)
#pragma unroll
      for (int kc = 0; kc < NK; kc++)
#pragma unroll
        for (int s = 0; s < 8; s++)
          w[bt][kc][s] = *(const h8v *)(wfrag + ((((wrow_bt + bt) * NK + kc) * 8 + s) * 64 + lane) * 16);
#pragma unroll
    for (int bt = 0; bt < NB; bt++) sc[bt] = wscale[sc_off + bt * 32 + n];
  }
  // One K step of every beam tile (NK == 1), for a rolling prefetch that
  // refills step s as soon as the last MFMA reading it has issued.
  __device__ __forceinline__ void load_step(const uint8_t *wfrag, long wrow_bt, int s, int lane) {
    static_assert(NK == 1, "rolling prefetch is for single-chunk tiles");
#pragma unroll
    for (int bt = 0; bt < NB; bt++)
      w[bt][0][s] = *(const h8v *)(wfrag + (((wrow_bt + bt) * 8 + s) * 64 + lane) * 16);
  }
  // Step s of chunk kc of beam tile bt, for a rolling prefetch that refills
  // it as soon as the last MFMA reading it has issued.
  __device__ __forceinline__ void load_k(const uint8_t *wfrag, long wrow_bt, int bt, int kc, int s, int lane) {
    w[bt][kc][s] = *(const h8v *)(wfrag + ((((wrow_bt + bt) * NK + kc) * 8 + s) * 64 + lane) * 16);
  }
  __device__ __forceinline__ void load_one(const uint8_t *wfrag, long wrow_bt, int bt, int s, int lane) {
    static_assert(NK == 1, "single-chunk tiles");
    load_k(wfrag, wrow_bt, bt, 0, s, lane);
  }
  __device__ __forceinline__ void load_scale(const float *wscale, long sc_off, int n) {
#pragma unroll
    for (int bt = 0; bt < NB; bt++) sc[bt] = wscale[sc_off + bt * 32 + n];
  }
  // load_one and load_scale through WSrc (wrow_bt and sc_off wave-uniform)
  __device__ __forceinline__ void load_one_rb(const WSrc &src, uint32_t wrow_bt, int bt, int s) {
    static_assert(NK == 1, "single-chunk tiles");
    w[bt][0][s] = __builtin_bit_cast(
        h8v, __builtin_amdgcn_raw_buffer_load_b128(src.rw, src.vl, ((wrow_bt + bt) * 8u + s) * 1024u, 0));
  }
  __device__ __forceinline__ void load_scale_rb(const WSrc &src, uint32_t sc_off) { rb_scale<NB>(sc, src, sc_off); }
  __device__ __forceinline__ void load_buf(__amdgpu_buffer_rsrc_t rw, __amdgpu_buffer_rsrc_t rs,
                                           uint32_t wrow_bt, uint32_t sc_off, uint32_t vlane16,
                                           uint32_t vn4) {
    // the scales first: a register copy of one then waits only for it
#pragma unroll
    for (int bt = 0; bt < NB; bt++)
      sc[bt] = __builtin_bit_cast(
          float, __builtin_amdgcn_raw_buffer_load_b32(rs, vn4, (sc_off + bt * 32) * 4u, 0));
#pragma unroll
    for (int bt = 0; bt < NB; bt++)
#pragma unroll
      for (int kc = 0; kc < NK; kc++)
#pragma unroll
        for (int s = 0; s < 8; s++)
          w[bt][kc][s] = __builtin_bit_cast(
              h8v, __builtin_amdgcn_raw_buffer_load_b128(
                       rw, vlane16, (((wrow_bt + bt) * NK + kc) * 8 + s) * 1024u, 0));
  }
};
template <int NB, int NK> struct WFrag<1, NB, NK> {
  i8v hi[NB][NK][2], lo[NB][NK][2];
  float sc[NB];
  // E8M0 operand scales of the hi / lo parts: FP8 fragments have none
  __device__ __forceinline__ int shi(int, int, int) const { return 0; }
  __device__ __forceinline__ int slo(int, int, int) const { return 0; }
  __device__ __forceinline__ void load(const uint8_t *wfrag, const float *wscale, long wrow_bt,
                                       long sc_off, int lane, int n) {
#pragma unroll
    for (int bt = 0; bt < NB; bt++)
#pragma unroll
      for (int kc = 0; kc < NK; kc++)
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++) {
          const long rec = ((((wrow_bt + bt) * NK + kc) * 2 + s2) * 2) * 64 + lane;
          hi[bt][kc][s2] = *(const i8v *)(wfrag + rec * 32);
          lo[bt][kc][s2] = *(const i8v *)(wfrag + (rec + 64) * 32);
        }
#pragma unroll
    for (int bt = 0; bt < NB; bt++) sc[bt] = wscale[sc_off + bt * 32 + n];
  }
  // Step s2 (hi and lo) of every beam tile (NK == 1); see WFrag<0>::load_step.
  __device__ __forceinline__ void load_step(const uint8_t *wfrag, long wrow_bt, int s2, int lane) {
    static_assert(NK == 1, "rolling prefetch is for single-chunk tiles");
#pragma unroll
    for (int bt = 0; bt < NB; bt++) {
      const long rec = (((wrow_bt + bt) * 2 + s2) * 2) * 64 + lane;
      hi[bt][0][s2] = *(const i8v *)(wfrag + rec * 32);
      lo[bt][0][s2] = *(const i8v *)(wfrag + (rec + 64) * 32);
    }
  }
  __device__ __forceinline__ void load_k(const uint8_t *wfrag, long wrow_bt, int bt, int kc, int s2, int lane) {
    const long rec = ((((wrow_bt + bt) * NK + kc) * 2 + s2) * 2) * 64 + lane;
    hi[bt][kc][s2] = *(const i8v *)(wfrag + rec * 32);
    lo[bt][kc][s2] = *(const i8v *)(wfrag + (rec + 64) * 32);
  }
  __device__ __forceinline__ void load_one(const uint8_t *wfrag, long wrow_bt, int bt, int s2, int lane) {
    static_assert(NK == 1, "single-chunk tiles");
    load_k(wfrag, wrow_bt, bt, 0, s2, lane);
  }
  __device__ __forceinline__ void load_scale(const float *wscale, long sc_off, int n) {
#pragma unroll
    for (int bt = 0; bt < NB; bt++) sc[bt] = wscale[sc_off + bt * 32 + n];
  }
  // load_one and load_scale through WSrc (wrow_bt and sc_off wave-uniform)
  __device__ __forceinline__ void load_one_rb(const WSrc &src, uint32_t wrow_bt, int bt, int s2) {
    static_assert(NK == 1, "single-chunk tiles");
    const uint32_t so = ((wrow_bt + bt) * 2u + s2) * 4096u;
    hi[bt][0][s2] = rb_ld32(src, src.vl, so);
    lo[bt][0][s2] = rb_ld32(src, src.vl + 2048u, so);
  }
  __device__ __forceinline__ void load_scale_rb(const WSrc &src, uint32_t sc_off) { rb_scale<NB>(sc, src, sc_off); }
  __device__ __forceinline__ void load_buf(__amdgpu_buffer_rsrc_t rw, __amdgpu_buffer_rsrc_t rs,
                                           uint32_t wrow_bt, uint32_t sc_off, uint32_t vlane16,
                                           uint32_t vn4) {
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int bt = 0; bt < NB; bt++)
      sc[bt] = __builtin_bit_cast(
          float, __builtin_amdgcn_raw_buffer_load_b32(rs, vn4, (sc_off + bt * 32) * 4u, 0));
#pragma unroll
    for (int bt = 0; bt < NB; bt++)
#pragma unroll
      for (int kc = 0; kc < NK; kc++)
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++) {
          // 32-byte records: lane l at rec*32 = base + 2*(16*l) -> two b128
          // loads at 32*l and 32*l + 16 (voffset 2*vlane16)
          const uint32_t base = ((((wrow_bt + bt) * NK + kc) * 2 + s2) * 2) * 64u * 32u;
#pragma unroll
          for (int hl = 0; hl < 2; hl++) {
            const uint32_t so = base + hl * 64u * 32u;
            const u4 a = __builtin_amdgcn_raw_buffer_load_b128(rw, 2 * vlane16, so, 0);
            const u4 b = __builtin_amdgcn_raw_buffer_load_b128(rw, 2 * vlane16 + 16, so, 0);
            const i8v v = {(int)a.x, (int)a.y, (int)a.z, (int)a.w,
                           (int)b.x, (int)b.y, (int)b.z, (int)b.w};
            if (hl == 0) hi[bt][kc][s2] = v;
            else lo[bt][kc][s2] = v;
          }
        }
  }
};

// FP6 hi + lo fragments: 6 operand dwords each; the hi fragment's dwords 6
// and 7 (which the MFMA does not read) carry the lane's E8M0 scales of both
// parts, the lo fragment's are unused.
template <int NB, int NK> struct WFrag<2, NB, NK> {
  i8v hi[NB][NK][2], lo[NB][NK][2];
  float sc[NB];
  __device__ __forceinline__ int shi(int bt, int kc, int s2) const { return hi[bt][kc][s2][6]; }
  __device__ __forceinline__ int slo(int bt, int kc, int s2) const { return hi[bt][kc][s2][7]; }
  static __device__ __forceinline__ i8v ld6(const uint8_t *p) {
    typedef int i4v __attribute__((ext_vector_type(4)));
    typedef int i2v __attribute__((ext_vector_type(2)));
    typedef int i6v __attribute__((ext_vector_type(6)));
    const i4v a = *(const i4v *)p;
    const i2v b = *(const i2v *)(p + 16);
    const i6v v = __builtin_shufflevector(a, __builtin_shufflevector(b, b, 0, 1, -1, -1), 0, 1, 2, 3, 4, 5);
    return __builtin_shufflevector(v, v, 0, 1, 2, 3, 4, 5, -1, -1);
  }
  __device__ __forceinline__ void load_k(const uint8_t *wfrag, long wrow_bt, int bt, int kc, int s2, int lane) {
    const long rec = ((((wrow_bt + bt) * NK + kc) * 2 + s2) * 2) * 64 + lane;
    hi[bt][kc][s2] = *(const i8v *)(wfrag + rec * 32);
    lo[bt][kc][s2] = ld6(wfrag + (rec + 64) * 32);
  }
  __device__ __forceinline__ void load_scale(const float *wscale, long sc_off, int n) {
#pragma unroll
    for (int bt = 0; bt < NB; bt++) sc[bt] = wscale[sc_off + bt * 32 + n];
  }
  __device__ __forceinline__ void load(const uint8_t *wfrag, const float *wscale, long wrow_bt,
                                       long sc_off, int lane, int n) {
#pragma unroll
    for (int bt = 0; bt < NB; bt++)
#pragma unroll
      for (int kc = 0; kc < NK; kc++)
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++) load_k(wfrag, wrow_bt, bt, kc, s2, lane);
    load_scale(wscale, sc_off, n);
  }
  __device__ __forceinline__ void load_one(const uint8_t *wfrag, long wrow_bt, int bt, int s2, int lane) {
    static_assert(NK == 1, "single-chunk tiles");
    load_k(wfrag, wrow_bt, bt, 0, s2, lane);
  }
  // load_one and load_scale through WSrc (wrow_bt and sc_off wave-uniform)
  __device__ __forceinline__ void load_one_rb(const WSrc &src, uint32_t wrow_bt, int bt, int s2) {
    static_assert(NK == 1, "single-chunk tiles");
    const uint32_t so = ((wrow_bt + bt) * 2u + s2) * 4096u;
    hi[bt][0][s2] = rb_ld32(src, src.vl, so);
    lo[bt][0][s2] = rb_ld24(src, src.vl + 2048u, so);
  }
  __device__ __forceinline__ void load_scale_rb(const WSrc &src, uint32_t sc_off) { rb_scale<NB>(sc, src, sc_off); }
};

// FP6 hi + mid + lo fragments (three-part split): the hi fragment's dwords 6
// and 7 carry the E8M0 scales of the hi and mid parts, the mid fragment's
// dword 6 that of the lo part.
template <int NB, int NK> struct WFrag<3, NB, NK> {
  i8v hi[NB][NK][2], mid[NB][NK][2], lo[NB][NK][2];
  float sc[NB];
  __device__ __forceinline__ int shi(int bt, int kc, int s2) const { return hi[bt][kc][s2][6]; }
  __device__ __forceinline__ int smid(int bt, int kc, int s2) const { return hi[bt][kc][s2][7]; }
  __device__ __forceinline__ int slo(int bt, int kc, int s2) const { return mid[bt][kc][s2][6]; }
  __device__ __forceinline__ void load_k(const uint8_t *wfrag, long wrow_bt, int bt, int kc, int s2, int lane) {
    const long rec = ((((wrow_bt + bt) * NK + kc) * 2 + s2) * 3) * 64 + lane;
    hi[bt][kc][s2] = *(const i8v *)(wfrag + rec * 32);
    mid[bt][kc][s2] = *(const i8v *)(wfrag + (rec + 64) * 32);
    lo[bt][kc][s2] = WFrag<2, NB, NK>::ld6(wfrag + (rec + 128) * 32);
  }
  __device__ __forceinline__ void load_scale(const float *wscale, long sc_off, int n) {
#pragma unroll
    for (int bt = 0; bt < NB; bt++) sc[bt] = wscale[sc_off + bt * 32 + n];
  }
  __device__ __forceinline__ void load(const uint8_t *wfrag, const float *wscale, long wrow_bt,
                                       long sc_off, int lane, int n) {
#pragma unroll
    for (int bt = 0; bt < NB; bt++)
#pragma unroll
      for (int kc = 0; kc < NK; kc++)
#pragma unroll
        for (int s2 = 0; s2 < 2; s2++) load_k(wfrag, wrow_bt, bt, kc, s2, lane);
    load_scale(wscale, sc_off, n);
  }
  __device__ __forceinline__ void load_one(const uint8_t *wfrag, long wrow_bt, int bt, int s2, int lane) {
    static_assert(NK == 1, "single-chunk tiles");
    load_k(wfrag, wrow_bt, bt, 0, s2, lane);
  }
  // load_one and load_scale through WSrc (wrow_bt and sc_off wave-uniform);
  // the lo part's 4096-byte offset is past the instruction's offset field, so
  // it goes in the scalar offset
  __device__ __forceinline__ void load_one_rb(const WSrc &src, uint32_t wrow_bt, int bt, int s2) {
    static_assert(NK == 1, "single-chunk tiles");
    const uint32_t so = ((wrow_bt + bt) * 2u + s2) * 6144u;
    hi[bt][0][s2] = rb_ld32(src, src.vl, so);
    mid[bt][0][s2] = rb_ld32(src, src.vl + 2048u, so);
    lo[bt][0][s2] = rb_ld24(src, src.vl, so + 4096u);
  }
  __device__ __forceinline__ void load_scale_rb(const WSrc &src, uint32_t sc_off) { rb_scale<NB>(sc, src, sc_off); }
};

// One 32x32x64 MFMA of the FP8 (PREC 1) or FP6 (PREC 2, 3) chains.  FP6
// weight parts are stored 2^k larger per lane; the lane's E8M0 B scale
// sb = 127 - k takes that back (A scale 2^0).  FP8 uses no scales.
template <int PREC>
__device__ __forceinline__ f16v mfma64(const i8v &a, const i8v &b, const f16v &c, int sb) {
  if constexpr (PREC == 1) return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 0, 0, 0, 0, 0, 0);
  else return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 2, 2, 0, 127, 0, sb);
}

// FP6 A_re of 16 antennas of one sample row (4 packed dwords, byte k of x[i]
// = antenna 4i + k) as six dwords: slots 0..15 = Dr, 16..31 = -Di.  A byte
// permute makes each nibble n the exact half 128 + (n^8)/8, minus 129 = n/8,
// and the converter packs the 32 halves.
__device__ __forceinline__ void f6_operand(const uint32_t (&x)[4], uint32_t (&o)[6]) {
  typedef _Float16 h32v __attribute__((ext_vector_type(32)));
  typedef int i6v __attribute__((ext_vector_type(6)));
  const h2v off = {(_Float16)-129.0f, (_Float16)-129.0f};
  h2v pr[8], pi[8];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t x8 = x[k] ^ 0x88888888u;
    const uint32_t lo = x8 & 0x0F0F0F0Fu, hi = (x8 >> 4) & 0x0F0F0F0Fu;
    pr[2 * k] = __builtin_bit_cast(h2v, __builtin_amdgcn_perm(0x58585858u, lo, 0x04010400u)) + off;
    pr[2 * k + 1] = __builtin_bit_cast(h2v, __builtin_amdgcn_perm(0x58585858u, lo, 0x04030402u)) + off;
    pi[2 * k] = __builtin_bit_cast(h2v, __builtin_amdgcn_perm(0x58585858u, hi, 0x04010400u)) + off;
    pi[2 * k + 1] = __builtin_bit_cast(h2v, __builtin_amdgcn_perm(0x58585858u, hi, 0x04030402u)) + off;
  }
  h32v src;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    src[2 * j] = pr[j][0];
    src[2 * j + 1] = pr[j][1];
    src[16 + 2 * j] = pi[j][0];
    src[16 + 2 * j + 1] = pi[j][1];
  }
  const i6v c = __builtin_amdgcn_cvt_scalef32_pk32_fp6_f16(src, 1.0f);
  o[0] = (uint32_t)c[0];
  o[1] = (uint32_t)c[1];
  o[2] = (uint32_t)c[2];
  o[3] = (uint32_t)c[3] ^ E2M3_SIGN0;   // -Di
  o[4] = (uint32_t)c[4] ^ E2M3_SIGN1;
  o[5] = (uint32_t)c[5] ^ E2M3_SIGN2;
}

// A_im = [Di | Dr] from A_re = [Dr | -Di] (16 six-bit slots each; dwords 6
// and 7 of the MFMA operand are unused).
__device__ __forceinline__ void f6_re_im(const uint4 &r0, const uint2 &r1, i8v &are, i8v &aim) {
  are = (i8v){(int)r0.x, (int)r0.y, (int)r0.z, (int)r0.w, (int)r1.x, (int)r1.y, 0, 0};
  aim = (i8v){(int)(r0.w ^ E2M3_SIGN0), (int)(r1.x ^ E2M3_SIGN1), (int)(r1.y ^ E2M3_SIGN2),
              (int)r0.x, (int)r0.y, (int)r0.z, 0, 0};
}

// One GEMM unit: 32 staged sample rows of one (channel, pol) against the
// wave's weights.  Rows are rowb bytes apart; lane (n, h) reads row n,
// antennas 64kc + 32h .. 64kc + 32h + 31 for each chunk kc.
template <int NB = BTW> struct Acc {
  f16v r[NB], i[NB];
};

// With PF, each K step of W is reloaded from row wnext (beam tile index of
// the next (channel, pol)) right after its last MFMA has issued: a rolling
// prefetch that needs no extra registers, pinned there by scheduling barriers
// that VALU, SALU and LDS work may still cross.
template <bool PF = false, int PREC, int NB, int NK>
__device__ __forceinline__ void gemm_unit(Acc<NB> &A, WFrag<PREC, NB, NK> &W, const uint8_t *rows,
                                          int rowb, int n, int h, const uint8_t *wfrag = nullptr,
                                          long wnext = 0) {
  // A wave issuing its MFMA chain runs at raised issue priority, so on a SIMD
  // shared by two waves the matrix pipe is fed first and the other wave's
  // detect/staging VALU work fills the gaps (measured: 8 % faster FP16 block,
  // and the scheduling boundary keeps register use at ~170).
  __builtin_amdgcn_s_setprio(1);
#pragma unroll
  for (int bt = 0; bt < NB; bt++) { A.r[bt] = (f16v){}; A.i[bt] = (f16v){}; }
#pragma unroll
  for (int kc = 0; kc < NK; kc++) {
    const uint8_t *src = rows + n * rowb + kc * 64 + h * 32;
    const uint4 d0 = *(const uint4 *)src, d1 = *(const uint4 *)(src + 16);
    const uint32_t dw[8] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
    if constexpr (PREC == 0) {
#pragma unroll
      for (int s = 0; s < 8; s++) {
        h8v are, aim;
        unpack_f16(dw[s], are, aim);
#pragma unroll
        for (int bt = 0; bt < NB; bt++) {
          A.r[bt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(are, W.w[bt][kc][s], A.r[bt], 0, 0, 0);
          A.i[bt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(aim, W.w[bt][kc][s], A.i[bt], 0, 0, 0);
        }
        if constexpr (PF) {
          const int lane = h * 32 + n;
          __builtin_amdgcn_sched_barrier(0x386);
#pragma unroll
          for (int bt = 0; bt < NB; bt++) W.load_k(wfrag, wnext, bt, kc, s, lane);
          __builtin_amdgcn_sched_barrier(0x386);
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
        // hi parts of every accumulator, then lo parts: 2*NB - 1 MFMAs sit
        // between two that update the same accumulator
#pragma unroll
        for (int bt = 0; bt < NB; bt++) {
          A.r[bt] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(are, W.hi[bt][kc][s2], A.r[bt], 0, 0, 0, 0, 0, 0);
          A.i[bt] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(aim, W.hi[bt][kc][s2], A.i[bt], 0, 0, 0, 0, 0, 0);
        }
#pragma unroll
        for (int bt = 0; bt < NB; bt++) {
          A.r[bt] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(are, W.lo[bt][kc][s2], A.r[bt], 0, 0, 0, 0, 0, 0);
          A.i[bt] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(aim, W.lo[bt][kc][s2], A.i[bt], 0, 0, 0, 0, 0, 0);
        }
        if constexpr (PF) {
          const int lane = h * 32 + n;
          __builtin_amdgcn_sched_barrier(0x386);
#pragma unroll
          for (int bt = 0; bt < NB; bt++) W.load_k(wfrag, wnext, bt, kc, s2, lane);
          __builtin_amdgcn_sched_barrier(0x386);
        }
      }
    }
  }
  __builtin_amdgcn_s_setprio(0);
}

// |B|^2 summed over the accumulator rows of each bin (four independent FMA
// chains per bin), scaled into the partial power.  Lane (n, h) of a 32x32
// accumulator holds beam column n, rows 8*(v/4) + 4h + v%4.
template <int TAVG>
__device__ __forceinline__ void detect1(const f16v &r, const f16v &i, float sc,
                                        float (&part)[TAVG <= 32 ? 32 / TAVG : 1]) {
  constexpr int NP = TAVG <= 32 ? 32 / TAVG : 1;
  constexpr int VPB = 16 / NP;
#pragma unroll
  for (int j = 0; j < NP; j++) {
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
#pragma unroll
    for (int v = j * VPB; v < (j + 1) * VPB; v += 2) {
      s0 = __builtin_fmaf(r[v], r[v], s0);
      s1 = __builtin_fmaf(i[v], i[v], s1);
      s2 = __builtin_fmaf(r[v + 1], r[v + 1], s2);
      s3 = __builtin_fmaf(i[v + 1], i[v + 1], s3);
    }
    part[j] = __builtin_fmaf((s0 + s1) + (s2 + s3), sc, part[j]);
  }
}

template <int TAVG, int NB>
__device__ __forceinline__ void detect(const Acc<NB> &A, const float (&sc)[NB],
                                       float (&part)[NB][TAVG <= 32 ? 32 / TAVG : 1]) {
#pragma unroll
  for (int bt = 0; bt < NB; bt++) detect1<TAVG>(A.r[bt], A.i[bt], sc[bt], part[bt]);
}

// 4x4 byte transposes: the NCI dwords (each 1 channel x 2 times x 2 pols) of
// 4 antennas (aq*4..aq*4+3) of packet pl -> one dword (4 antennas) in each of
// the 4*NCI staged rows ([channel][pol][sample], rowb bytes apart) they feed.
template <int NCI>
__device__ __forceinline__ void stage_store(uint8_t *stage, const uint32_t (&v)[4][NCI], int pl, int aq,
                                            int rowb) {
  uint8_t *base = stage + aq * 4;
#pragma unroll
  for (int ci = 0; ci < NCI; ci++) {
    const uint32_t x0 = v[0][ci], x1 = v[1][ci], x2 = v[2][ci], x3 = v[3][ci];
    const uint32_t l01 = __builtin_amdgcn_perm(x1, x0, 0x05010400u);
    const uint32_t h01 = __builtin_amdgcn_perm(x1, x0, 0x07030602u);
    const uint32_t l23 = __builtin_amdgcn_perm(x3, x2, 0x05010400u);
    const uint32_t h23 = __builtin_amdgcn_perm(x3, x2, 0x07030602u);
    *(uint32_t *)(base + ((ci * 2 + 0) * TCU + 2 * pl) * rowb) = __builtin_amdgcn_perm(l23, l01, 0x05040100u);
    *(uint32_t *)(base + ((ci * 2 + 1) * TCU + 2 * pl) * rowb) = __builtin_amdgcn_perm(l23, l01, 0x07060302u);
    *(uint32_t *)(base + ((ci * 2 + 0) * TCU + 2 * pl + 1) * rowb) = __builtin_amdgcn_perm(h23, h01, 0x05040100u);
    *(uint32_t *)(base + ((ci * 2 + 1) * TCU + 2 * pl + 1) * rowb) = __builtin_amdgcn_perm(h23, h01, 0x07060302u);
  }
}

// The tile's outputs.  part[k] holds this wave's bins of 32-sample unit
// tt_first + k (k < ntt); lanes h = 0 and h = 1 hold the two row halves of each
// bin.  TAVG 64 has one bin per tile, the sum of both units' halves, so it
// needs ntt = 2.  Each unit's sum per beam (the AGC input) goes to bpart
// [nbeam][nsamp/32][ngroup].
template <int TAVG, int NB>
__device__ __forceinline__ void tile_output(const BfArgs &a, const TileId &t, int btw, int n, int h,
                                            float (&part)[2][NB][TAVG <= 32 ? 32 / TAVG : 1],
                                            int tt_first, int ntt) {
  constexpr int NP = TAVG <= 32 ? 32 / TAVG : 1;
#pragma unroll
  for (int bt = 0; bt < NB; bt++) {
    const int b = (btw + bt) * 32 + n;
    const bool own = h == 0 && b < a.nbeam;
    float x[2][NP];
#pragma unroll
    for (int k = 0; k < 2; k++) {
      if (k >= ntt) continue;
      float unit_sum = 0.f;
#pragma unroll
      for (int j = 0; j < NP; j++) {
        x[k][j] = part[k][bt][j] + __shfl_xor(part[k][bt][j], 32, 64);
        unit_sum += x[k][j];
      }
      const int tt = tt_first + k;
      if (own) a.bpart[((long)b * 2 * a.ntblk + 2 * t.tblk + tt) * a.ngroup + t.g] = unit_sum;
      if (TAVG < 64 && own)
#pragma unroll
        for (int j = 0; j < NP; j++) {
          const int o = t.tblk * (TCU / TAVG) + tt * (32 / TAVG) + j;
          const long idx = out_index(a.layout, a.nout, a.nbeam, a.ngroup, o, b, t.g);
          if (a.power) a.power[idx] = x[k][j];
          a.out[idx] = requant_u8(x[k][j], a.scale[b]);
        }
    }
    if (TAVG == 64 && own) {
      const float y = x[0][0] + x[1][0];
      const long idx = out_index(a.layout, a.nout, a.nbeam, a.ngroup, t.tblk, b, t.g);
      if (a.power) a.power[idx] = y;
      a.out[idx] = requant_u8(y, a.scale[b]);
    }
  }
}

// Host side: fill the tile arguments shared by the fused schedules.
BfArgs bf_make_args(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag,
                    const float *d_wscale, const float *d_scale, uint8_t *d_out, float *d_power,
                    float *d_bpart, int bset_beams);

#endif

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
// Both formats use 128 bytes per lane per (cq, beam tile, antenna chunk).
// The matching A (voltage) slots are built by unpack_f16 / nib_e4m3.
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

#endif

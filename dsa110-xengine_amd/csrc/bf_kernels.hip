// bf_kernels.hip -- gfx950 kernels of the DSA X-engine beamformer.
//
// Reference path: src/dsaX_beamformer.cu (promoter :170-193, beamformer
// :267-416, badder :78-112, adder :120-158, calc_weights :420-449).  The
// reference runs five launches per 32-sample "int" (640 per block) and keeps
// unpacked FP16 planes and FP32 partial powers in HBM between them.  Here one
// fused kernel does, per workgroup tile (one 8-channel output group x TCU
// samples x 256 beams):
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
// What bounds it: every (channel, pol) weight fragment a wave loads feeds
// only that wave's samples, so the weight stream from L2 (fine-channel
// weights: 64 KB per (channel, pol) per 256 beams) sets the pace.  A
// workgroup therefore owns BF_TH x 64 samples with BF_TH waves sharing each
// weight fragment through L1.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include "bf_launch.h"

typedef _Float16 h2v __attribute__((ext_vector_type(2)));
typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef int i8v __attribute__((ext_vector_type(8)));

static thread_local char g_err[256];
extern "C" const char *bfk_last_error(void) { return g_err; }

static int bfk_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return 0;
  snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
  return -3; /* DSAXBF_EHIP */
}

// ---------------------------------------------------------------------------
// Fragment layouts (both written by bf_weights_kernel, read by the fused
// kernel).  Antenna a of 64-antenna chunk kc maps to lane half h = (a%64)/32.
// FP16 (v_mfma_f32_32x32x16_f16, 8 K-slots per lane per step):
//   step s = (a%32)/4, slots [Wr a0, Wr a1, Wi a0, Wi a1, Wr a2, Wr a3, Wi a2,
//   Wi a3] for a_i = 64kc + 32h + 4s + i; one 16-byte record per lane,
//   index (((cq*nbt + bt)*nkc + kc)*8 + s)*64 + lane, cq = wc*2 + pol.
// FP8 (v_mfma_f32_32x32x64_f8f6f4, 32 K-slots per lane per step):
//   step s2 = (a%32)/16, dword 2u = Wr of ants 4u..4u+3 of the 16,
//   dword 2u+1 = Wi; hi and lo parts are separate 32-byte records,
//   index ((((cq*nbt + bt)*nkc + kc)*2 + s2)*2 + hl)*64 + lane.
// The matching A (voltage) slots are built in the fused kernel.
// ---------------------------------------------------------------------------

extern "C" size_t bfk_wfrag_bytes(const bfk_dims *d) {
  // 64 lanes x (8 steps x 16 B) per (cq, bt, kc) for FP16; 64 x (2 x 2 x 32 B) for FP8.
  return (size_t)d->ncal * 2 * d->nbt * d->nkc * 64 * 128;
}
extern "C" size_t bfk_wscale_bytes(const bfk_dims *d) {
  return (size_t)d->ncal * 2 * d->nbeam_pad * sizeof(float);
}

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

// One thread per (weight channel wc, pol, beam): FP64 phase and complex
// product (calc_weights, dsaX_beamformer.cu:420-449), a power-of-two scale
// per (wc, pol, beam) so the largest component sits just below the format's
// top binade, then FP16 or FP8 hi/lo fragments.
__global__ void bf_weights_kernel(bfk_dims d, const float *__restrict__ antpos,
                                  const float *__restrict__ cal,
                                  const double *__restrict__ freqs,
                                  const double *__restrict__ theta,
                                  uint8_t *__restrict__ wfrag,
                                  float *__restrict__ wscale) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long total = (long)d.ncal * 2 * d.nbeam_pad;
  if (idx >= total) return;
  int b = (int)(idx % d.nbeam_pad);
  int q = (int)((idx / d.nbeam_pad) & 1);
  int wc = (int)(idx / (2L * d.nbeam_pad));
  int cq = wc * 2 + q;
  double f = d.wmode ? freqs[wc] : freqs[wc * 8 + 4];
  bool live = b < d.nbeam;
  double ang = live ? theta[b] : 0.0;
  if (d.amode) ang = sin(ang);
  double afac = -2.0 * M_PI * f * ang / 299792458.0;

  double mx = 0.0;
  if (live)
    for (int a = 0; a < d.nant; a++) {
      const float *cw = cal + (((long)a * d.ncal + wc) * 2 + q) * 2;
      double s, c;
      sincos(afac * (double)antpos[a], &s, &c);
      double wr = c * cw[0] - s * cw[1], wi = s * cw[0] + c * cw[1];
      mx = fmax(mx, fmax(fabs(wr), fabs(wi)));
    }
  int k = 0;
  if (mx > 0.0) frexp(mx, &k);
  int e = (d.prec == 0 ? 15 : 8) - k;     // max component in [2^14,2^15) or [128,256)
  double sc = ldexp(1.0, e);
  // power scale that undoes the weight scale (and the FP8 voltage 2^-9)
  wscale[idx] = (float)(d.prec == 0 ? ldexp(1.0, -2 * e) : ldexp(1.0, 18 - 2 * e));

  int bt = b >> 5, n = b & 31;
  for (int kc = 0; kc < d.nkc; kc++)
    for (int h = 0; h < 2; h++) {
      int lane = h * 32 + n;
      double wr[32], wi[32];
      for (int j = 0; j < 32; j++) {
        int a = kc * 64 + h * 32 + j;
        wr[j] = wi[j] = 0.0;
        if (live && a < d.nant) {
          const float *cw = cal + (((long)a * d.ncal + wc) * 2 + q) * 2;
          double s, c;
          sincos(afac * (double)antpos[a], &s, &c);
          wr[j] = (c * cw[0] - s * cw[1]) * sc;
          wi[j] = (s * cw[0] + c * cw[1]) * sc;
        }
      }
      if (d.prec == 0) {
        for (int s = 0; s < 8; s++) {
          h8v v;
          int j = 4 * s;
          v[0] = (_Float16)wr[j];     v[1] = (_Float16)wr[j + 1];
          v[2] = (_Float16)wi[j];     v[3] = (_Float16)wi[j + 1];
          v[4] = (_Float16)wr[j + 2]; v[5] = (_Float16)wr[j + 3];
          v[6] = (_Float16)wi[j + 2]; v[7] = (_Float16)wi[j + 3];
          long rec = (((long)(cq * d.nbt + bt) * d.nkc + kc) * 8 + s) * 64 + lane;
          *(h8v *)(wfrag + rec * 16) = v;
        }
      } else {
        for (int s2 = 0; s2 < 2; s2++) {
          uint32_t hi[8], lo[8];
          for (int u = 0; u < 4; u++) {
            uint32_t hr = 0, hi_i = 0, lr = 0, li = 0;
            for (int i = 0; i < 4; i++) {
              int j = 16 * s2 + 4 * u + i;
              uint32_t br = e4m3_encode(wr[j]), bi = e4m3_encode(wi[j]);
              uint32_t cr = e4m3_encode(wr[j] - e4m3_decode(br));
              uint32_t ci = e4m3_encode(wi[j] - e4m3_decode(bi));
              hr |= br << (8 * i);  hi_i |= bi << (8 * i);
              lr |= cr << (8 * i);  li |= ci << (8 * i);
            }
            hi[2 * u] = hr; hi[2 * u + 1] = hi_i;
            lo[2 * u] = lr; lo[2 * u + 1] = li;
          }
          long rec = ((((long)(cq * d.nbt + bt) * d.nkc + kc) * 2 + s2) * 2) * 64 + lane;
          uint32_t *ph = (uint32_t *)(wfrag + rec * 32);
          uint32_t *pl = (uint32_t *)(wfrag + (rec + 64) * 32);
          for (int u = 0; u < 8; u++) { ph[u] = hi[u]; pl[u] = lo[u]; }
        }
      }
    }
}

// ---------------------------------------------------------------------------
// Voltage unpack to MFMA A-operand slots.
// ---------------------------------------------------------------------------

// FP16: 4 packed samples (4 antennas) -> A_re = [Dr0,Dr1,-Di0,-Di1,Dr2,Dr3,
// -Di2,-Di3] and A_im = [Di0,Di1,Dr0,Dr1,Di2,Di3,Dr2,Dr3].  A nibble n is
// formed exactly as half(0x6400 | (n ^ 8)) - 1032 = n.
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

// FP8: 4 packed samples -> E4M3 bytes of Dr, Di, -Di.  E4M3 byte k (k <= 15,
// sign in bit 7) is exactly k * 2^-9, so a nibble only needs its two's
// complement turned into sign-magnitude: |n| = (s ^ smear) + sign.
__device__ __forceinline__ uint32_t nib_e4m3(uint32_t xs) {
  const uint32_t s = xs & 0x0F0F0F0Fu;
  const uint32_t m = xs & 0x08080808u;     // sign bit of each nibble
  const uint32_t t = m >> 3;               // 1 per negative byte
  const uint32_t sm = (m << 1) - t;        // 0x0F per negative byte
  return ((s ^ sm) + t) | (m << 4);
}

// ---------------------------------------------------------------------------
// Fused beamformer.
// ---------------------------------------------------------------------------
// A workgroup of BF_BQ x BF_TH waves owns one output channel group g,
// BF_TH x 64 samples and BF_BQ x BTW x 32 beams.  Wave w covers beam quarter
// w % BF_BQ and time half w / BF_BQ; it holds the weights of its BTW beam
// tiles for one (channel, pol) in registers and streams its staged rows
// from LDS.  The BF_TH waves of a beam quarter load the same weight
// fragments at about the same time, so L1 serves all but the first.
#ifndef BF_BQ
#define BF_BQ 4          // beam quarters (waves across beams)
#endif
#ifndef BF_TH
#define BF_TH 1          // time halves (waves across samples)
#endif
#ifndef BF_BTW
#define BF_BTW 2         // 32-beam tiles per wave
#endif
#ifndef BF_WPE
#define BF_WPE 2         // waves per SIMD the register budget is sized for
#endif
namespace {
constexpr int TH = BF_TH, BQ = BF_BQ, BTW = BF_BTW;
constexpr int NW = BQ * TH, NT = NW * 64;
constexpr int TCU = 64 * TH;               // samples per tile
constexpr int PCU = TCU / 2;               // packets per tile
constexpr int ROWB = 80;                   // LDS bytes per (chan,pol,sample) row
constexpr int LDS_BYTES = 16 * TCU * ROWB;
constexpr int BSET = BQ * BTW * 32;        // beams per workgroup
}

struct BfArgs {
  const uint8_t *in;
  const uint8_t *wfrag;
  const float *wscale;
  const float *scale;
  uint8_t *out;
  float *power;   // optional
  float *bpart;   // [nbeam][nsamp/64][ngroup] per-64-sample beam power sums
  int nant, nchan, ngroup, nbeam, nbeam_pad, nbt, wmode, order;
  int ntblk, nbsets, nwg_logical;
};

__device__ __forceinline__ uint8_t requant_u8(float x, float scale) {
  // (unsigned char)(__float2int_rn(x*scale)/2), dsaX_beamformer.cu:150-156
  float v = x * scale;
  float r = __builtin_rintf(v);
  int iv;
  if (!(r == r)) iv = 0;
  else if (r >= 2147483648.0f) iv = 2147483647;
  else if (r <= -2147483648.0f) iv = (-2147483647 - 1);
  else iv = (int)r;
  return (uint8_t)(unsigned)(iv / 2);
}

// Weight fragments of one (channel, pol) for the wave's BTW beam tiles, and
// the per-beam power scale that undoes the weight (and FP8 voltage) scaling.
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

// Sum of |B|^2 over the accumulator rows of each bin, two independent FMA
// chains (real, imaginary) per bin, scaled into the partial power.
template <int TAVG>
__device__ __forceinline__ void detect_accumulate(const f16v (&ar)[BTW], const f16v (&ai)[BTW],
                                                  const float (&sc)[BTW],
                                                  float (&part)[BTW][TAVG <= 32 ? 32 / TAVG : 1]) {
  constexpr int NP = TAVG <= 32 ? 32 / TAVG : 1;
  constexpr int VPB = 16 / NP;
#pragma unroll
  for (int bt = 0; bt < BTW; bt++)
#pragma unroll
    for (int j = 0; j < NP; j++) {
      float sr = 0.f, si = 0.f;
#pragma unroll
      for (int v = j * VPB; v < (j + 1) * VPB; v++) {
        sr = __builtin_fmaf(ar[bt][v], ar[bt][v], sr);
        si = __builtin_fmaf(ai[bt][v], ai[bt][v], si);
      }
      part[bt][j] = __builtin_fmaf(sr + si, sc[bt], part[bt][j]);
    }
}

// One (channel, pol): the wave's two 32-sample row tiles against its weights;
// detect and accumulate into the partial powers.
template <int PREC, int TAVG>
__device__ __forceinline__ void beam_cq(const WFrag<PREC> &W, const uint8_t *rows, int n, int h,
                                        float (&part)[TAVG == 64 ? 1 : 2][BTW]
                                                     [TAVG <= 32 ? 32 / TAVG : 1]) {
  constexpr int NTT = TAVG == 64 ? 1 : 2;
  uint4 dv[2][2];
#pragma unroll
  for (int tt = 0; tt < 2; tt++) {
    const uint8_t *src = rows + (tt * 32 + n) * ROWB + h * 32;
    dv[tt][0] = *(const uint4 *)src;
    dv[tt][1] = *(const uint4 *)(src + 16);
  }
#pragma unroll
  for (int tt = 0; tt < 2; tt++) {
    const uint32_t dw[8] = {dv[tt][0].x, dv[tt][0].y, dv[tt][0].z, dv[tt][0].w,
                            dv[tt][1].x, dv[tt][1].y, dv[tt][1].z, dv[tt][1].w};
    f16v ar[BTW], ai[BTW];
#pragma unroll
    for (int bt = 0; bt < BTW; bt++) { ar[bt] = (f16v){}; ai[bt] = (f16v){}; }
    if constexpr (PREC == 0) {
#pragma unroll
      for (int s = 0; s < 8; s++) {
        h8v are, aim;
        unpack_f16(dw[s], are, aim);
#pragma unroll
        for (int bt = 0; bt < BTW; bt++) {
          ar[bt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(are, W.w[bt][s], ar[bt], 0, 0, 0);
          ai[bt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(aim, W.w[bt][s], ai[bt], 0, 0, 0);
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
          ar[bt] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(are, W.hi[bt][s2], ar[bt], 0, 0, 0, 0, 0, 0);
          ar[bt] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(are, W.lo[bt][s2], ar[bt], 0, 0, 0, 0, 0, 0);
          ai[bt] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(aim, W.hi[bt][s2], ai[bt], 0, 0, 0, 0, 0, 0);
          ai[bt] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(aim, W.lo[bt][s2], ai[bt], 0, 0, 0, 0, 0, 0);
        }
      }
    }
    detect_accumulate<TAVG>(ar, ai, W.sc, part[NTT == 1 ? 0 : tt]);
  }
}

template <int PREC, int TAVG>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(BF_WPE, BF_WPE)))
void bf_fused_kernel(BfArgs a) {
  __shared__ __attribute__((aligned(16))) uint8_t lds[LDS_BYTES];
  constexpr int NP = TAVG <= 32 ? 32 / TAVG : 1;   // bins per 32-row tile
  constexpr int NTT = TAVG == 64 ? 1 : 2;          // tile-row groups with own bins

  // XCD-aware tile order: workgroup P runs on XCD P % 8 and logical tiles
  // are dealt so that consecutive L run on one XCD.  Time blocks are the
  // fastest index, so the workgroups resident on an XCD share one group's
  // weights (16 (channel,pol) x 64 KB with fine weights) in its L2.
  const int nwg = gridDim.x;
  const int P = blockIdx.x;
  const int L = (P & 7) * (nwg >> 3) + (P >> 3);
  if (L >= a.nwg_logical) return;
  const int tblk = L % a.ntblk;
  const int rest = L / a.ntblk;
  const int g = rest % a.ngroup;
  const int bset = rest / a.ngroup;

  const int tid = threadIdx.x;
  const int p0 = tblk * PCU;

  // (a) stage: 32-byte chunk per (packet, antenna) = 8 channels x 2 t x 2 pol,
  // four antennas per task, 4x4 byte transposes, one dword per LDS row.
  for (int task = tid; task < PCU * 16; task += NT) {
    const int pl = task >> 4, aq = task & 15;
    const int p = p0 + pl;
    uint32_t x[4][8];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int ant = aq * 4 + j;
      if (ant < a.nant) {
        long off = a.order == 0
                       ? (((long)p * a.nant + ant) * a.nchan + g * 8) * 4
                       : (((long)p * a.ngroup + g) * a.nant + ant) * 32;
        const uint4 *src = (const uint4 *)(a.in + off);
        uint4 v0 = src[0], v1 = src[1];
        x[j][0] = v0.x; x[j][1] = v0.y; x[j][2] = v0.z; x[j][3] = v0.w;
        x[j][4] = v1.x; x[j][5] = v1.y; x[j][6] = v1.z; x[j][7] = v1.w;
      } else {
#pragma unroll
        for (int w = 0; w < 8; w++) x[j][w] = 0u;
      }
    }
#pragma unroll
    for (int ci = 0; ci < 8; ci++) {
      const uint32_t l01 = __builtin_amdgcn_perm(x[1][ci], x[0][ci], 0x05010400u);
      const uint32_t h01 = __builtin_amdgcn_perm(x[1][ci], x[0][ci], 0x07030602u);
      const uint32_t l23 = __builtin_amdgcn_perm(x[3][ci], x[2][ci], 0x05010400u);
      const uint32_t h23 = __builtin_amdgcn_perm(x[3][ci], x[2][ci], 0x07030602u);
      uint8_t *base = lds + aq * 4;
      *(uint32_t *)(base + ((ci * 2 + 0) * TCU + 2 * pl) * ROWB) = __builtin_amdgcn_perm(l23, l01, 0x05040100u);
      *(uint32_t *)(base + ((ci * 2 + 1) * TCU + 2 * pl) * ROWB) = __builtin_amdgcn_perm(l23, l01, 0x07060302u);
      *(uint32_t *)(base + ((ci * 2 + 0) * TCU + 2 * pl + 1) * ROWB) = __builtin_amdgcn_perm(h23, h01, 0x05040100u);
      *(uint32_t *)(base + ((ci * 2 + 1) * TCU + 2 * pl + 1) * ROWB) = __builtin_amdgcn_perm(h23, h01, 0x07060302u);
    }
  }
  __syncthreads();

  const int lane = tid & 63, wave = tid >> 6;
  const int n = lane & 31, h = lane >> 5;
  const int bq = wave % BQ, th = wave / BQ;
  const int btw = bset * (BSET / 32) + bq * BTW;   // first beam tile of this wave

  float part[NTT][BTW][NP];
#pragma unroll
  for (int i = 0; i < NTT; i++)
#pragma unroll
    for (int bt = 0; bt < BTW; bt++)
#pragma unroll
      for (int j = 0; j < NP; j++) part[i][bt][j] = 0.f;

  // (b) 16 (channel, pol) GEMMs; cq = 2*channel + pol indexes the staged rows.
  for (int cq = 0; cq < 16; cq++) {
    const long wr = (long)((a.wmode ? g * 8 + (cq >> 1) : g) * 2 + (cq & 1));
    WFrag<PREC> W;
    W.load(a.wfrag, a.wscale, wr * a.nbt + btw, wr * a.nbeam_pad + btw * 32, lane, n);
    beam_cq<PREC, TAVG>(W, lds + (cq * TCU + th * 64) * ROWB, n, h, part);
  }

  // (c) lanes h=0 and h=1 hold the two row halves of each bin.  Besides the
  // output, each wave leaves the sum of its bins per beam (the AGC input,
  // badder in the reference) in bpart[beam][64-sample block][group].
  const int t64 = tblk * TH + th;                  // this wave's 64-sample block
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
          const int o = TAVG == 64 ? t64 : t64 * (64 / TAVG) + tt * (32 / TAVG) + j;
          const long idx = ((long)o * a.nbeam + b) * a.ngroup + g;
          if (a.power) a.power[idx] = x;
          a.out[idx] = requant_u8(x, a.scale[b]);
        }
      }
    if (h == 0 && b < a.nbeam) a.bpart[((long)b * (a.ntblk * TH) + t64) * a.ngroup + g] = tile_sum;
  }
}

// ---------------------------------------------------------------------------
// -s / -q outputs: one workgroup per (64-sample block, group), one thread
// per antenna sums |d|^2 per output bin over the block's packets, 8 channels,
// 2 times and 2 pols; integer-valued sums are exact in float.
//  stuff 1: beam b <- antenna ((b/16)%4)*16 + b%16  (dsaX_beamformer.cu:326-337)
//  stuff 2: beam 0 <- sum over all antennas          (dsaX_beamformer.cu:307-323)
// The per-block beam sums in bpart are rewritten for the beams written here.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void bf_stuff_kernel(bfk_dims d, const uint8_t *__restrict__ in,
                                                       const float *__restrict__ scale,
                                                       uint8_t *__restrict__ out,
                                                       float *__restrict__ power,
                                                       float *__restrict__ bpart) {
  __shared__ float S[256][8];
  __shared__ float IB[8];
  const int nblk = d.nsamp / 64;
  const int t64 = blockIdx.x % nblk, g = blockIdx.x / nblk;
  const int nbin = 64 / d.tavg < 1 ? 1 : 64 / d.tavg;
  const int ppb = d.tavg / 2 < 32 ? d.tavg / 2 : 32;      // packets per bin
  const int tid = threadIdx.x;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (tid < d.nant) {
    for (int pl = 0; pl < 32; pl++) {
      const long p = (long)t64 * 32 + pl;
      const long off = d.order == 0 ? ((p * d.nant + tid) * d.nchan + g * 8) * 4
                                    : ((p * d.ngroup + g) * d.nant + tid) * 32;
      const uint4 *src = (const uint4 *)(in + off);
      const uint4 v0 = src[0], v1 = src[1];
      const uint32_t w[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < 8; k++)
#pragma unroll
        for (int by = 0; by < 4; by++) {
          const uint32_t bv = (w[k] >> (8 * by)) & 0xFFu;
          const int lo = bv & 15, hi = bv >> 4;
          const int re = lo >= 8 ? lo - 16 : lo, im = hi >= 8 ? hi - 16 : hi;
          acc += (float)(re * re + im * im);
        }
      const int bin = pl / ppb;
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (k == bin) s[k] += acc;
    }
  }
#pragma unroll
  for (int k = 0; k < 8; k++) S[tid][k] = s[k];
  __syncthreads();
  if (tid < nbin) {
    float x = 0.f;
    for (int a = 0; a < d.nant; a++) x += S[a][tid];
    IB[tid] = x;
  }
  __syncthreads();
  const int bend = d.stuff == 2 ? 1 : d.nbeam;
  for (int b = tid; b < bend; b += 256) {
    const int ant = ((b / 16) % 4) * 16 + b % 16;
    float tile = 0.f;
    for (int k = 0; k < nbin; k++) {
      const float x = d.stuff == 2 ? IB[k] : (ant < d.nant ? S[ant][k] : 0.f);
      const long idx = ((long)(t64 * nbin + k) * d.nbeam + b) * d.ngroup + g;
      if (power) power[idx] = x;
      out[idx] = requant_u8(x, scale[b]);
      tile += x;
    }
    bpart[((long)b * nblk + t64) * d.ngroup + g] = tile;
  }
}

// beam_power[b] = sum over 64-sample blocks and groups of bpart[b][.][.]
// (badder + calc_allbp, dsaX_beamformer.cu:78-112,482-495).  One workgroup
// per beam, coalesced reads, fixed-order FP64 tree: deterministic.
__global__ __launch_bounds__(256) void bf_beam_power_kernel(bfk_dims d, const float *__restrict__ bpart,
                                                            float *__restrict__ bp) {
  __shared__ double red[256];
  const int b = blockIdx.x, tid = threadIdx.x;
  const long n = (long)(d.nsamp / 64) * d.ngroup;
  const float *src = bpart + (long)b * n;
  double s = 0.0;
  for (long i = tid; i < n; i += 256) s += src[i];
  red[tid] = s;
  __syncthreads();
  for (int m = 128; m > 0; m >>= 1) {
    if (tid < m) red[tid] += red[tid + m];
    __syncthreads();
  }
  if (tid == 0) bp[b] = (float)red[0];
}

// ---------------------------------------------------------------------------
// Launchers (C ABI).
// ---------------------------------------------------------------------------
extern "C" int bfk_weights(const bfk_dims *d, const float *d_antpos, const float *d_cal,
                           const double *d_freqs, const double *d_theta, void *d_wfrag,
                           float *d_wscale, void *stream) {
  long total = (long)d->ncal * 2 * d->nbeam_pad;
  int nblk = (int)((total + 127) / 128);
  hipLaunchKernelGGL(bf_weights_kernel, dim3(nblk), dim3(128), 0, (hipStream_t)stream, *d,
                     d_antpos, d_cal, d_freqs, d_theta, (uint8_t *)d_wfrag, d_wscale);
  return bfk_check(hipGetLastError(), "bf_weights_kernel");
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
    snprintf(g_err, sizeof(g_err), "bfk_beamform: unsupported shape nant_pad=%d nbeam_pad=%d nsamp=%d",
             d->nant_pad, d->nbeam_pad, d->nsamp);
    return -1;
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
  a.nbsets = d->nbeam_pad / BSET;
  a.nwg_logical = a.nbsets * d->ngroup * a.ntblk;
  const int nwg = (a.nwg_logical + 7) / 8 * 8;
  hipStream_t s = (hipStream_t)stream;
  if (d->prec == 0) launch_fused<0>(d->tavg, dim3(nwg), s, a);
  else launch_fused<1>(d->tavg, dim3(nwg), s, a);
  return bfk_check(hipGetLastError(), "bf_fused_kernel");
}

extern "C" int bfk_stuff(const bfk_dims *d, const uint8_t *d_in, const float *d_scale,
                         uint8_t *d_out, float *d_power, float *d_bpart, void *stream) {
  if (d->nant > 256 || d->nsamp % 64 != 0) {
    snprintf(g_err, sizeof(g_err), "bfk_stuff: unsupported shape nant=%d nsamp=%d", d->nant, d->nsamp);
    return -1;
  }
  hipLaunchKernelGGL(bf_stuff_kernel, dim3((d->nsamp / 64) * d->ngroup), dim3(256), 0,
                     (hipStream_t)stream, *d, d_in, d_scale, d_out, d_power, d_bpart);
  return bfk_check(hipGetLastError(), "bf_stuff_kernel");
}

extern "C" size_t bfk_bpart_bytes(const bfk_dims *d) {
  return (size_t)d->nbeam * (d->nsamp / 64) * d->ngroup * sizeof(float);
}

extern "C" int bfk_beam_power(const bfk_dims *d, const float *d_bpart, float *d_beam_power,
                              void *stream) {
  hipLaunchKernelGGL(bf_beam_power_kernel, dim3(d->nbeam), dim3(256), 0, (hipStream_t)stream, *d,
                     d_bpart, d_beam_power);
  return bfk_check(hipGetLastError(), "bf_beam_power_kernel");
}

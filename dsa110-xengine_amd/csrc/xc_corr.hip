// xc_corr.hip -- the correlator X-engine on the beamformer's 4-bit input:
// for every channel and pol, the visibility of each antenna pair (i, j <= i)
// summed over the block, on int8 matrix cores.
//
// Reference: src/dsaX_bfCorr.cu dcorrelator (:472-569): the block
// [packet][antenna][channel][2 times][2 pols] is transposed and unpacked to
// FP16 planes in HBM (:316-356), four strided batched FP16 GEMMs per
// (channel, time, pol) form C = r r^T + i i^T and C' = r i^T - i r^T with FP16
// accumulation (:512-556), and two more kernels transpose and gather the
// upper triangle into [baseline][channel][pol][re, im] (:366-457).
//
// Here:
//   - xc_turn_kernel turns the packed block, still one byte (4-bit re, 4-bit
//     im) per sample, into (channel, pol) planes [channel*2 + pol][antenna]
//     [sample]: each lane reads one channel's 4-byte words of 64 packets
//     (32 lanes cover a 128-byte run of 32 channels) and writes its 128-byte
//     row segment.
//   - xc_kernel: one wave per ((channel, pol), share of the triangle, time
//     slice); waves are independent (no LDS, no barriers).  Per 64 samples a
//     lane loads 32 contiguous bytes of each antenna block its tiles read.
//   - Operands are int8: a 4-bit sample v is taken as u = v + 8 = nibble ^ 8
//     in 0..15 (masks, no conversion), so the matrix op sums products of
//     small non-negative integers in int32, exactly.  Lane (n, h) of an
//     operand holds antenna 32J + n and 8 samples as 16 K slots X = [ur | ui];
//     W = [ui | -ur] (-u is representable).  Re' = X_j . X_i and
//     Im' = X_j . W_i; with the per-antenna sums S = sum (ur + ui) and
//     E = sum (ur - ui) over the N samples, the visibilities of v = u - 8 are
//     Re V(j, i) = Re' - 8 S_j - 8 S_i + 128 N and
//     Im V(j, i) = Im' - 8 E_j + 8 E_i, integers, exact in FP32 below 2^24.
//   - With several time slices each slice writes its partial triangle and
//     xc_reduce_kernel adds them (exact).
#include "bf_common.h"

namespace {
constexpr int XU = 64;                    // samples per unit
constexpr int XUP = XU / 2;               // packets per unit
constexpr int TCH = 32;                   // channels per turn workgroup (one 128-byte run)
constexpr int XPD = 2;                    // units of voltages in flight per wave

struct XcArgs {
  const uint8_t *in;                      // (channel, pol) planes [channel*2 + pol][antenna][sample]
  float *out;                             // visibilities, or slice partials
  int nant, nchan, npkt, nslice;
  uint32_t in_bytes;
  long slice_floats;
};
typedef int i4v __attribute__((ext_vector_type(4)));
typedef int i16v __attribute__((ext_vector_type(16)));
typedef uint32_t u4v __attribute__((ext_vector_type(4)));
}  // namespace

// The block to (channel, pol) planes, unit-blocked: plane pq = channel*2 + pol,
// unit u (64 samples), antenna a: 64 bytes at ((pq * nunit + u) * nant + a) * 64.
// Workgroup = (unit, 16 antennas, 32 channels).  (1) Its 512 (packet, antenna)
// 128-byte runs of 32 channels, 16 bytes a lane, go to LDS as words
// [antenna][channel][packet] (row pitch 34 words: a wave's stores, lanes over
// (16-byte piece, packet), fall on 64 distinct banks).  (2) Each lane builds
// 16 bytes (8 packets) of one (channel, pol) of one antenna with byte
// permutes and stores them; four lanes cover an antenna's 64-byte segment.
constexpr int TA16 = 16, TPS = 34;
__global__ __launch_bounds__(256) void xc_turn_kernel(const uint8_t *__restrict__ in, uint8_t *__restrict__ turn,
                                                      int nant, int nchan, int npkt, int order, uint32_t in_bytes) {
  __shared__ uint32_t lds[TA16 * TCH * TPS];
  const int ncb = (nchan + TCH - 1) / TCH, nunit = npkt / XUP;
  const long w = blockIdx.x;
  const int u = (int)(w % nunit), cb = (int)((w / nunit) % ncb), ab = (int)(w / ((long)nunit * ncb));
  const int tid = threadIdx.x;
  const int ngroup = nchan / 8;
  const __amdgpu_buffer_rsrc_t rin = buffer_rsrc(in, in_bytes);
  // (1) piece i = tid + 256k: quad = i & 7 (channels 4 quad .. + 3), packet
  // (i >> 3 & 7) + 8 (i >> 10), antenna (i >> 6) & 15
  uint4 v[16];
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const int i = tid + 256 * k, quad = i & 7, pl = ((i >> 3) & 7) + 8 * (i >> 10), al = (i >> 6) & 15;
    const int c = cb * TCH + 4 * quad, ant = ab * TA16 + al, p = u * XUP + pl;
    const uint32_t off = c < nchan && ant < nant
                             ? (uint32_t)(chunk_offset(order, p, ant, c >> 3, nant, nchan, ngroup) + (c & 7) * 4)
                             : 0x80000000u;
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    const u4 x = __builtin_amdgcn_raw_buffer_load_b128(rin, off, 0, 0);
    v[k] = make_uint4(x.x, x.y, x.z, x.w);
  }
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const int i = tid + 256 * k, quad = i & 7, pl = ((i >> 3) & 7) + 8 * (i >> 10), al = (i >> 6) & 15;
    uint32_t *r = lds + (al * TCH + 4 * quad) * TPS + pl;
    r[0] = v[k].x;
    r[TPS] = v[k].y;
    r[2 * TPS] = v[k].z;
    r[3 * TPS] = v[k].w;
  }
  __syncthreads();
  // (2) piece i = tid + 256k: 16 bytes (packets 8 (i & 3) .. + 7) of channel
  // (i >> 2) & 31, pol (i >> 7) & 1, antenna k; word (p, a, c) = bytes
  // [t0 q0, t0 q1, t1 q0, t1 q1]
#pragma unroll 4
  for (int k = 0; k < 16; k++) {
    const int i = tid + 256 * k, pc = i & 3, cl = (i >> 2) & 31, q = (i >> 7) & 1;
    const int c = cb * TCH + cl, ant = ab * TA16 + k;
    if (c >= nchan || ant >= nant) continue;
    const uint32_t *r = lds + (k * TCH + cl) * TPS + 8 * pc;
    const uint2 w0 = *(const uint2 *)r, w1 = *(const uint2 *)(r + 2), w2 = *(const uint2 *)(r + 4),
                w3 = *(const uint2 *)(r + 6);
    const uint32_t sel = q ? 0x07050301u : 0x06040200u;
    const uint4 o = make_uint4(__builtin_amdgcn_perm(w0.y, w0.x, sel), __builtin_amdgcn_perm(w1.y, w1.x, sel),
                               __builtin_amdgcn_perm(w2.y, w2.x, sel), __builtin_amdgcn_perm(w3.y, w3.x, sel));
    *(uint4 *)(turn + ((long)((2 * c + q) * nunit + u) * nant + ant) * XU + 16 * pc) = o;
  }
}

// The triangle's 32x32 tiles (J <= I) in column-major order, (0,0), (0,1),
// (1,1), (0,2), ..., and the share of them that wave part P of NS takes: in
// this order the first shares need fewer antenna blocks (96 antennas, two
// shares: blocks 0-1 with two W operands, and blocks 0-2 with one, where
// row-major order gives 3 + 3 and 2 + 2).
namespace {
template <int NB> constexpr int ntile() { return NB * (NB + 1) / 2; }
template <int NB> constexpr int tile_j(int t) {
  int k = 0;
  for (int I = 0; I < NB; I++)
    for (int J = 0; J <= I; J++, k++)
      if (k == t) return J;
  return -1;
}
template <int NB> constexpr int tile_i(int t) {
  int k = 0;
  for (int I = 0; I < NB; I++)
    for (int J = 0; J <= I; J++, k++)
      if (k == t) return I;
  return -1;
}
template <int NB, int NS, int P> constexpr int tile_lo() { return P * ntile<NB>() / NS; }
template <int NB, int NS, int P> constexpr int tile_hi() { return (P + 1) * ntile<NB>() / NS; }
template <int NB, int NS> constexpr int tiles_max() {
  int m = 0;
  for (int P = 0; P < NS; P++) {
    const int n = (P + 1) * ntile<NB>() / NS - P * ntile<NB>() / NS;
    m = n > m ? n : m;
  }
  return m;
}
// block b is read by a tile of the share: as a row block (X) or a column block (X and W)
template <int NB, int NS, int P> constexpr bool reads(int b) {
  for (int t = tile_lo<NB, NS, P>(); t < tile_hi<NB, NS, P>(); t++)
    if (tile_j<NB>(t) == b || tile_i<NB>(t) == b) return true;
  return false;
}
template <int NB, int NS, int P> constexpr bool is_col(int b) {
  for (int t = tile_lo<NB, NS, P>(); t < tile_hi<NB, NS, P>(); t++)
    if (tile_i<NB>(t) == b) return true;
  return false;
}

// Share P of one (channel, pol) plane over units [ub, ub + nunit).
template <int NB, int NS, int P, int PD>
__device__ __forceinline__ void xc_share(const XcArgs &a, int pq, int ub, int nunit, float *out, int lane) {
  constexpr int T0 = tile_lo<NB, NS, P>(), T1 = tile_hi<NB, NS, P>(), NTW = tiles_max<NB, NS>();
  const int n = lane & 31, h = lane >> 5;
  const int nsamp = 2 * a.npkt;
  const __amdgpu_buffer_rsrc_t rin = buffer_rsrc(a.in, a.in_bytes);
  // lane part of each block's offset in a unit (sentinel for padding
  // antennas: zeros); unit u of the plane starts at (pq * nunit + u) * nant * 64
  const int ntot = nsamp / XU;
  uint32_t off[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const int ant = 32 * b + n;
    off[b] = ant < a.nant ? (uint32_t)(ant * XU + 32 * h) : 0x80000000u;
  }
  // (units past the slice read at offset in_bytes and up, which the buffer's
  // range check answers with zeros: every pass of the loop below issues the
  // same loads, so each unit waits only for its own)
  auto load = [&](int u, uint32_t (&r)[NB][8]) {
    const uint32_t so = (uint32_t)__builtin_amdgcn_readfirstlane(
        u < nunit ? (int)(((long)pq * ntot + ub + u) * a.nant * XU) : (int)a.in_bytes);
#pragma unroll
    for (int b = 0; b < NB; b++) {
      if (!reads<NB, NS, P>(b)) continue;
      const u4v v0 = __builtin_amdgcn_raw_buffer_load_b128(rin, off[b], so, 0);
      const u4v v1 = __builtin_amdgcn_raw_buffer_load_b128(rin, off[b] + 16, so, 0);
      r[b][0] = v0.x; r[b][1] = v0.y; r[b][2] = v0.z; r[b][3] = v0.w;
      r[b][4] = v1.x; r[b][5] = v1.y; r[b][6] = v1.z; r[b][7] = v1.w;
    }
  };
  i16v accr[NTW], acci[NTW];
#pragma unroll
  for (int t = 0; t < NTW; t++) { accr[t] = (i16v){}; acci[t] = (i16v){}; }
  uint32_t sr[NB], si[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) sr[b] = si[b] = 0;

  // four 16-sample K steps per unit: lane half h takes samples 32h + 8s .. + 7
  auto compute = [&](const uint32_t (&r)[NB][8]) {
#pragma unroll
    for (int st = 0; st < 4; st++) {
      i4v X[NB], W[NB];
#pragma unroll
      for (int b = 0; b < NB; b++) {
        if (!reads<NB, NS, P>(b)) continue;
        const uint32_t y0 = r[b][2 * st] ^ 0x88888888u, y1 = r[b][2 * st + 1] ^ 0x88888888u;
        const uint32_t ur0 = y0 & 0x0F0F0F0Fu, ur1 = y1 & 0x0F0F0F0Fu;
        const uint32_t ui0 = (y0 >> 4) & 0x0F0F0F0Fu, ui1 = (y1 >> 4) & 0x0F0F0F0Fu;
        sr[b] = __builtin_amdgcn_sad_u8(ur0, 0u, sr[b]);
        sr[b] = __builtin_amdgcn_sad_u8(ur1, 0u, sr[b]);
        si[b] = __builtin_amdgcn_sad_u8(ui0, 0u, si[b]);
        si[b] = __builtin_amdgcn_sad_u8(ui1, 0u, si[b]);
        X[b] = (i4v){(int)ur0, (int)ur1, (int)ui0, (int)ui1};
        if (is_col<NB, NS, P>(b)) {
          // -u bytewise, borrow-free: (0x80 - u) ^ 0x80 for u in 0..15
          const uint32_t nr0 = (0x80808080u - ur0) ^ 0x80808080u, nr1 = (0x80808080u - ur1) ^ 0x80808080u;
          W[b] = (i4v){(int)ui0, (int)ui1, (int)nr0, (int)nr1};
        }
      }
#pragma unroll
      for (int t = T0; t < T1; t++) {
        const int J = tile_j<NB>(t), I = tile_i<NB>(t);
        accr[t - T0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(X[J], X[I], accr[t - T0], 0, 0, 0);
        acci[t - T0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(X[J], W[I], acci[t - T0], 0, 0, 0);
      }
    }
  };

  uint32_t r[PD][NB][8];
  // in unit order, as the loop issues them
#pragma unroll
  for (int k = 0; k < PD; k++) {
    load(k, r[k]);
    __builtin_amdgcn_sched_barrier(0);
  }
  // whole rounds of PD units, then the rest
  int u0 = 0;
  for (; u0 + PD <= nunit; u0 += PD) {
#pragma unroll
    for (int k = 0; k < PD; k++) {
      compute(r[k]);
      load(u0 + k + PD, r[k]);
      // (the loads stay here: sunk below the next units' work, they would be
      // waited for with those units')
      __builtin_amdgcn_sched_barrier(0);
    }
  }
#pragma unroll
  for (int k = 0; k < PD - 1; k++)
    if (u0 + k < nunit) compute(r[k]);

  // per-antenna sums over both lane halves: S = sum (ur + ui), E = sum (ur - ui)
  int S[NB], E[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) {
    S[b] = E[b] = 0;
    if (!reads<NB, NS, P>(b)) continue;
    const int s0 = (int)(sr[b] + si[b]), e0 = (int)sr[b] - (int)si[b];
    S[b] = s0 + __shfl_xor(s0, 32, 64);
    E[b] = e0 + __shfl_xor(e0, 32, 64);
  }
  const int N128 = 128 * nunit * XU;
  // lane (n, h) of tile (J, I) holds column i = 32I + n and rows
  // j = 32J + 8(v/4) + 4h + v%4
  const int pol = pq & 1, ch = pq >> 1;
#pragma unroll
  for (int t = T0; t < T1; t++) {
    const int J = tile_j<NB>(t), I = tile_i<NB>(t);
    const int ii = 32 * I + n;
#pragma unroll
    for (int v = 0; v < 16; v++) {
      const int jl = 8 * (v >> 2) + 4 * h + (v & 3), jj = 32 * J + jl;
      const int Sj = __shfl(S[J], jl, 64), Ej = __shfl(E[J], jl, 64);
      if (ii < a.nant && jj <= ii) {
        const long bl = (long)ii * (ii + 1) / 2 + jj;
        const int re = accr[t - T0][v] - 8 * Sj - 8 * S[I] + N128;
        const int im = acci[t - T0][v] - 8 * Ej + 8 * E[I];
        *(float2 *)(out + ((bl * a.nchan + ch) * 2 + pol) * 2) = make_float2((float)re, (float)im);
      }
    }
  }
}
}  // namespace

// Share `part` of NS (a wave-uniform branch to the share's own code).
template <int NB, int NS, int P, int PD>
__device__ __forceinline__ void xc_share_of(int part, const XcArgs &a, int pq, int ub, int nunit, float *out,
                                            int lane) {
  if constexpr (P + 1 < NS) {
    if (part != P) {
      xc_share_of<NB, NS, P + 1, PD>(part, a, pq, ub, nunit, out, lane);
      return;
    }
  }
  xc_share<NB, NS, P, PD>(a, pq, ub, nunit, out, lane);
}

// NB antenna blocks of 32, NS waves per (channel, pol) plane, PD units of
// voltages in flight.  Wave g of the grid: share g % NS of plane (g / NS) %
// nplane of slice g / (NS nplane); the NS shares of a plane are neighbours in
// one workgroup or the next (same XCD: blockIdx is spread round-robin, so
// workgroup P takes work item (P % 8) * (grid / 8) + P / 8).
template <int NB, int NS, int PD>
__global__ __launch_bounds__(256) void xc_kernel(XcArgs a) {
  const int P = blockIdx.x;
  const int L = (P & 7) * (gridDim.x >> 3) + (P >> 3);
  const int lane = threadIdx.x & 63;
  // wave-uniform (a scalar): the plane's offset goes in the loads' scalar offset
  const int gw = L * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int nplane = 2 * a.nchan;
  // (readfirstlane: the runtime division leaves its result in vector registers)
  const int part = gw % NS;
  const int pq = __builtin_amdgcn_readfirstlane((gw / NS) % nplane);
  const int slice = __builtin_amdgcn_readfirstlane(gw / (NS * nplane));
  if (slice >= a.nslice) return;          // whole waves; nothing below synchronises
  const int ntot = a.npkt / XUP;
  const int ub = (int)((long)ntot * slice / a.nslice), nunit = (int)((long)ntot * (slice + 1) / a.nslice) - ub;
  float *out = a.out + slice * a.slice_floats;
  xc_share_of<NB, NS, 0, PD>(part, a, pq, ub, nunit, out, lane);
}

extern "C" size_t bfk_xcorr_vis_bytes(int nant, int nchan) {
  return (size_t)nant * (nant + 1) / 2 * nchan * 2 * 2 * sizeof(float);
}

// Sum of the slice partials (integer-valued floats: exact in any order).
__global__ __launch_bounds__(256) void xc_reduce_kernel(const float4 *__restrict__ part, float4 *__restrict__ vis,
                                                        long n4, int nslice) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n4) return;
  float4 s = part[i];
  for (int k = 1; k < nslice; k++) {
    const float4 v = part[k * n4 + i];
    s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
  }
  vis[i] = s;
}

// Slices of the block's time for a shape: one when the (channel, pol) planes
// alone give about a wave per SIMD (2 nchan planes x 2 waves >= 1024; measured
// on the 384-channel block: 1, 2, 4 and 8 slices took 0.235, 0.254, 0.299 and
// 0.409 ms: the partial sums' traffic outweighs the extra waves), else enough
// for that, with at least 8 units (512 samples) per slice.
extern "C" int bfk_xcorr_default_slices(int nchan, int npkt) {
  const int nunit = npkt / XUP;
  int s = nchan >= 256 ? 1 : (256 + nchan - 1) / nchan;
  if (s > nunit / 8) s = nunit / 8;
  return s < 1 ? 1 : s;
}

extern "C" size_t bfk_xcorr_part_bytes(int nant, int nchan, int nslice) {
  return nslice > 1 ? (size_t)nslice * bfk_xcorr_vis_bytes(nant, nchan) : 0;
}

// d_turn: the input's size; d_part: bfk_xcorr_part_bytes (nslice > 1).
extern "C" int bfk_xcorr(int nant, int nchan, int npkt, int order, const uint8_t *d_in, uint8_t *d_turn,
                         float *d_part, int nslice, float *d_vis, void *stream) {
  if (nant < 1 || nant > 96 || nchan < 8 || nchan % 8 || npkt < XUP || npkt % XUP || order < 0 || order > 1) {
    bfk_set_error("bfk_xcorr: unsupported shape nant=%d nchan=%d npkt=%d order=%d", nant, nchan, npkt, order);
    return -1;
  }
  if ((size_t)npkt * nant * nchan * 4 > 0x7fffffffu) {
    bfk_set_error("bfk_xcorr: block above 2 GiB");
    return -1;
  }
  if (nslice < 1 || nslice > npkt / XUP || (nslice > 1 && !d_part) || !d_turn) {
    bfk_set_error("bfk_xcorr: %d slices of %d units (partials %p)", nslice, npkt / XUP, (void *)d_part);
    return -1;
  }
  hipStream_t s = (hipStream_t)stream;
  const uint32_t in_bytes = (uint32_t)((size_t)npkt * nant * nchan * 4);
  const long nturn = (long)(npkt / XUP) * ((nchan + TCH - 1) / TCH) * ((nant + TA16 - 1) / TA16);
  hipLaunchKernelGGL(xc_turn_kernel, dim3((unsigned)nturn), dim3(256), 0, s, d_in, d_turn, nant, nchan, npkt, order,
                     in_bytes);
  XcArgs a;
  a.in = d_turn;
  a.out = nslice > 1 ? d_part : d_vis;
  a.nant = nant;
  a.nchan = nchan;
  a.npkt = npkt;
  a.nslice = nslice;
  a.in_bytes = in_bytes;
  a.slice_floats = (long)(bfk_xcorr_vis_bytes(nant, nchan) / sizeof(float));
  // two waves per plane above 32 antennas (measured on the 96-antenna block,
  // whole block: one, two, three and six waves per plane took 0.242, 0.238,
  // 0.249 and 0.259 ms; more waves per SIMD do not pay for the duplicated
  // loads and conversions.  With each unit waiting for its own loads: two
  // waves 0.222 ms, three 0.252 ms, or 0.266 ms held to 168 registers for
  // three waves per SIMD; XPD = 3 0.229 ms)
  const int ns = nant <= 32 ? 1 : 2;
  const long nwave = (long)ns * 2 * nchan * nslice;
  const dim3 grid((unsigned)(((nwave + 3) / 4 + 7) / 8 * 8)), block(256);
  if (nant <= 32) hipLaunchKernelGGL((xc_kernel<1, 1, XPD>), grid, block, 0, s, a);
  else if (nant <= 64) hipLaunchKernelGGL((xc_kernel<2, 2, XPD>), grid, block, 0, s, a);
  else hipLaunchKernelGGL((xc_kernel<3, 2, XPD>), grid, block, 0, s, a);
  if (nslice > 1) {
    const long n4 = a.slice_floats / 4;
    hipLaunchKernelGGL(xc_reduce_kernel, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, s,
                       (const float4 *)d_part, (float4 *)d_vis, n4, nslice);
  }
  return bfk_check(hipGetLastError(), "xc_kernel");
}

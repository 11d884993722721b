// xc_corr.hip -- the correlator X-engine on the beamformer's 4-bit input:
// for every channel and pol, the visibility of each antenna pair (i, j <= i)
// summed over the block, on FP6 matrix cores.
//
// Reference: src/dsaX_bfCorr.cu dcorrelator (:472-569): the block
// [packet][antenna][channel][2 times][2 pols] is transposed and unpacked to
// FP16 planes in HBM (:316-356), four strided batched FP16 GEMMs per
// (channel, time, pol) form C = r r^T + i i^T and C' = r i^T - i r^T with FP16
// accumulation (:512-556), and two more kernels transpose and gather the
// upper triangle into [baseline][channel][pol][re, im] (:366-457).
//
// Here:
//   - xc_turn_kernel turns the packed block, still 4-bit, into channel-pair
//     planes [group*4 + pair][packet][antenna][8 bytes] (2 channels x 2 times
//     x 2 pols per antenna and packet) through LDS: reads are whole channel
//     runs of one (packet, antenna), writes whole antenna runs of one pair.
//   - xc_kernel: workgroup = one channel pair (2 channels x 2 pols = 4
//     (channel, pol), one wave each) of one 8-channel group over one time
//     slice of the block, so each antenna pair's sum is finished in registers.
//     Its loads read one pair plane, consecutive lanes on consecutive
//     antennas (a wave covers 512 contiguous bytes).  With one slice the
//     triangle goes straight to the visibilities; with more, each slice
//     writes its partial triangle and xc_reduce_kernel adds them (exact:
//     integer-valued floats below 2^24).
//   - Per unit of 64 samples, every thread loads 4 packets x 8 bytes of one
//     antenna (buffer loads: padding antennas read zeros) and byte-permutes
//     them into per-(channel, pol) rows [antenna][64 samples] in LDS (double
//     buffered; 80-byte rows, so the 16-byte operand reads are
//     conflict-free).  The next units' loads are in flight during this one's
//     matrix work.
//   - K = time.  Lane (n, h) of an operand holds antenna 32J + n and 16
//     samples (half h of a 32-sample step) as 32 E2M3 slots: [Dr | Di] = X.
//     A voltage v is the E2M3 code v/8 exactly (the beamformer's converter);
//     the MFMA's E8M0 scales 2^3 on both operands make each product the
//     integer d_j d_i, so FP32 accumulation is exact (|V| < 2^24).
//     Re V(j, i) = X_j . X_i,  Im V(j, i) = X_j . W_i with W = [Di | -Dr]
//     (a swap and sign flip of X), for the upper-triangle tiles J <= I.
#include "bf_common.h"

namespace {
constexpr int XNW = 4, XNT = XNW * 64;    // one wave per (channel, pol) of a channel pair
constexpr int XU = 64;                    // samples per staged unit
constexpr int XUP = XU / 2;               // packets per unit
constexpr int XROW = 80;                  // LDS bytes per (channel-pol, antenna) row
constexpr int XPD = 3;                    // units of voltages in flight per thread

constexpr int TA = 32;                    // antennas per turn workgroup
constexpr int TK = 64;                    // channel pairs per turn workgroup

struct XcArgs {
  const uint8_t *in;                      // pair planes [group*4 + pair][packet][antenna][8]
  float *out;                             // visibilities, or slice partials
  int nant, nchan, ngroup, npkt, nslice;
  uint32_t in_bytes;
  long slice_floats;
};
typedef uint32_t u2v __attribute__((ext_vector_type(2)));
}  // namespace

// The block's 8-byte pieces (packet p, antenna a, channel pair k = 4g + cp)
// to the pair planes.  Workgroup = (packet, 32 antennas, 64 pairs).
__global__ __launch_bounds__(256) void xc_turn_kernel(const uint8_t *__restrict__ in, uint8_t *__restrict__ turn,
                                                      int nant, int nchan, int npkt, int order) {
  __shared__ u2v lds[TA][TK + 1];
  const int npair = nchan / 2, ngroup = nchan / 8;
  const int nkb = (npair + TK - 1) / TK, nab = (nant + TA - 1) / TA;
  const long w = blockIdx.x;
  const int kb = (int)(w % nkb), ab = (int)((w / nkb) % nab);
  const long p = w / ((long)nkb * nab);
  const int tid = threadIdx.x;
  // reads: a wave covers 64 consecutive pairs of one antenna
#pragma unroll
  for (int i = 0; i < TA / 4; i++) {
    const int kk = tid & 63, aa = (tid >> 6) + 4 * i;
    const int k = kb * TK + kk, a = ab * TA + aa;
    if (k < npair && a < nant)
      lds[aa][kk] = *(const u2v *)(in + chunk_offset(order, p, a, k >> 2, nant, nchan, ngroup) + (k & 3) * 8);
  }
  __syncthreads();
  // writes: half a wave covers 32 consecutive antennas of one pair
#pragma unroll
  for (int i = 0; i < TK / 8; i++) {
    const int aa = tid & 31, kk = (tid >> 5) + 8 * i;
    const int k = kb * TK + kk, a = ab * TA + aa;
    if (k < npair && a < nant) *(u2v *)(turn + (((long)k * npkt + p) * nant + a) * 8) = lds[aa][kk];
  }
}

// NB antenna blocks of 32; PD units of voltages in flight in registers.
template <int NB, int PD>
__global__ __launch_bounds__(XNT) void xc_kernel(XcArgs a) {
  constexpr int NA = NB * 32;             // antennas, padded
  constexpr int NT = NB * (NB + 1) / 2;   // upper-triangle tiles
  constexpr int BUF = 4 * NA * XROW;
  __shared__ __attribute__((aligned(16))) uint8_t stage[2 * BUF];
  const int P = blockIdx.x;
  const int L = (P & 7) * (gridDim.x >> 3) + (P >> 3);   // consecutive L on one XCD
  const int npair = a.ngroup * 4;
  if (L >= npair * a.nslice) return;
  const int kp = L % npair, slice = L / npair;
  const int cp = kp & 3, g = kp >> 2;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int n = lane & 31, h = lane >> 5;
  const __amdgpu_buffer_rsrc_t rin = buffer_rsrc(a.in, a.in_bytes);
  const int ntot = a.npkt / XUP;
  const int ub = (int)((long)ntot * slice / a.nslice), nunit = (int)((long)ntot * (slice + 1) / a.nslice) - ub;
  float *out = a.out + slice * a.slice_floats;

  // staging: task = (antenna, packet quad); NA * 8 tasks = NB per thread.
  // Unit u's loads go to register slot u % PD, PD units ahead of its use.
  u2v r[PD][NB][4];
  auto load = [&](int u, u2v (&rs)[NB][4]) {
#pragma unroll
    for (int k = 0; k < NB; k++) {
      const int task = tid + k * XNT, ant = task % NA, pq = task / NA;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const long p = (long)(ub + u) * XUP + pq * 4 + j;
        const uint32_t off = ant < a.nant ? (uint32_t)(((kp * (long)a.npkt + p) * a.nant + ant) * 8) : 0x80000000u;
        rs[k][j] = __builtin_amdgcn_raw_buffer_load_b64(rin, off, 0, 0);
      }
    }
  };
  // packet j of a task: dword c = channel 2cp + c, bytes [t0p0, t0p1, t1p0,
  // t1p1]; row (c, pol) gets the 8 samples (2 per packet) of the quad
  auto store = [&](const u2v (&rs)[NB][4], uint8_t *buf) {
#pragma unroll
    for (int k = 0; k < NB; k++) {
      const int task = tid + k * XNT, ant = task % NA, pq = task / NA;
#pragma unroll
      for (int cq = 0; cq < 4; cq++) {
        const int c = cq >> 1;
        const uint32_t sel = (cq & 1) ? 0x07050301u : 0x06040200u;
        const uint32_t lo = __builtin_amdgcn_perm(rs[k][1][c], rs[k][0][c], sel);
        const uint32_t hi = __builtin_amdgcn_perm(rs[k][3][c], rs[k][2][c], sel);
        *(u2v *)(buf + (cq * NA + ant) * XROW + pq * 8) = (u2v){lo, hi};
      }
    }
  };

  f16v acc_re[NT], acc_im[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) { acc_re[t] = (f16v){}; acc_im[t] = (f16v){}; }

  // one unit from LDS: two 32-sample K steps
  auto compute = [&](const uint8_t *buf) {
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
      i8v X[NB], W[NB];
#pragma unroll
      for (int J = 0; J < NB; J++) {
        const uint4 v = *(const uint4 *)(buf + (wave * NA + 32 * J + n) * XROW + ks * 32 + h * 16);
        const uint32_t x[4] = {v.x, v.y, v.z, v.w};
        uint32_t o[6];
        f6_operand(x, o);                 // [Dr | -Di]
        const uint32_t d3 = o[3] ^ E2M3_SIGN0, d4 = o[4] ^ E2M3_SIGN1, d5 = o[5] ^ E2M3_SIGN2;
        X[J] = (i8v){(int)o[0], (int)o[1], (int)o[2], (int)d3, (int)d4, (int)d5, 0, 0};
        W[J] = (i8v){(int)d3, (int)d4, (int)d5, (int)(o[0] ^ E2M3_SIGN0), (int)(o[1] ^ E2M3_SIGN1),
                     (int)(o[2] ^ E2M3_SIGN2), 0, 0};
      }
      int t = 0;
#pragma unroll
      for (int J = 0; J < NB; J++)
#pragma unroll
        for (int I = J; I < NB; I++, t++) {
          acc_re[t] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(X[J], X[I], acc_re[t], 2, 2, 0, 130, 0, 130);
          acc_im[t] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(X[J], W[I], acc_im[t], 2, 2, 0, 130, 0, 130);
        }
    }
  };

#pragma unroll
  for (int k = 0; k < PD; k++)
    if (k < nunit) load(k, r[k]);
  store(r[0], stage);
  __syncthreads();
  for (int u0 = 0; u0 < nunit; u0 += PD) {
#pragma unroll
    for (int k = 0; k < PD; k++) {
      const int u = u0 + k;              // in slot k, already in LDS buffer u & 1
      if (u < nunit) {
        if (u + PD < nunit) load(u + PD, r[k]);
        compute(stage + (u & 1) * BUF);
        if (u + 1 < nunit) store(r[(k + 1) % PD], stage + ((u + 1) & 1) * BUF);
        __syncthreads();
      }
    }
  }

  // the upper triangle: lane (n, h) of tile (J, I) holds column i = 32I + n
  // and rows j = 32J + 8(v/4) + 4h + v%4
  const int ch = 8 * g + 2 * cp + (wave >> 1), q = wave & 1;
  const int i = n;
  int t = 0;
#pragma unroll
  for (int J = 0; J < NB; J++)
#pragma unroll
    for (int I = J; I < NB; I++, t++) {
      const int ii = 32 * I + i;
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int jj = 32 * J + 8 * (v >> 2) + 4 * h + (v & 3);
        if (ii < a.nant && jj <= ii) {
          const long b = (long)ii * (ii + 1) / 2 + jj;
          *(float2 *)(out + ((b * a.nchan + ch) * 2 + q) * 2) = make_float2(acc_re[t][v], acc_im[t][v]);
        }
      }
    }
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

// Slices of the block's time for a shape: one when the channel pairs alone
// give most of a workgroup per CU (measured on the 384-channel block: 1, 2, 3,
// 4 and 8 slices took 0.30, 0.32, 0.34, 0.32 and 0.40 ms: the partial sums'
// traffic outweighs the fuller chip), else enough for one per CU, at least 8
// units (512 samples) per slice.
extern "C" int bfk_xcorr_default_slices(int nchan, int npkt) {
  const int npair = nchan / 2, nunit = npkt / XUP;
  int s = npair >= 128 ? 1 : 256 / npair;
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
  const int npair = nchan / 2;
  const long nturn = (long)npkt * ((nant + TA - 1) / TA) * ((npair + TK - 1) / TK);
  hipLaunchKernelGGL(xc_turn_kernel, dim3((unsigned)nturn), dim3(256), 0, s, d_in, d_turn, nant, nchan, npkt, order);
  XcArgs a;
  a.in = d_turn;
  a.out = nslice > 1 ? d_part : d_vis;
  a.nant = nant;
  a.nchan = nchan;
  a.ngroup = nchan / 8;
  a.npkt = npkt;
  a.nslice = nslice;
  a.in_bytes = (uint32_t)((size_t)npkt * nant * nchan * 4);
  a.slice_floats = (long)(bfk_xcorr_vis_bytes(nant, nchan) / sizeof(float));
  const dim3 grid((npair * nslice + 7) / 8 * 8);
  if (nant <= 32) hipLaunchKernelGGL((xc_kernel<1, XPD>), grid, dim3(XNT), 0, s, a);
  else if (nant <= 64) hipLaunchKernelGGL((xc_kernel<2, XPD>), grid, dim3(XNT), 0, s, a);
  else hipLaunchKernelGGL((xc_kernel<3, XPD>), grid, dim3(XNT), 0, s, a);
  if (nslice > 1) {
    const long n4 = a.slice_floats / 4;
    hipLaunchKernelGGL(xc_reduce_kernel, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, s,
                       (const float4 *)d_part, (float4 *)d_vis, n4, nslice);
  }
  return bfk_check(hipGetLastError(), "xc_kernel");
}

// fl_kernels.hip -- device work of the search-side RFI flagger on a
// [nbeam][ntime][nchan] uint8 filterbank block (gpu_flagger.cu:113-520).
//
// The reference reads the 1 GiB block (64 beams x 16384 times x 1024
// channels) once per statistic: calc_spectrum, calc_varspec, calc_ppspec and
// (as its loop calls it) calc_spectrum again, each with one 32-thread block
// per (beam, channel) walking a column at a stride of 1 KB, after scaley has
// read and written it.  Here:
//
//   fl_stats_kernel   one pass for every statistic, optionally scaling the
//                     block in place first (scaley fused, as a per-channel
//                     table of the 256 byte values built in LDS with the
//                     reference's FP64 formula): a workgroup owns a
//                     beam's 256-channel slab, each lane 4 channels read as
//                     one dword per row (a wave reads 256 contiguous bytes),
//                     the 16 waves the 32 time segments of the reference's
//                     threads.  Sums are integers (exact in any order); the
//                     variance comes from exact integer sums (the reference's
//                     float sums differ from it by their rounding, < 1e-5
//                     relative); ppspec keeps the reference's 32-segment
//                     semantics (max over segments of the largest 128-sample
//                     bin mean, minus the minimum over segments 16..31: what
//                     its unsynchronised max-then-min tree leaves in slot 0).
//   fl_flag_ts_kernel one pass that replaces the masked (beam, channel)
//                     columns and sums each row of the flagged block (flag +
//                     make_ts): a wave per row, 16 bytes per lane.
//   fl_flagts_kernel  the masked (beam, time) rows, channels 128..nchan-129.
//   and elementwise kernels for sumpulse, genrand (own counter-based
//   generator), the noise replacement copy and fix_zspec's median.
#include <stdint.h>

#include "bf_common.h"

namespace {
constexpr int FL_SLAB = 256;   // channels per stats workgroup
constexpr int FL_SW = 16;      // waves per stats workgroup
constexpr int FL_NSEG = 32;    // the reference's threads per (beam, channel)
constexpr int FL_TBIN = 128;   // ppspec bin
constexpr double FL_MN = 48.0, FL_SIG = 6.0;
constexpr int FL_RPW = 16;     // rows per wave in fl_flag_ts_kernel

// (unsigned char)(__float2uint_rn(2 * val) / 2): round half to even, the
// saturating conversion (NaN and negatives to 0), halve, keep the low byte.
__device__ inline uint32_t fl_u8(float val) {
  const float r = __builtin_rintf(2.0f * val);
  const uint32_t u = r > 0.f ? (r >= 4294967296.f ? 0xFFFFFFFFu : (uint32_t)r) : 0u;
  return (u >> 1) & 0xFFu;
}

// The replacement byte of flag / flagts: MN * bpwr stored to a byte
// (truncation toward zero, clamped to 0..255, NaN to 0).
__device__ inline uint32_t fl_rep(float bpwr) {
  const double v = FL_MN * (double)bpwr;
  return v > 0. ? (v >= 255. ? 255u : (uint32_t)v) : 0u;
}

// scaley on one byte: (float)fma((double)(x - spec0), SIG / sqrtf(var0), MN).
__device__ inline uint32_t fl_scale(uint32_t x, float s0, double k) {
  const float a = (float)x - s0;
  return fl_u8((float)__builtin_fma((double)a, k, FL_MN));
}

struct FlStats {
  const uint8_t *data;
  uint8_t *wdata;          // scaled block (scaley fused), else NULL
  const float *spec0, *var0;
  const float *spec_in;    // varspec's spectrum argument (NULL: the block's own mean)
  float *spec, *var, *mx, *pp;
  int nb, nt, nc;
};
}  // namespace

template <bool SCALE>
__global__ __launch_bounds__(1024) void fl_stats_kernel(FlStats a) {
  __shared__ uint32_t lc[FL_NSEG][FL_SLAB];            // per segment: largest bin sum
  __shared__ uint32_t ls[FL_SW][FL_SLAB], lq[FL_SW][FL_SLAB], lm[FL_SW][FL_SLAB];
  // scaley as a table: the scaled byte of every input byte x for each channel
  // of the slab, [x][channel] so that a wave's 64 lanes (consecutive channel
  // quads) read 64 distinct banks whatever their bytes
  __shared__ uint8_t lut[SCALE ? 256 * FL_SLAB : 4];
  const int nslab = (a.nc + FL_SLAB - 1) / FL_SLAB;
  const int bm = blockIdx.x / nslab, slab = blockIdx.x % nslab;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c0 = slab * FL_SLAB + lane * 4;
  const bool live = c0 < a.nc;
  const int per = a.nt / FL_NSEG;
  const size_t nc = (size_t)a.nc;
  if (SCALE) {
    const int c = threadIdx.x & (FL_SLAB - 1), x0 = (threadIdx.x / FL_SLAB) * (256 / 4);
    const int ch = slab * FL_SLAB + c;
    if (ch < a.nc) {
      const float s0 = a.spec0[(size_t)bm * nc + ch];
      const double k = __ddiv_rn(FL_SIG, (double)__fsqrt_rn(a.var0[(size_t)bm * nc + ch]));
      for (int x = x0; x < x0 + 256 / 4; x++) lut[x * FL_SLAB + c] = (uint8_t)fl_scale((uint32_t)x, s0, k);
    }
    __syncthreads();
  }
  uint32_t S[4] = {0, 0, 0, 0}, Q[4] = {0, 0, 0, 0}, M[4] = {0, 0, 0, 0};
  // each wave walks its two segments' rows with PF loads in flight: row
  // i + PF is loaded as row i is used (the last PF loads re-read the
  // segment's last row).  Buffer loads over the beam's plane: 32-bit offsets,
  // and non-live lanes read zeros past its end.
  constexpr int PF = 16;
  static_assert(FL_TBIN % PF == 0, "bins end on a prefetch block");
  const size_t plane = (size_t)a.nt * nc;
  // (the beam's base address through readfirstlane: the scaley path also uses
  // bm in per-lane addresses, and a resource built from vector registers puts
  // a waterfall loop around every load)
  const uint64_t base = (uint64_t)(a.data + (size_t)bm * plane);
  const uint64_t sbase = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)base) |
                         (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(base >> 32)) << 32;
  const __amdgpu_buffer_rsrc_t rin = buffer_rsrc((const void *)sbase, (uint32_t)plane);
  // the scaled block through a buffer store over the same plane of wdata:
  // non-live lanes' stores (offset past the plane) are dropped, so the rows
  // need no branch
  __amdgpu_buffer_rsrc_t rout = rin;
  if (SCALE) {
    const uint64_t wbase = (uint64_t)(a.wdata + (size_t)bm * plane);
    const uint64_t swbase = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)wbase) |
                            (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(wbase >> 32)) << 32;
    rout = buffer_rsrc((const void *)swbase, (uint32_t)plane);
  }
  const uint32_t cl = live ? (uint32_t)c0 : 0x80000000u;
  for (int h = 0; h < 2; h++) {
    const int seg = w + FL_SW * h;
    const uint32_t s0off = (uint32_t)(seg * per) * (uint32_t)a.nc;
    uint32_t cm[4] = {0, 0, 0, 0}, bs[4] = {0, 0, 0, 0};
    uint32_t buf[PF];
#pragma unroll
    for (int d = 0; d < PF; d++)
      buf[d] = __builtin_amdgcn_raw_buffer_load_b32(rin, cl + s0off + (uint32_t)d * (uint32_t)a.nc, 0, 0);
    for (int i0 = 0; i0 < per; i0 += PF) {
#pragma unroll
      for (int d = 0; d < PF; d++) {
        uint32_t x = buf[d];
        if (SCALE) {
          // (every lane: lane * 4 + 3 < FL_SLAB keeps a non-live lane's
          // lookups in the table, and its statistics are never written)
          const uint8_t *t = lut + lane * 4;
          x = (uint32_t)t[(x & 0xFF) * FL_SLAB] | (uint32_t)t[((x >> 8) & 0xFF) * FL_SLAB + 1] << 8 |
              (uint32_t)t[((x >> 16) & 0xFF) * FL_SLAB + 2] << 16 | (uint32_t)t[(x >> 24) * FL_SLAB + 3] << 24;
          __builtin_amdgcn_raw_buffer_store_b32(x, rout, cl + s0off + (uint32_t)(i0 + d) * (uint32_t)a.nc, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const uint32_t v = (x >> (8 * j)) & 0xFF;
          bs[j] += v;
          Q[j] += v * v;
          M[j] = max(M[j], v);
        }
        // the next load into buf[d] once x is dead, so that it can take x's
        // register: a load issued while x is live needs another one, and the
        // loop then rotates registers whose loads are still in flight
        const uint32_t nx = (uint32_t)min(i0 + PF + d, per - 1);
        buf[d] = __builtin_amdgcn_raw_buffer_load_b32(rin, cl + s0off + nx * (uint32_t)a.nc, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
      }
      if (((i0 + PF) & (FL_TBIN - 1)) == 0) {          // end of a 128-sample bin
#pragma unroll
        for (int j = 0; j < 4; j++) {
          S[j] += bs[j];
          cm[j] = max(cm[j], bs[j]);
          bs[j] = 0;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) lc[seg][lane * 4 + j] = cm[j];
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    ls[w][lane * 4 + j] = S[j];
    lq[w][lane * 4 + j] = Q[j];
    lm[w][lane * 4 + j] = M[j];
  }
  __syncthreads();
  const int c = threadIdx.x, ch = slab * FL_SLAB + c;
  if (c >= FL_SLAB || ch >= a.nc) return;
  uint64_t St = 0, Qt = 0;
  uint32_t Mt = 0, cmax = 0, cmin = 0xFFFFFFFFu;
  for (int v = 0; v < FL_SW; v++) {
    St += ls[v][c];
    Qt += lq[v][c];
    Mt = max(Mt, lm[v][c]);
  }
  for (int s = 0; s < FL_NSEG; s++) {
    cmax = max(cmax, lc[s][c]);
    if (s >= FL_NSEG / 2) cmin = min(cmin, lc[s][c]);
  }
  const size_t i = (size_t)bm * nc + ch;
  const double n = (double)a.nt;
  const float sp = (float)__ddiv_rn((double)St, n);
  if (a.spec) a.spec[i] = sp;
  if (a.var) {
    // sum (x - s)^2 = [(n Q - S^2) + (S - s n)^2] / n, exact up to its last
    // rounding when s is the mean
    const double s = (double)(a.spec_in ? a.spec_in[i] : sp);
    const double dl = __dsub_rn((double)St, __dmul_rn(s, n));
    const int64_t q = (int64_t)a.nt * (int64_t)Qt - (int64_t)St * (int64_t)St;
    const double num = __dadd_rn((double)q, __dmul_rn(dl, dl));
    a.var[i] = (float)__ddiv_rn(__ddiv_rn(num, n), n);
  }
  if (a.mx) a.mx[i] = (float)Mt;
  if (a.pp) a.pp[i] = (float)cmax / (float)FL_TBIN - (float)cmin / (float)FL_TBIN;
}

__device__ inline uint32_t fl_wave_sum(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// bytes != 0 -> 0xFF, == 0 -> 0x00
__device__ inline uint32_t fl_bytemask(uint32_t x) {
  x |= (x >> 4) & 0x0F0F0F0Fu;
  x |= (x >> 2) & 0x33333333u;
  x |= (x >> 1) & 0x55555555u;
  return (x & 0x01010101u) * 0xFFu;
}

// Replace the masked columns (mask[bm][ch] != 0) with the beam's byte and,
// when ts is given, write ts[bm][t] = (sum over channels of the flagged row)
// / nchan.  A wave per FL_RPW consecutive rows of one beam, 16 bytes per lane.
// A lane stores its chunk when the chunk holds a masked column, or, when at
// least a quarter of the wave's chunks do, every lane stores (unchanged chunks
// are written back as read): on the 64 x 16384 x 1024 block, 5 % of columns
// masked at random touch every line of a row, and whole rows then take
// 0.436 ms against 0.459 ms as scattered chunks, while 0.5 % at random or a
// 64-channel band run faster as chunks (0.357 and 0.343 ms, against 0.431
// whole).
__global__ __launch_bounds__(256) void fl_flag_ts_kernel(uint8_t *data, const uint8_t *mask,
                                                         const float *bpwr, float *ts, int nb, int nt,
                                                         int nc) {
  const int lane = threadIdx.x & 63;
  const long gw = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int wpb = nt / FL_RPW;                 // waves per beam
  if (gw >= (long)nb * wpb) return;            // the whole wave leaves together
  const int bm = (int)(gw / wpb), t0 = (int)(gw % wpb) * FL_RPW;
  const int nq = nc / 16;                      // 16-byte chunks per row
  const uint32_t rep = mask ? fl_rep(bpwr[bm]) * 0x01010101u : 0u;
  uint32_t rs[FL_RPW];
#pragma unroll
  for (int r = 0; r < FL_RPW; r++) rs[r] = 0;
  for (int q = lane; q - lane < nq; q += 64) {
    if (q >= nq) break;
    uint4 m = make_uint4(0, 0, 0, 0);
    if (mask) {
      m = *(const uint4 *)(mask + (size_t)bm * nc + 16 * q);
      m = make_uint4(fl_bytemask(m.x), fl_bytemask(m.y), fl_bytemask(m.z), fl_bytemask(m.w));
    }
    const bool own = (m.x | m.y | m.z | m.w) != 0;
    // (both ballots in every active lane: under `own ||` they would run only
    // in the lanes whose chunk is clean)
    const bool whole = 4 * __popcll(__ballot(own)) >= __popcll(__ballot(1));
    const bool any = own || whole;
#pragma unroll
    for (int r = 0; r < FL_RPW; r++) {
      uint4 *pv = (uint4 *)(data + ((size_t)bm * nt + t0 + r) * nc + 16 * q);
      uint4 v = *pv;
      if (any) {
        v.x = (v.x & ~m.x) | (rep & m.x);
        v.y = (v.y & ~m.y) | (rep & m.y);
        v.z = (v.z & ~m.z) | (rep & m.z);
        v.w = (v.w & ~m.w) | (rep & m.w);
        *pv = v;
      }
      uint32_t s = __builtin_amdgcn_sad_u8(v.x, 0u, rs[r]);
      s = __builtin_amdgcn_sad_u8(v.y, 0u, s);
      s = __builtin_amdgcn_sad_u8(v.z, 0u, s);
      rs[r] = __builtin_amdgcn_sad_u8(v.w, 0u, s);
    }
  }
  if (ts) {
#pragma unroll
    for (int r = 0; r < FL_RPW; r++) {
      const uint32_t s = fl_wave_sum(rs[r]);
      if (lane == 0) ts[(size_t)bm * nt + t0 + r] = (float)__ddiv_rn((double)s, (double)nc);
    }
  }
}

// flagts (gpu_flagger.cu:478-494): row midx[i] = bm * ntime + t, channels
// 128 .. nchan-129 set to the beam's byte.  A wave per listed row.
__global__ __launch_bounds__(256) void fl_flagts_kernel(uint8_t *data, const int *midx, int n,
                                                        const float *bpwr, int nb, int nt, int nc) {
  const int lane = threadIdx.x & 63;
  const long i = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const int m = midx[i];
  if (m < 0 || m >= nb * nt) return;
  const int bm = m / nt;
  const uint32_t rep = fl_rep(bpwr[bm]) * 0x01010101u;
  uint8_t *row = data + (size_t)m * nc;
  for (int q = 8 + lane; q < nc / 16 - 8; q += 64) *(uint4 *)(row + 16 * q) = make_uint4(rep, rep, rep, rep);
}

// flag (gpu_flagger.cu:458-474): column midx[i] = bm * nchan + ch, every time.
__global__ __launch_bounds__(256) void fl_flag_cols_kernel(uint8_t *data, const int *midx, int n,
                                                           const float *bpwr, int nb, int nt, int nc) {
  const int i = blockIdx.y;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (i >= n || t >= nt) return;
  const int m = midx[i];
  if (m < 0 || m >= nb * nc) return;
  const int bm = m / nc, ch = m % nc;
  data[((size_t)bm * nt + t) * nc + ch] = (uint8_t)fl_rep(bpwr[bm]);
}

// sumpulse (gpu_flagger.cu:398-406): 4 bytes per thread, a tail thread for
// the last n % 4.
__global__ __launch_bounds__(256) void fl_sumpulse_kernel(uint8_t *data, const float *summand, size_t n) {
  const size_t i = ((size_t)blockIdx.x * 256 + threadIdx.x) * 4;
  if (i + 4 <= n) {
    const uint32_t x = *(const uint32_t *)(data + i);
    const float4 s = *(const float4 *)(summand + i);
    *(uint32_t *)(data + i) = fl_u8((float)(x & 0xFF) + s.x) | fl_u8((float)((x >> 8) & 0xFF) + s.y) << 8 |
                              fl_u8((float)((x >> 16) & 0xFF) + s.z) << 16 | fl_u8((float)(x >> 24) + s.w) << 24;
  } else {
    for (size_t j = i; j < n; j++) data[j] = (uint8_t)fl_u8((float)data[j] + summand[j]);
  }
}

__device__ inline uint64_t fl_mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// genrand (gpu_flagger.cu:499-520) with a counter-based generator: element i
// from the splitmix64 finaliser of (seed * 2^32 + i), uniforms
// (32-bit half % 16384) / 16384, Box-Muller in FP64.
__global__ __launch_bounds__(256) void fl_genrand_kernel(uint8_t *rep, size_t n, uint32_t seed) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint64_t r = fl_mix(((uint64_t)seed << 32) + i);
  const double u1 = (double)((uint32_t)r % 16384u) / 16384.0;
  const double u2 = (double)((uint32_t)(r >> 32) % 16384u) / 16384.0;
  const double va = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
  rep[i] = (uint8_t)fl_u8((float)__builtin_fma(va, FL_SIG, FL_MN));
}

// Every beam of the block replaced by the [ntime][nchan] noise block
// (the reference's per-beam device copies, gpu_flagger.cu:1319-1320,1356-1358).
__global__ __launch_bounds__(256) void fl_noise_kernel(uint8_t *data, const uint8_t *rep, size_t per_beam,
                                                       int nb) {
  const size_t i = ((size_t)blockIdx.x * 256 + threadIdx.x) * 16;
  if (i >= per_beam) return;
  const uint4 v = *(const uint4 *)(rep + i);
  for (int b = 0; b < nb; b++) *(uint4 *)(data + (size_t)b * per_beam + i) = v;
}

// fix_zspec (gpu_flagger.cu:113-168): entry i of s0 / v0 becomes the value
// among the naver stacked blocks with exactly naver/2 others <= it; with no
// such value (ties) entry i keeps block 0's.
__global__ __launch_bounds__(256) void fl_fix_zspec_kernel(float *s0, float *v0, int nbc, int naver) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nbc) return;
  const int place = naver / 2;
  for (int h = 0; h < 2; h++) {
    float *a = h ? v0 : s0;
    float pick = a[i];
    for (int t = 0; t < naver; t++) {
      const float x = a[(size_t)t * nbc + i];
      int ct = 0;
      for (int k = 0; k < naver; k++) ct += (k != t && a[(size_t)k * nbc + i] <= x);
      if (ct == place) pick = x;
    }
    a[i] = pick;
  }
}

// ---- launchers --------------------------------------------------------------

static bool fl_shape_ok(int nb, int nt, int nc, const char *who) {
  if (nb < 1 || nt < 4096 || nt % 4096 || nc < 16 || nc % 16 || (long)nb * nt > (1L << 30) / 16 ||
      (long)nt * nc >= (1L << 31)) {
    bfk_set_error("%s: shape %d x %d x %d unsupported (ntime a multiple of 4096, nchan of 16)", who, nb,
                  nt, nc);
    return false;
  }
  return true;
}

extern "C" int flk_stats(const uint8_t *d_data, uint8_t *d_scaled, const float *d_spec0, const float *d_var0,
                         const float *d_spec_in, float *d_spec, float *d_var, float *d_max, float *d_pp,
                         int nb, int nt, int nc, void *stream) {
  if (!fl_shape_ok(nb, nt, nc, "flk_stats")) return -1;
  if (d_scaled && (!d_spec0 || !d_var0)) {
    bfk_set_error("flk_stats: scaling needs spec0 and var0");
    return -1;
  }
  FlStats a{d_data, d_scaled, d_spec0, d_var0, d_spec_in, d_spec, d_var, d_max, d_pp, nb, nt, nc};
  const dim3 grid(nb * ((nc + FL_SLAB - 1) / FL_SLAB)), block(64 * FL_SW);
  if (d_scaled)
    hipLaunchKernelGGL(fl_stats_kernel<true>, grid, block, 0, (hipStream_t)stream, a);
  else
    hipLaunchKernelGGL(fl_stats_kernel<false>, grid, block, 0, (hipStream_t)stream, a);
  return bfk_check(hipGetLastError(), "fl_stats_kernel");
}

extern "C" int flk_flag_ts(uint8_t *d_data, const uint8_t *d_mask, const float *d_bpwr, float *d_ts, int nb,
                           int nt, int nc, void *stream) {
  if (!fl_shape_ok(nb, nt, nc, "flk_flag_ts")) return -1;
  const long waves = (long)nb * (nt / FL_RPW);
  hipLaunchKernelGGL(fl_flag_ts_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, (hipStream_t)stream,
                     d_data, d_mask, d_bpwr, d_ts, nb, nt, nc);
  return bfk_check(hipGetLastError(), "fl_flag_ts_kernel");
}

extern "C" int flk_flagts(uint8_t *d_data, const int *d_midx, int n, const float *d_bpwr, int nb, int nt, int nc,
                          void *stream) {
  if (!fl_shape_ok(nb, nt, nc, "flk_flagts")) return -1;
  if (n <= 0) return 0;
  hipLaunchKernelGGL(fl_flagts_kernel, dim3((n + 3) / 4), dim3(256), 0, (hipStream_t)stream, d_data, d_midx, n,
                     d_bpwr, nb, nt, nc);
  return bfk_check(hipGetLastError(), "fl_flagts_kernel");
}

extern "C" int flk_flag_cols(uint8_t *d_data, const int *d_midx, int n, const float *d_bpwr, int nb, int nt,
                             int nc, void *stream) {
  if (!fl_shape_ok(nb, nt, nc, "flk_flag_cols")) return -1;
  if (n <= 0) return 0;
  if (n > 65535 * 16) {
    bfk_set_error("flk_flag_cols: %d columns", n);
    return -1;
  }
  for (int i0 = 0; i0 < n; i0 += 65535) {
    const int k = n - i0 < 65535 ? n - i0 : 65535;
    hipLaunchKernelGGL(fl_flag_cols_kernel, dim3((nt + 255) / 256, k), dim3(256), 0, (hipStream_t)stream,
                       d_data, d_midx + i0, k, d_bpwr, nb, nt, nc);
  }
  return bfk_check(hipGetLastError(), "fl_flag_cols_kernel");
}

extern "C" int flk_sumpulse(uint8_t *d_data, const float *d_summand, size_t n, void *stream) {
  if (n == 0) return 0;
  const size_t threads = (n + 3) / 4;
  hipLaunchKernelGGL(fl_sumpulse_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, d_data, d_summand, n);
  return bfk_check(hipGetLastError(), "fl_sumpulse_kernel");
}

extern "C" int flk_genrand(uint8_t *d_rep, size_t n, uint32_t seed, void *stream) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(fl_genrand_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     d_rep, n, seed);
  return bfk_check(hipGetLastError(), "fl_genrand_kernel");
}

extern "C" int flk_noise(uint8_t *d_data, const uint8_t *d_rep, int nb, int nt, int nc, void *stream) {
  if (!fl_shape_ok(nb, nt, nc, "flk_noise")) return -1;
  const size_t per = (size_t)nt * nc;
  hipLaunchKernelGGL(fl_noise_kernel, dim3((unsigned)((per / 16 + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, d_data, d_rep, per, nb);
  return bfk_check(hipGetLastError(), "fl_noise_kernel");
}

extern "C" int flk_fix_zspec(float *d_s0, float *d_v0, int nbc, int naver, void *stream) {
  if (nbc < 1 || naver < 1) {
    bfk_set_error("flk_fix_zspec: %d entries x %d blocks", nbc, naver);
    return -1;
  }
  hipLaunchKernelGGL(fl_fix_zspec_kernel, dim3((nbc + 255) / 256), dim3(256), 0, (hipStream_t)stream, d_s0, d_v0,
                     nbc, naver);
  return bfk_check(hipGetLastError(), "fl_fix_zspec_kernel");
}

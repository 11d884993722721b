// bf_aux.hip -- the small kernels around the fused beamformer: antenna
// stuffing (-s) and the incoherent beam (-q), and the per-beam power sums
// the AGC consumes.
#include "bf_common.h"

// One workgroup per (64-sample block, group), one thread per antenna sums
// |d|^2 per output bin over the block's packets, 8 channels, 2 times and 2
// pols; integer-valued sums are exact in float.
//  stuff 1: beam b <- antenna ((b/16)%4)*16 + b%16  (dsaX_beamformer.cu:326-337)
//  stuff 2: beam 0 <- sum over all antennas          (dsaX_beamformer.cu:307-323)
// The per-32-sample beam sums in bpart are rewritten for the beams written here.
__global__ __launch_bounds__(256) void bf_stuff_kernel(bfk_dims d, const uint8_t *__restrict__ in,
                                                       const float *__restrict__ scale,
                                                       uint8_t *__restrict__ out,
                                                       float *__restrict__ power,
                                                       float *__restrict__ bpart) {
  __shared__ float S[256][10];      // 8 bins, then the two 32-sample halves
  __shared__ float IB[10];
  const int nblk = d.nsamp / 64;
  const int t64 = blockIdx.x % nblk, g = blockIdx.x / nblk;
  const int nbin = 64 / d.tavg < 1 ? 1 : 64 / d.tavg;
  const int ppb = d.tavg / 2 < 32 ? d.tavg / 2 : 32;      // packets per bin
  const int tid = threadIdx.x;
  float s[10] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (tid < d.nant) {
    for (int pl = 0; pl < 32; pl++) {
      const long p = (long)t64 * 32 + pl;
      const uint4 *src = (const uint4 *)(in + chunk_offset(d.order, p, tid, g, d.nant, d.nchan, d.ngroup));
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
      if (pl < 16) s[8] += acc;
      else s[9] += acc;
    }
  }
#pragma unroll
  for (int k = 0; k < 10; k++) S[tid][k] = s[k];
  __syncthreads();
  if (tid < 10) {
    float x = 0.f;
    for (int a = 0; a < d.nant; a++) x += S[a][tid];
    IB[tid] = x;
  }
  __syncthreads();
  const int bend = d.stuff == 2 ? 1 : d.nbeam;
  for (int b = tid; b < bend; b += 256) {
    const int ant = ((b / 16) % 4) * 16 + b % 16;
    for (int k = 0; k < nbin; k++) {
      const float x = d.stuff == 2 ? IB[k] : (ant < d.nant ? S[ant][k] : 0.f);
      const long idx = out_index(d.layout, d.nout, d.nbeam, d.ngroup, t64 * nbin + k, b, g);
      if (power) power[idx] = x;
      out[idx] = requant_u8(x, scale[b]);
    }
    for (int hh = 0; hh < 2; hh++)
      bpart[((long)b * 2 * nblk + 2 * t64 + hh) * d.ngroup + g] =
          d.stuff == 2 ? IB[8 + hh] : (ant < d.nant ? S[ant][8 + hh] : 0.f);
  }
}

// beam_power[b] = sum over 32-sample units and groups of bpart[b][.][.]
// (badder + calc_allbp, dsaX_beamformer.cu:78-112,482-495).  One workgroup
// per beam, FP32 throughout: each thread sums its strided share in eight
// independent partial sums (all loads of a round issued before the adds),
// the waves reduce with cross-lane swaps and the four wave sums are added in
// LDS, all in a fixed order, so the result is deterministic.
__device__ __forceinline__ void beam_sum(const float *__restrict__ src, int n, float *__restrict__ dst) {
  __shared__ float red[4];
  const int tid = threadIdx.x;
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  int i = tid;
  for (; i + 7 * 256 < n; i += 8 * 256) {
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = src[i + k * 256];
#pragma unroll
    for (int k = 0; k < 8; k++) acc[k] += v[k];
  }
  for (; i < n; i += 256) acc[0] += src[i];
  float s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) s += __shfl_xor(s, m);
  if ((tid & 63) == 0) red[tid >> 6] = s;
  __syncthreads();
  if (tid == 0) *dst = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void bf_beam_power_kernel(bfk_dims d, const float *__restrict__ bpart,
                                                            float *__restrict__ bp) {
  const int n = (d.nsamp / 32) * d.ngroup;
  beam_sum(bpart + (long)blockIdx.x * n, n, bp + blockIdx.x);
}

// The beam-power sums of every block of a batch in one launch: workgroup B
// is beam B - beam_start[j] of job j, reduced exactly as bf_beam_power_kernel
// does, so the sums are bit-identical to a per-block launch.
__global__ __launch_bounds__(256) void bf_beam_power_batch_kernel(const bfk_bp_job *__restrict__ jobs,
                                                                  const int *__restrict__ beam_start) {
  const int B = blockIdx.x;
  int j = 0;
  while (B >= beam_start[j + 1]) j++;
  const bfk_bp_job &job = jobs[j];
  const int b = B - beam_start[j];
  beam_sum(job.bpart + (long)b * job.n, job.n, job.out + b);
}

extern "C" int bfk_beam_power_batch(const void *d_jobs, const int *d_beam_start, int nbeams_total,
                                    void *stream) {
  if (nbeams_total < 1) return 0;
  hipLaunchKernelGGL(bf_beam_power_batch_kernel, dim3(nbeams_total), dim3(256), 0, (hipStream_t)stream,
                     (const bfk_bp_job *)d_jobs, d_beam_start);
  return bfk_check(hipGetLastError(), "bf_beam_power_batch_kernel");
}

extern "C" int bfk_stuff(const bfk_dims *d, const uint8_t *d_in, const float *d_scale,
                         uint8_t *d_out, float *d_power, float *d_bpart, void *stream) {
  if (d->nant > 256 || d->nsamp % 64 != 0) {
    bfk_set_error("bfk_stuff: unsupported shape nant=%d nsamp=%d", d->nant, d->nsamp);
    return -1;
  }
  hipLaunchKernelGGL(bf_stuff_kernel, dim3((d->nsamp / 64) * d->ngroup), dim3(256), 0,
                     (hipStream_t)stream, *d, d_in, d_scale, d_out, d_power, d_bpart);
  return bfk_check(hipGetLastError(), "bf_stuff_kernel");
}

extern "C" size_t bfk_bpart_bytes(const bfk_dims *d) {
  return (size_t)d->nbeam * (d->nsamp / 32) * d->ngroup * sizeof(float);
}

extern "C" int bfk_beam_power(const bfk_dims *d, const float *d_bpart, float *d_beam_power,
                              void *stream) {
  hipLaunchKernelGGL(bf_beam_power_kernel, dim3(d->nbeam), dim3(256), 0, (hipStream_t)stream, *d,
                     d_bpart, d_beam_power);
  return bfk_check(hipGetLastError(), "bf_beam_power_kernel");
}

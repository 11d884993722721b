// bf_fused.hip -- the "pipelined" schedule of the fused beamformer and the
// schedule dispatch.
//
// Reference path: src/dsaX_beamformer.cu promoter (:170-193), beamformer
// (:267-416), badder (:78-112) and adder (:120-158).  The reference runs five
// launches per 32-sample "int" (640 per block) and keeps unpacked FP16 planes
// and FP32 partial powers in HBM between them.  Here one kernel does, per
// tile (one 8-channel output group x 64 samples x 256 beams):
//   (a) the packed 4-bit block is copied in coalesced 16-byte slices into a
//       raw [packet][antenna] LDS buffer and transposed LDS -> LDS into
//       [channel][pol][sample][antenna] rows;
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
// Schedule.  One persistent workgroup of 4 waves per CU (one wave per SIMD,
// 512 registers each) walks an XCD-local list of tiles and overlaps, by
// construction, everything that would stall the MFMA pipe:
//   - the next tile's raw copy is issued in slices between the GEMM units and
//     lands while they run; it is transposed at the tile boundary;
//   - a tile is 32 units (16 (channel, pol) x two 32-sample halves); unit u
//     accumulates into one of two accumulator sets while the detect of unit
//     u-1 drains the other;
//   - weights of (channel, pol) c+1 load into the second of two register sets
//     two units before they are needed.
// All loads are buffer loads whose offsets beyond the lane part are
// wave-uniform, so the loop keeps no per-load address registers.
#include "bf_common.h"

namespace {
constexpr int NW = 4, NT = NW * 64;       // waves, threads per workgroup
constexpr int BSET = NW * BTW * 32;       // beams per tile
constexpr int RAW_BYTES = PCU * 64 * 32;  // [packet][64 antenna slots][32 B]
constexpr int RAW_LOADS = RAW_BYTES / 16 / NT;   // 16-byte direct loads per thread
constexpr int TASKS = PCU * 16 / NT;      // (packet, antenna quad) transposes per thread
}

// Raw copy slice i (of RAW_LOADS) of a tile: wave w fills slots
// [(w*RAW_LOADS + i)*64, +64) of 16 bytes; slot -> packet slot>>7, antenna
// (slot>>1)&63, half slot&1.  Antennas past nant read antenna nant-1 (in
// bounds) and are zeroed by the transpose.  A slice is fetched into four
// registers and written to LDS one slice later, so its latency hides behind
// the GEMM units in between.  (A direct global->LDS copy would need a full
// vmcnt drain before every later LDS read of the stage: the compiler cannot
// keep the two apart across the loop.)
__device__ __forceinline__ uint4 raw_fetch(__amdgpu_buffer_rsrc_t rin, uint32_t tile_off,
                                           const BfArgs &a, int wave, uint32_t voff_even,
                                           uint32_t voff_odd, int i) {
  const int blk = wave * RAW_LOADS + i;          // wave-uniform
  const uint32_t soff = tile_off + (uint32_t)(blk >> 1) * a.pstride;
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  const u4 v = __builtin_amdgcn_raw_buffer_load_b128(rin, (blk & 1) ? voff_odd : voff_even, soff, 0);
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void raw_put(uint8_t *raw, int wave, int lane, int i, uint4 v) {
  *(uint4 *)(raw + ((wave * RAW_LOADS + i) * 64 + lane) * 16) = v;
}

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
    stage_store(stage, v, pl, aq, ROWB);
  }
}

template <int PREC, int TAVG>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(1, 1)))
void bf_pipelined_kernel(BfArgs a) {
  __shared__ __attribute__((aligned(16))) uint8_t stage[STAGE_BYTES];
  __shared__ __attribute__((aligned(16))) uint8_t raw[RAW_BYTES];
  constexpr int NP = TAVG <= 32 ? 32 / TAVG : 1;   // bins per 32-row unit
  static_assert(RAW_LOADS == 16, "raw slices are issued two per (channel, pol) pair");

  // Persistent, XCD-aware: workgroup P runs on XCD P % 8 and walks that XCD's
  // contiguous share of the tile list with stride = workgroups per XCD.
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3, nslot = gridDim.x >> 3;
  const int tbeg = xcd * a.tiles_per_xcd;
  const int tend = min(tbeg + a.tiles_per_xcd, a.ntiles);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int n = lane & 31, h = lane >> 5;

  int L = tbeg + slot;
  if (L >= tend) return;

  const __amdgpu_buffer_rsrc_t rin = buffer_rsrc(a.in, a.in_bytes);
  const __amdgpu_buffer_rsrc_t rw = buffer_rsrc(a.wfrag, a.wfrag_bytes);
  const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(a.wscale, a.wscale_bytes);
  const uint32_t vlane16 = (uint32_t)lane * 16u, vn4 = (uint32_t)n * 4u;
  // lane part of the raw copy offsets: even slices hold antennas 0..31, odd
  // ones 32..63 (two lanes per antenna, 16 bytes each)
  const uint32_t voff_even = (uint32_t)min(lane >> 1, a.nant - 1) * a.astride + (lane & 1) * 16u;
  const uint32_t voff_odd = (uint32_t)min(32 + (lane >> 1), a.nant - 1) * a.astride + (lane & 1) * 16u;
  auto tile_off = [&](const TileId &t) -> uint32_t {
    return (uint32_t)t.tblk * PCU * a.pstride + (uint32_t)t.g * a.gstride;
  };

  {
    const uint32_t to = tile_off(tile_id(a, L));
    uint4 v[RAW_LOADS];
#pragma unroll
    for (int i = 0; i < RAW_LOADS; i++) v[i] = raw_fetch(rin, to, a, wave, voff_even, voff_odd, i);
#pragma unroll
    for (int i = 0; i < RAW_LOADS; i++) raw_put(raw, wave, lane, i, v[i]);
  }

  while (L < tend) {
    const TileId t = tile_id(a, L);
    const int Lnext = L + nslot;
    // after the last tile the slices re-read this tile (in bounds, unused)
    const uint32_t to_next = tile_off(tile_id(a, Lnext < tend ? Lnext : L));
    // this tile's raw copy is in LDS (the barrier orders everyone's
    // writes); the stage was last read by the previous tile
    __syncthreads();
    raw_transpose(raw, stage, a.nant, tid);
    __syncthreads();

    const int btw = __builtin_amdgcn_readfirstlane(t.bset * (BSET / 32) + wave * BTW);
    auto wrow_bt = [&](int cq) -> uint32_t { return (uint32_t)weight_row(a, t.g, cq) * a.nbt + btw; };
    auto sc_off = [&](int cq) -> uint32_t {
      return (uint32_t)weight_row(a, t.g, cq) * a.nbeam_pad + btw * 32;
    };

    float part[2][BTW][NP];          // per 32-sample unit
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int bt = 0; bt < BTW; bt++)
#pragma unroll
        for (int j = 0; j < NP; j++) part[i][bt][j] = 0.f;

    // Unit u = 2*cq + tt.  Weights of cq+1 are requested when unit 2cq
    // starts (two units of slack); unit u fills accumulator set u&1 while the
    // detect of unit u-1 drains the other.  Raw slice cq of the next tile goes
    // out with the weight request of (channel, pol) cq.  The loop body has no
    // branches, so the wait counts the compiler derives stay exact.
    // sched_barrier keeps each (gemm u, detect u-1) pair in its own
    // scheduling region so the scheduler interleaves within it instead of
    // hoisting later units.
    WFrag<PREC> W0, W1;
    Acc<> A0, A1;
    float sc_prev[BTW];
    W0.load_buf(rw, rs, wrow_bt(0), sc_off(0), vlane16, vn4);
    W1.load_buf(rw, rs, wrow_bt(1), sc_off(1), vlane16, vn4);
    uint4 slice = raw_fetch(rin, to_next, a, wave, voff_even, voff_odd, 0);
    gemm_unit(A0, W0, stage, ROWB, n, h);                                       // unit 0
    __builtin_amdgcn_sched_barrier(0);
    gemm_unit(A1, W0, stage + 32 * ROWB, ROWB, n, h);                           // unit 1
    detect<TAVG>(A0, W0.sc, part[0]);                                           // unit 0
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int bt = 0; bt < BTW; bt++) sc_prev[bt] = W0.sc[bt];
#pragma unroll 1
    for (int cq = 1; cq < 15; cq += 2) {
      // ---- (channel, pol) cq with W1; W0 <- cq+1
      raw_put(raw, wave, lane, cq - 1, slice);
      W0.load_buf(rw, rs, wrow_bt(cq + 1), sc_off(cq + 1), vlane16, vn4);
      slice = raw_fetch(rin, to_next, a, wave, voff_even, voff_odd, cq);
      gemm_unit(A0, W1, stage + (cq * TCU) * ROWB, ROWB, n, h);                // unit 2cq
      detect<TAVG>(A1, sc_prev, part[1]);                                 // unit 2cq-1
      __builtin_amdgcn_sched_barrier(0);
      gemm_unit(A1, W1, stage + (cq * TCU + 32) * ROWB, ROWB, n, h);           // unit 2cq+1
      detect<TAVG>(A0, W1.sc, part[0]);                                          // unit 2cq
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int bt = 0; bt < BTW; bt++) sc_prev[bt] = W1.sc[bt];
      // ---- (channel, pol) cq+1 with W0; W1 <- cq+2
      raw_put(raw, wave, lane, cq, slice);
      W1.load_buf(rw, rs, wrow_bt(cq + 2), sc_off(cq + 2), vlane16, vn4);
      slice = raw_fetch(rin, to_next, a, wave, voff_even, voff_odd, cq + 1);
      gemm_unit(A0, W0, stage + ((cq + 1) * TCU) * ROWB, ROWB, n, h);          // unit 2cq+2
      detect<TAVG>(A1, sc_prev, part[1]);                                 // unit 2cq+1
      __builtin_amdgcn_sched_barrier(0);
      gemm_unit(A1, W0, stage + ((cq + 1) * TCU + 32) * ROWB, ROWB, n, h);     // unit 2cq+3
      detect<TAVG>(A0, W0.sc, part[0]);                                          // unit 2cq+2
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int bt = 0; bt < BTW; bt++) sc_prev[bt] = W0.sc[bt];
    }
    // ---- (channel, pol) 15 with W1
    raw_put(raw, wave, lane, 14, slice);
    slice = raw_fetch(rin, to_next, a, wave, voff_even, voff_odd, 15);
    gemm_unit(A0, W1, stage + (15 * TCU) * ROWB, ROWB, n, h);                   // unit 30
    detect<TAVG>(A1, sc_prev, part[1]);                                   // unit 29
    __builtin_amdgcn_sched_barrier(0);
    gemm_unit(A1, W1, stage + (15 * TCU + 32) * ROWB, ROWB, n, h);              // unit 31
    detect<TAVG>(A0, W1.sc, part[0]);                                           // unit 30
    __builtin_amdgcn_sched_barrier(0);
    detect<TAVG>(A1, W1.sc, part[1]);                                     // unit 31
    raw_put(raw, wave, lane, 15, slice);

    tile_output<TAVG>(a, t, btw, n, h, part, 0, 2);
    L = Lnext;
  }
}

template <int PREC>
static void launch(int tavg, dim3 grid, hipStream_t s, const BfArgs &a) {
  switch (tavg) {
    case 8:  hipLaunchKernelGGL((bf_pipelined_kernel<PREC, 8>), grid, dim3(NT), 0, s, a); break;
    case 16: hipLaunchKernelGGL((bf_pipelined_kernel<PREC, 16>), grid, dim3(NT), 0, s, a); break;
    case 32: hipLaunchKernelGGL((bf_pipelined_kernel<PREC, 32>), grid, dim3(NT), 0, s, a); break;
    default: hipLaunchKernelGGL((bf_pipelined_kernel<PREC, 64>), grid, dim3(NT), 0, s, a); break;
  }
}

BfArgs bf_make_args(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag,
                    const float *d_wscale, const float *d_scale, uint8_t *d_out, float *d_power,
                    float *d_bpart, int bset_beams) {
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
  a.nbsets = (d->nbeam + bset_beams - 1) / bset_beams;   // beams past nbeam: zero weights, no output
  a.ntiles = a.nbsets * d->ngroup * a.ntblk;
  a.tiles_per_xcd = (a.ntiles + 7) / 8;
  a.pstride = (uint32_t)d->nant * d->nchan * 4;
  a.astride = d->order == 0 ? (uint32_t)d->nchan * 4 : 32u;
  a.gstride = d->order == 0 ? 32u : (uint32_t)d->nant * 32;
  a.in_bytes = (uint32_t)((size_t)d->npkt * d->nant * d->nchan * 4);
  a.wfrag_bytes = (uint32_t)bfk_wfrag_bytes(d);
  a.wscale_bytes = (uint32_t)bfk_wscale_bytes(d);
  a.split = 0;
  return a;
}

extern "C" int bfk_beamform_classic(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag,
                                    const float *d_wscale, const float *d_scale, uint8_t *d_out,
                                    float *d_power, float *d_bpart, void *stream);
extern "C" int bfk_beamform_shared(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag,
                                   const float *d_wscale, const float *d_scale, uint8_t *d_out,
                                   float *d_power, float *d_bpart, void *stream);
extern "C" int bfk_beamform_shared8(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag,
                                   const float *d_wscale, const float *d_scale, uint8_t *d_out,
                                   float *d_power, float *d_bpart, void *stream);

extern "C" int bfk_classic_beams(int variant, int nkc);

extern "C" int bfk_variant_supported(const bfk_dims *d, int v) {
  if (d->nant_pad != 64 * d->nkc || (d->nkc != 1 && d->nkc != 2 && d->nkc != 4)) return 0;
  if (d->prec == 2) return v == BFK_VAR_SHARED_256;   // FP6: shared tiles only
  if (v == BFK_VAR_PIPELINED_256) return d->nkc == 1;
  if (v == BFK_VAR_SHARED_256) return 1;       // 4-wave tiles to 64 antennas, 8-wave above
  return bfk_classic_beams(v, d->nkc) > 0;
}

static int beamform_pipelined(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag,
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
  const BfArgs a = bf_make_args(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart, BSET);
  // one workgroup per CU (144 KB of LDS each), spread evenly over 8 XCDs
  int per_xcd = ncu / 8;
  if (per_xcd > a.tiles_per_xcd) per_xcd = a.tiles_per_xcd;
  const dim3 grid(8 * per_xcd);
  if (d->prec == 0) launch<0>(d->tavg, grid, (hipStream_t)stream, a);
  else launch<1>(d->tavg, grid, (hipStream_t)stream, a);
  return bfk_check(hipGetLastError(), "bf_pipelined_kernel");
}

extern "C" int bfk_beamform(const bfk_dims *d, const uint8_t *d_in, const void *d_wfrag,
                            const float *d_wscale, const float *d_scale, uint8_t *d_out,
                            float *d_power, float *d_bpart, void *stream) {
  // buffer offsets are 32-bit
  if ((size_t)d->npkt * d->nant * d->nchan * 4 > 0x7fffffffu || bfk_wfrag_bytes(d) > 0x7fffffffu) {
    bfk_set_error("bfk_beamform: block or weights above 2 GiB");
    return -1;
  }
  if (!bfk_variant_supported(d, d->variant)) {
    bfk_set_error("bfk_beamform: variant %d cannot run this shape (nant_pad %d)", d->variant, d->nant_pad);
    return -1;
  }
  if (d->variant == BFK_VAR_PIPELINED_256)
    return beamform_pipelined(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart, stream);
  if (d->variant == BFK_VAR_SHARED_256)
    return d->nkc == 1 ? bfk_beamform_shared(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart, stream)
                       : bfk_beamform_shared8(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart, stream);
  return bfk_beamform_classic(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart, stream);
}

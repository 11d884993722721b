// bf_dispatch.hip -- the fused beamformer's launch arguments and the
// dispatch to its schedules (classic: bf_classic.hip; shared: bf_shared.hip
// to 64 antennas, bf_shared8.hip above).
//
// Reference path: src/dsaX_beamformer.cu promoter (:170-193), beamformer
// (:267-416), badder (:78-112) and adder (:120-158).  The reference runs five
// launches per 32-sample "int" (640 per block) and keeps unpacked FP16 planes
// and FP32 partial powers in HBM between them.  Here one kernel does, per
// tile (one 8-channel output group x 64 samples x a set of beams):
//   (a) the packed 4-bit block is loaded in coalesced 32-byte chunks and
//       transposed into [channel][pol][sample][antenna] rows in LDS;
//   (b) per (channel, pol) a complex GEMM in the real 4M form on
//       v_mfma_f32_32x32x16_f16 (FP16) or v_mfma_f32_32x32x64_f8f6f4
//       (FP8 E4M3 or FP6 E2M3 voltages x split weights), FP32 accumulation;
//   (c) |B|^2, the sums over 2 pols x 8 channels x tavg samples, scaling and
//       the uint8 requantisation in registers, with no HBM round trip.
//
// GEMM orientation: rows (M) = samples, columns (N) = beams, K = antennas x
// {re, im}.  Each lane of a 32x32 accumulator holds one beam column and 16
// sample rows, so the time average is an in-lane sum plus one cross-lane add.
// The complex structure sits on the voltage side: with B-operand K-slots
// [Wr, Wi], the real part uses A = [Dr, -Di] and the imaginary part
// A = [Di, Dr]; weights are stored once.
#include "bf_common.h"

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
  a.layout = d->layout;
  a.nout = d->nout;
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
  if (d->prec == 3 && d->nkc > 2) return 0;          // three-part FP6: up to 128 antennas
  if (d->prec >= 2) return v == BFK_VAR_SHARED_256;   // FP6: shared tiles only
  if (v == BFK_VAR_SHARED_256) return 1;       // 4-wave tiles to 64 antennas, 8-wave above
  return bfk_classic_beams(v, d->nkc) > 0;
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
  if (d->variant == BFK_VAR_SHARED_256)
    return d->nkc == 1 ? bfk_beamform_shared(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart, stream)
                       : bfk_beamform_shared8(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart, stream);
  return bfk_beamform_classic(d, d_in, d_wfrag, d_wscale, d_scale, d_out, d_power, d_bpart, stream);
}

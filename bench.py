#!/usr/bin/env python3
"""Beamformer benchmark: one step = one full DSA block beamformed on one GPU.

Workload (synthetic, seeded): 64 antennas x 2 pols x 384 fine channels x 4096
samples of 4-bit complex voltages (Gaussian sigma 2, clipped to +-7, a tone
at SNR 10 in 4 channels from one beam direction), in capture packet order;
256 beams over +-1.75 deg on a 1.2 km E-W line with one weight per fine
channel; detect, sum 2 pols x 32 samples x 8 channels, requantise to uint8
-> [128 t][256 beams][48 ch].  Input blocks are resident in HBM (three
distinct blocks in rotation, larger than the 256 MB last-level cache) when
the timed region starts.

Three precisions run; "value" is the FP8-split path (FP8 E4M3 voltages x
hi+lo FP8 weights, FP32 accumulate) unless --precision names another.  The
line also carries the FP16 path and the FP6-split path (E2M3 voltages x hi+lo
E2M3 weights, twice the FP8 matrix rate), the error of each against the FP64
oracle on a subset of channel groups, the MFMA roofline of the fused kernel
(HIP events on the kernel's own stream) and a multithreaded FP32 CPU
baseline.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "dsa110-xengine_amd", "python"), os.path.join(ROOT, "oracle")]

METRIC = "beamformed input samples/s + % MFMA peak, FP8-split vs FP16, at +-1 LSB"
PRECISIONS = ("fp16", "fp8split", "fp6split", "fp6split3")
PEAK_TFLOPS = {"fp16": 2500.0, "fp8split": 5000.0, "fp6split": 10100.0,
               "fp6split3": 10100.0}   # MI355X dense MFMA
WEIGHT_PARTS = {"fp16": 1, "fp8split": 2, "fp6split": 2, "fp6split3": 3}   # MFMA passes
DTYPES = {"fp16": "fp16 x fp16, fp32 accumulate",
          "fp8split": "fp8-e4m3 voltages x hi+lo fp8-e4m3 weights, fp32 accumulate",
          "fp6split": "fp6-e2m3 voltages x hi+lo fp6-e2m3 weights, fp32 accumulate",
          "fp6split3": "fp6-e2m3 voltages x hi+mid+lo fp6-e2m3 weights, fp32 accumulate"}
NANT, NCHAN, NPKT, NBEAM, TAVG = 64, 384, 2048, 256, 32
NSAMP = 2 * NPKT
SAMPLES_PER_BLOCK = NANT * 2 * NCHAN * NSAMP
FLOP_PER_BLOCK = 8.0 * NBEAM * NANT * NCHAN * 2 * NSAMP
# algorithmic HBM bytes per block: voltages + fine-channel weight fragments (one
# 128-byte record per lane, channel, pol, 32-beam tile) + uint8 output + per-32-sample
# beam sums
ALG_BYTES_PER_BLOCK = (SAMPLES_PER_BLOCK + NCHAN * 2 * (NBEAM // 32) * 64 * 128
                       + (NSAMP // TAVG) * NBEAM * (NCHAN // 8) + 4 * NBEAM * (NSAMP // 32) * (NCHAN // 8))


class HipEvents:
    """hipEvent timing through the HIP runtime's C API."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.h.hipEventSynchronize.argtypes = [C.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.h.hipEventDestroy.argtypes = [C.c_void_p]
        self.h.hipDeviceSynchronize.argtypes = []

    def create(self, n):
        evs = []
        for _ in range(n):
            e = C.c_void_p()
            assert self.h.hipEventCreate(C.byref(e)) == 0
            evs.append(e)
        return evs

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, stream) == 0

    def ms(self, a, b):
        self.h.hipEventSynchronize(b)
        x = C.c_float()
        assert self.h.hipEventElapsedTime(C.byref(x), a, b) == 0
        return x.value

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0


def workload(dsaxbf, nblocks, seed=2024):
    cfg = dsaxbf.config(nant=NANT, nchan=NCHAN, npkt=NPKT, nbeam=NBEAM, tavg=TAVG,
                        weight_mode=dsaxbf.WEIGHTS_FINE, angle_mode=dsaxbf.ANGLE_SIN)
    antpos, cal, theta = dsaxbf.synth_geometry(NANT, NCHAN, NBEAM, seed, 1200.0, 0.1, 1.75)
    freqs = dsaxbf.reference_freqs(1498.75, NCHAN)
    tone_beam = 77
    gain = cal[:, 0, 0, :].copy()
    blocks = [dsaxbf.synth_block(cfg, seed + 17 * k, 2.0, 7, 10.0, theta[tone_beam],
                                 (40, 41, 200, 301), antpos, gain, freqs) for k in range(nblocks)]
    return cfg, antpos, cal, theta, freqs, blocks, tone_beam


def time_path(dsaxbf, ev, prec_name, cfg, antpos, cal, theta, freqs, blocks, steps, warmup,
              variant="auto"):
    cfg.precision = {"fp16": dsaxbf.PREC_FP16, "fp8split": dsaxbf.PREC_FP8SPLIT,
                     "fp6split": dsaxbf.PREC_FP6SPLIT, "fp6split3": dsaxbf.PREC_FP6SPLIT3}[prec_name]
    plan = dsaxbf.Plan(cfg)
    plan.set_weights(antpos, cal, freqs, theta)
    tune_ms = None
    if variant == "auto":       # the library's autotuner picks the kernel variant
        tune_ms = plan.autotune(5)
    else:
        plan.variant = {v: k for k, v in dsaxbf.VARIANT_NAMES.items()}[variant]
    var_name = dsaxbf.VARIANT_NAMES[plan.variant]
    stream = plan.stream
    nout = cfg.nout * cfg.nbeam * cfg.ngroup
    d_in = [dsaxbf.DeviceBuffer(b.nbytes) for b in blocks]
    for buf, b in zip(d_in, blocks):
        buf.upload(b)
    d_scale = dsaxbf.DeviceBuffer(4 * NBEAM)
    d_out = dsaxbf.DeviceBuffer(nout)
    d_pow = dsaxbf.DeviceBuffer(4 * nout)
    d_bp = dsaxbf.DeviceBuffer(4 * NBEAM)
    # scale: the AGC's first-block rule (median beam power) on block 0
    d_scale.upload(np.ones(NBEAM, np.float32))
    plan.execute(d_in[0], d_scale, d_out, d_pow, d_bp, stream)
    bp = d_bp.download(np.empty(NBEAM, np.float32), stream)
    agc = dsaxbf.Agc(NBEAM, cfg.ngroup, NPKT // 16)
    scale, _ = agc.update(bp)
    d_scale.upload(scale)

    def step(i, k0=None, k1=None):
        # one block: fused beamformer (uint8 out, no float power) + AGC sums
        if k0 is not None:
            ev.record(k0, stream)
        plan.execute(d_in[i % len(d_in)], d_scale, d_out, None, None, stream)
        if k1 is not None:
            ev.record(k1, stream)
        plan.beam_power(d_bp, stream)

    for i in range(warmup):
        step(i)
    ev.sync()
    e0, e1 = ev.create(2)
    ks = ev.create(2 * steps)
    ev.record(e0, stream)
    for i in range(steps):
        step(i, ks[2 * i], ks[2 * i + 1])
    ev.record(e1, stream)
    ev.sync()
    total_ms = ev.ms(e0, e1)
    kern_ms = [ev.ms(ks[2 * i], ks[2 * i + 1]) for i in range(steps)]
    # outputs of block 0 for the error report
    plan.execute(d_in[0], d_scale, d_out, d_pow, None, stream)
    out = d_out.download(np.empty((cfg.nout, NBEAM, cfg.ngroup), np.uint8), stream)
    pw = d_pow.download(np.empty((cfg.nout, NBEAM, cfg.ngroup), np.float32), stream)
    wdev = plan.get_weights()            # the weights the kernel applied, for the error report
    plan.close()
    for b in d_in + [d_scale, d_out, d_pow, d_bp]:
        b.free()
    return dict(ms_per_step=total_ms / steps, kernel_ms=float(np.mean(kern_ms)),
                kernel_ms_min=float(np.min(kern_ms)), out=out, power=pw, scale=scale,
                variant=var_name, wdev=wdev,
                autotune_ms=None if tune_ms is None else
                {dsaxbf.VARIANT_NAMES[i]: float(t) for i, t in enumerate(tune_ms) if t > 0})


def errors(orc, cfg, block, W, res, groups):
    d = orc.dims(NANT, NCHAN, NPKT, NBEAM, TAVG)
    ref = orc.power(d, 1, block, W, 0, groups[0], groups[1])
    sl = slice(groups[0], groups[1])
    r, p = ref[:, :, sl], res["power"][:, :, sl].astype(np.float64)
    rel = np.abs(p - r) / np.abs(r)
    ref_u8 = orc.requant_block(ref, res["scale"])[:, :, sl].astype(np.int64)
    du = np.abs(res["out"][:, :, sl].astype(np.int64) - ref_u8)
    du = np.minimum(du, 256 - du)          # distance modulo the uint8 wrap
    # the kernel alone: the oracle fed the weights the device applied (two groups)
    gk = (groups[0], min(groups[0] + 2, groups[1]))
    refk = orc.power(d, 1, block, res["wdev"], 0, *gk)[:, :, gk[0]:gk[1]]
    relk = np.abs(res["power"][:, :, gk[0]:gk[1]] - refk) / np.abs(refk)
    return dict(max_rel_power_err=float(rel.max()), rms_rel_power_err=float(np.sqrt((rel ** 2).mean())),
                meets_max_rel_1e3=bool(rel.max() <= 1e-3),
                frac_u8_diff_ge1=float((du >= 1).mean()), frac_u8_diff_gt1=float((du > 1).mean()),
                meets_lsb_gt1_0p1pct=bool((du > 1).mean() <= 1e-3),
                groups_checked=f"{groups[0]}..{groups[1] - 1}", mean_u8=float(res["out"].mean()),
                kernel_vs_own_weights_max_rel=float(relk.max()),
                kernel_groups_checked=f"{gk[0]}..{gk[1] - 1}")


def cpu_baseline(orc, cfg, block, W, scale, threads, target_s=10.0):
    import tempfile
    d = orc.dims(NANT, NCHAN, NPKT, NBEAM, TAVG)
    tmp = tempfile.mkdtemp(prefix="dsaxbf_cpu_")
    native = orc.native_cpu_lib(tmp)        # -march=native for this host's cores
    secs, reps = 0.0, 0
    while secs < target_s and reps < 40:
        s, _, _ = orc.cpu_beamform(d, 1, block, W, scale, 0, cfg.ngroup, threads, native)
        secs += s
        reps += 1
    build = "-O3 -march=native" if native is not None else "-O3 -march=x86-64-v4"
    return dict(value=reps * SAMPLES_PER_BLOCK / secs, unit="samples/s", cores=threads, kind="port",
                sample=f"{reps} full block(s) of the same workload, FP32 pthreads C loop "
                       f"(oracle/bf_cpu.c, {build}), {secs:.1f} s total, {threads} threads on "
                       f"{cpu_model()} ({os.cpu_count()} logical CPUs visible)",
                gflops=reps * FLOP_PER_BLOCK / secs / 1e9, s_per_block=secs / reps)


def cpu_model():
    """The host CPU's model name (lscpu's), or 'unknown CPU'."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def traffic_from_profiles(prec, kernel):
    """HBM bytes per launch of `kernel` (bf_shared_kernel, bf_classic_kernel)
    on the full block from its committed rocprof counter summary, or None when
    it has not been measured."""
    path = os.path.join(ROOT, "profiles", f"{kernel.replace('_kernel', '')}_full_block_hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"TRAFFIC_BYTES {prec} "):
            return float(line.split()[2])
    return None


def stream_bench(dsaxbf, ev, nblocks=8, seed=77):
    """The scaled array streamed from pinned host memory, FP8-split: 256
    antennas on a 2.5 km E-W line (seeded uniform positions) x 384 channels x
    4096 samples -> 1024 beams over +-3 deg, fine-channel weights, tavg 32.
    Two distinct synthetic blocks fill the pipe's two pinned slots once; then
    nblocks blocks go through it back-to-back (upload of block k+1 overlapping
    the beamforming of block k, the AGC between blocks)."""
    na, nc, npk, nb = 256, 384, 2048, 1024
    cfg = dsaxbf.config(nant=na, nchan=nc, npkt=npk, nbeam=nb, tavg=32,
                        precision=dsaxbf.PREC_FP8SPLIT, weight_mode=dsaxbf.WEIGHTS_FINE,
                        angle_mode=dsaxbf.ANGLE_SIN)
    antpos, cal, theta = dsaxbf.synth_geometry(na, nc, nb, seed, 2500.0, 0.1, 3.0)
    freqs = dsaxbf.reference_freqs(1498.75, nc)
    plan = dsaxbf.Plan(cfg)
    plan.set_weights(antpos, cal, freqs, theta)
    tune = plan.autotune(2)         # the library's autotuner picks the kernel variant
    variant = dsaxbf.VARIANT_NAMES[plan.variant]
    pipe = dsaxbf.Pipe(plan)
    t0 = time.time()
    for slot in range(2):
        pipe.input(slot)[:] = dsaxbf.synth_block(cfg, seed + slot)
    synth_s = time.time() - t0
    nin = dsaxbf.input_bytes(cfg)
    # upload alone (PCIe rate), then the kernel alone on device-resident data
    pipe.upload(0)
    t0 = time.perf_counter()
    pipe.upload(1)
    _, bp = pipe.run(0, np.full(nb, 1e-6, np.float32))   # waits for slot 0's upload and run
    _, bp = pipe.run(1, np.full(nb, 1e-6, np.float32))
    first_two_s = time.perf_counter() - t0
    d_in = dsaxbf.DeviceBuffer(nin)
    d_in.upload(pipe.input(0))
    d_sc, d_out = dsaxbf.DeviceBuffer(4 * nb), dsaxbf.DeviceBuffer(dsaxbf.output_bytes(cfg))
    d_sc.upload(np.full(nb, 1e-6, np.float32))
    stream = plan.stream
    plan.execute(d_in, d_sc, d_out, None, None, stream)
    ev.sync()
    k0, k1 = ev.create(2)
    ev.record(k0, stream)
    for _ in range(3):
        plan.execute(d_in, d_sc, d_out, None, None, stream)
    ev.record(k1, stream)
    kern_ms = ev.ms(k0, k1) / 3
    # the same block on the FP6-split path (kernel alone, device-resident data)
    cfg6 = dsaxbf.config(nant=na, nchan=nc, npkt=npk, nbeam=nb, tavg=32,
                         precision=dsaxbf.PREC_FP6SPLIT, weight_mode=dsaxbf.WEIGHTS_FINE,
                         angle_mode=dsaxbf.ANGLE_SIN)
    plan6 = dsaxbf.Plan(cfg6)
    plan6.set_weights(antpos, cal, freqs, theta)
    plan6.execute(d_in, d_sc, d_out, None, None, stream)
    ev.sync()
    k0, k1 = ev.create(2)
    ev.record(k0, stream)
    for _ in range(3):
        plan6.execute(d_in, d_sc, d_out, None, None, stream)
    ev.record(k1, stream)
    kern6_ms = ev.ms(k0, k1) / 3
    plan6.close()
    hu0, hu1 = ev.create(2)
    ev.record(hu0, stream)
    d_in.upload(pipe.input(1), stream)
    ev.record(hu1, stream)
    h2d_ms = ev.ms(hu0, hu1)
    for b in (d_in, d_sc, d_out):
        b.free()
    # the stream: nblocks back-to-back with the AGC between blocks
    agc = dsaxbf.Agc(nb, cfg.ngroup, npk // 16)
    scale, _ = agc.update(bp)
    pipe.upload(0)
    t0 = time.perf_counter()
    for k in range(nblocks):
        if k + 1 < nblocks:
            pipe.upload((k + 1) % 2)
        _, bp = pipe.run(k % 2, scale)
        scale, _ = agc.update(bp)
    wall = (time.perf_counter() - t0) / nblocks
    pipe.close()
    plan.close()
    flop = 8.0 * nb * na * nc * 2 * (2 * npk)
    wbytes = nc * 2 * (nb // 32) * (na // 64) * 64 * 128
    return {
        "workload": "256 ant x 2 pol x 384 ch x 4096 samples (805 MB packed) -> 1024 beams over "
                    "+-3 deg, fine-channel weights, tavg 32, FP8-split",
        "variant": variant,
        "autotune_ms": {dsaxbf.VARIANT_NAMES[i]: float(t) for i, t in enumerate(tune) if t > 0},
        "blocks_streamed": nblocks, "distinct_blocks": 2, "synth_s": synth_s,
        "s_per_block": wall, "realtime_margin": 0.134217728 / wall,
        "h2d_ms": h2d_ms, "h2d_GB_s": nin / (h2d_ms / 1e3) / 1e9,
        "kernel_ms": kern_ms, "kernel_tflops": flop / (kern_ms / 1e3) / 1e12,
        "kernel_mfma_frac_fp8": flop / (kern_ms / 1e3) / 1e12 / PEAK_TFLOPS["fp8split"],
        "kernel_ms_fp6split": kern6_ms,
        "kernel_algorithmic_GB_s": (nin + wbytes) / (kern_ms / 1e3) / 1e9,
        "first_two_blocks_s": first_two_s,
    }


def group_weights_bench(dsaxbf, ev, freqs, blocks, steps=10, warmup=2, seed=2024):
    """The same full blocks with group weights (one row per 8-channel group
    and pol: the reference's calc_weights at freqs[8g+4], the CLI's default),
    on the shared schedule: the FP8 and FP6 splits keep both pols' weight rows
    in registers and load no weights per (channel, pol); FP16 holds one row at
    a time, its steps pol-outer, and loads the second once per tile."""
    cfg = dsaxbf.config(nant=NANT, nchan=NCHAN, npkt=NPKT, nbeam=NBEAM, tavg=TAVG,
                        weight_mode=dsaxbf.WEIGHTS_GROUP, angle_mode=dsaxbf.ANGLE_SIN)
    antpos, cal, theta = dsaxbf.synth_geometry(NANT, cfg.ncal, NBEAM, seed, 1200.0, 0.1, 1.75)
    out = {"workload": "full DSA block as the headline, group weights (48 rows x 2 pols), tavg 32, "
                       "shared schedule (FP8/FP6 splits hold both pols' rows; FP16 one at a time, pol-outer)"}
    for prec in ("fp16", "fp8split", "fp6split"):
        r = time_path(dsaxbf, ev, prec, cfg, antpos, cal, theta, freqs, blocks, steps, warmup,
                      variant="shared256")
        out[prec] = {"kernel_ms": r["kernel_ms"], "kernel_ms_min": r["kernel_ms_min"],
                     "ms_per_step": r["ms_per_step"],
                     "samples_per_s": SAMPLES_PER_BLOCK / (r["ms_per_step"] / 1e3),
                     "mfma_frac_of_own_peak": FLOP_PER_BLOCK / (r["kernel_ms"] / 1e3) / 1e12
                     / PEAK_TFLOPS[prec]}
    for prec in ("fp8split", "fp6split"):
        out[f"{prec}_speedup_vs_fp16"] = out["fp16"]["ms_per_step"] / out[prec]["ms_per_step"]
    return out


def xengine_bench(dsaxbf, ev, reps=20, seed=91):
    """The correlator X-engine on a full block of the reference correlator's
    shape: 96 antennas x 384 channels x 4096 samples (2048 packets, 302 MB
    packed), all 4656 baselines x 2 pols, visibilities exact.  Time per block
    (the turn, the correlation and the slice sum) from HIP events on the
    plan's stream over reps executes, three distinct blocks in rotation
    (larger together than the last-level cache), for several time-slice
    counts; the library's default is the headline."""
    nant, nchan, npkt = 96, 384, 2048
    cfg = dsaxbf.config(nant=nant, nchan=nchan, npkt=npkt, nbeam=8)
    blocks = [dsaxbf.synth_block(cfg, seed + k) for k in range(3)]
    x = dsaxbf.XEngine(nant, nchan, npkt)
    default_slices = x.slices
    d_in = [dsaxbf.DeviceBuffer(b.nbytes) for b in blocks]
    for buf, b in zip(d_in, blocks):
        buf.upload(b)
    d_vis = dsaxbf.DeviceBuffer(x.vis_bytes)
    stream = x.stream
    sweep = {}
    for ns in sorted({1, 2, default_slices, 4, 8}):
        x.slices = ns
        for i in range(3):
            x.execute(d_in[i], d_vis, stream)
        ev.sync()
        marks = ev.create(2 * reps)
        for i in range(reps):
            ev.record(marks[2 * i], stream)
            x.execute(d_in[i % 3], d_vis, stream)
            ev.record(marks[2 * i + 1], stream)
        ms = [ev.ms(marks[2 * i], marks[2 * i + 1]) for i in range(reps)]
        sweep[ns] = (float(np.mean(ms)), float(np.min(ms)))
    kern_ms, kern_min = sweep[default_slices]
    nin = blocks[0].nbytes
    alg_bytes = nin + x.vis_bytes
    nbase = nant * (nant + 1) // 2
    # complex MACs of the triangle (4 real MACs each), and the int8 matrix work
    # issued: 6 32x32 tiles x 2 (re, im) 32x32x32 ops per 16 samples per (channel, pol)
    flop_alg = 8.0 * nbase * nchan * 2 * 2 * npkt
    flop_exec = 2.0 * 32 * 32 * 32 * 6 * 2 * (2 * npkt // 16) * nchan * 2
    x.close()
    for b in d_in + [d_vis]:
        b.free()
    return {"workload": "96 ant x 2 pol x 384 ch x 4096 samples (302 MB packed, capture order) -> "
                        "4656 baselines x 384 ch x 2 pols, exact integer visibilities",
            "kernels": "xc_turn_kernel + xc_kernel<3, 2, 2> (int8) + xc_reduce_kernel", "slices": default_slices,
            "ms_per_block": kern_ms, "ms_per_block_min": kern_min,
            "ms_by_slices": {str(k): v[0] for k, v in sweep.items()},
            "samples_per_s": nant * 2 * nchan * 2 * npkt / (kern_ms / 1e3),
            "roofline": {"bound": "hbm", "achieved": alg_bytes / (kern_ms / 1e3) / 1e9, "peak": 8000.0,
                         "unit": "GB/s", "frac": alg_bytes / (kern_ms / 1e3) / 1e9 / 8000.0,
                         "bytes_per_launch": alg_bytes,
                         "note": "algorithmic bytes (block in + visibilities out) over the whole "
                                 "three-launch sequence; the turn adds a block write and read"},
            "mfma_int8_executed_frac": flop_exec / (kern_ms / 1e3) / 1e15 / 5.0,
            "algorithmic_tflops": flop_alg / (kern_ms / 1e3) / 1e12,
            "realtime_fraction": kern_ms / 134.217728}


def flagger_bench(dsaxbf, ev, reps=6, seed=5, cpu_sample_beams=2):
    """The search-side RFI flagger on the reference's block: 64 beams x 16384
    times x 1024 channels uint8 (1 GiB), three synthetic blocks drawn by the
    library's noise generator (N(48, 6) bytes) in rotation.  Timed with HIP
    events on the plan's stream: the statistics pass (read 1 GiB), scaley
    fused with the statistics (read + write 1 GiB), flag fused with make_ts
    (read 1 GiB, write the masked columns), and the whole block step with its
    host work (wall clock).  The oracle's step on cpu_sample_beams beams,
    scaled to the block, is the CPU reference time."""
    import dsaxfl
    nb, nt, nc = 64, 16384, 1024
    n = nb * nt * nc
    nbc = nb * nc
    cfg = dsaxfl.config(nb, nt, nc, tsflag=1, pwr=1, seed=seed)
    plan = dsaxfl.Plan(cfg)
    s = plan.stream
    blocks = [dsaxbf.DeviceBuffer(n) for _ in range(3)]
    for k, b in enumerate(blocks):
        dsaxfl.genrand(b, n, seed + 1 + k, s)
    outs = [dsaxbf.DeviceBuffer(4 * nbc) for _ in range(4)]
    plan.calibrate(blocks)
    d_s0, d_v0 = dsaxbf.DeviceBuffer(4 * nbc), dsaxbf.DeviceBuffer(4 * nbc)
    s0, v0 = plan.reference()
    d_s0.upload(s0)
    d_v0.upload(v0)
    rng = np.random.default_rng(seed)
    mask = (rng.random((nb, nc)) < 0.05).astype(np.uint8)
    d_mask, d_bp, d_ts = dsaxbf.DeviceBuffer(nbc), dsaxbf.DeviceBuffer(4 * nb), dsaxbf.DeviceBuffer(4 * nb * nt)
    d_mask.upload(mask)
    d_bp.upload(np.ones(nb, np.float32))
    ev.sync()

    def timed(fn):
        fn(0)
        ev.sync()
        marks = ev.create(2 * reps)
        for i in range(reps):
            ev.record(marks[2 * i], s)
            fn(i % 3)
            ev.record(marks[2 * i + 1], s)
        return float(np.mean([ev.ms(marks[2 * i], marks[2 * i + 1]) for i in range(reps)]))

    t_stats = timed(lambda i: dsaxfl.stats(blocks[i], nb, nt, nc, outs[0], outs[1], outs[2], outs[3], stream=s))
    t_scaley = timed(lambda i: dsaxfl.scaley(blocks[i], nb, nt, nc, d_s0, d_v0, outs[0], outs[1], None, outs[3],
                                             stream=s))
    t_flag = timed(lambda i: dsaxfl.flag_mask(blocks[i], nb, nt, nc, d_mask, d_bp, d_ts, stream=s))
    plan.set_total_power(0.0)
    plan.step(blocks[0])                       # the first block after start-up is noise-replaced
    walls, infos = [], []
    for i in range(reps):
        t0 = time.perf_counter()
        infos.append(plan.step(blocks[i % 3]))
        walls.append(time.perf_counter() - t0)
    step_s = float(np.median(walls))
    # the oracle's step on a sample of beams, scaled to the block
    import orc
    x = np.random.default_rng(seed).normal(48, 6, (cpu_sample_beams, nt, nc))
    x = np.clip(np.rint(x), 0, 255).astype(np.uint8)
    rep = np.full(nt * nc, 48, np.uint8)
    t0 = time.perf_counter()
    orc.fl_step(x, s0[:cpu_sample_beams], v0[:cpu_sample_beams], rep, 1.0, 1e9, 5.0, 10.0, pwr=1, tsflag=1,
                exact_var=False)
    cpu_s = (time.perf_counter() - t0) * nb / cpu_sample_beams
    plan.close()
    for b in blocks + outs + [d_s0, d_v0, d_mask, d_bp, d_ts]:
        b.free()
    flagged_bytes = int(mask.sum()) * nt
    hbm = {"stats": n + 4 * 4 * nbc, "scaley_stats": 2 * n + 3 * 4 * nbc,
           "flag_make_ts": n + flagged_bytes + 4 * nb * nt}
    ms = {"stats": t_stats, "scaley_stats": t_scaley, "flag_make_ts": t_flag}
    return {"workload": "64 beams x 16384 times x 1024 channels uint8 (1 GiB), synthetic N(48,6) noise, "
                        "5% of (beam, channel) columns masked for the flag pass",
            "kernel_ms": ms,
            "roofline": {k: {"bound": "hbm", "achieved": hbm[k] / (ms[k] / 1e3) / 1e9, "peak": 8000.0,
                             "unit": "GB/s", "frac": hbm[k] / (ms[k] / 1e3) / 1e9 / 8000.0,
                             "bytes_per_launch": hbm[k]} for k in ms},
            "step_wall_s": step_s, "block_s": nt * 1e-3, "realtime_margin": nt * 1e-3 / step_s,
            "step_info": [{"tpwr": i.tpwr, "replaced": i.replaced, "n_mask": i.n_mask, "n_tsmask": i.n_tsmask}
                          for i in infos[:3]],
            "reference_kernel_passes": "scaley r+w, calc_spectrum x2, calc_varspec, calc_ppspec, flag, make_ts: "
                                       "6 reads + 1 write of the block (gpu_flagger.cu:1297-1384)",
            "cpu_oracle_step_s_per_block": cpu_s,
            "cpu_sample": f"oracle step (reference float order, bubble-sort medians) on {cpu_sample_beams} "
                          f"beams, scaled x{nb // cpu_sample_beams}, 1 thread"}


def batch_bench(dsaxbf, ev, seed=2025, nblocks=24, reps=5):
    """Uneven blocks in one graph launch: nblocks blocks of 64 antennas with
    channel counts from {48, 96, 192, 384}, beam counts from {64, 128, 256, 512}
    and time lengths from {1024, 2048, 4096} samples (seeded draws), 4-bit
    Gaussian voltages, geometry-derived fine-channel weights, tavg 32; run on
    each precision.  Reported: the graph replay time of the batch with the
    autotuned tile shapes, against the same blocks run one execute at a time
    with the default shape."""
    shapes = dsaxbf.uneven_batch_shapes(seed, nblocks)
    res = {"blocks": [{"nchan": c, "nbeam": b, "nsamp": t} for c, b, t in shapes]}
    flop = sum(8.0 * b * 64 * c * 2 * t for c, b, t in shapes)
    samples = sum(64 * 2 * c * t for c, b, t in shapes)
    for prec_name, prec in (("fp16", dsaxbf.PREC_FP16), ("fp8split", dsaxbf.PREC_FP8SPLIT),
                            ("fp6split", dsaxbf.PREC_FP6SPLIT), ("fp6split3", dsaxbf.PREC_FP6SPLIT3)):
        plans, bufs = [], []
        for i, (c, b, t) in enumerate(shapes):
            cfg = dsaxbf.config(nant=64, nchan=c, npkt=t // 2, nbeam=b, tavg=32, precision=prec,
                                weight_mode=dsaxbf.WEIGHTS_FINE, angle_mode=dsaxbf.ANGLE_SIN)
            antpos, cal, theta = dsaxbf.synth_geometry(64, c, b, seed + i, 1200.0, 0.1, 1.75)
            plan = dsaxbf.Plan(cfg)
            plan.set_weights(antpos, cal, dsaxbf.reference_freqs(1498.75, c), theta)
            d_in = dsaxbf.DeviceBuffer(dsaxbf.input_bytes(cfg))
            d_in.upload(dsaxbf.synth_block(cfg, seed + 100 + i))
            d_sc, d_out, d_bp = (dsaxbf.DeviceBuffer(4 * b), dsaxbf.DeviceBuffer(dsaxbf.output_bytes(cfg)),
                                 dsaxbf.DeviceBuffer(4 * b))
            d_sc.upload(np.full(b, 1e-4, np.float32))
            plans.append(plan)
            bufs.append((d_in, d_sc, d_out, d_bp))
        stream = plans[0].stream
        # one execute per block, default tile shape, same stream
        for p_, (d_in, d_sc, d_out, d_bp) in zip(plans, bufs):
            p_.execute(d_in, d_sc, d_out, None, d_bp, stream)
        ev.sync()
        e0, e1 = ev.create(2)
        ev.record(e0, stream)
        for _ in range(reps):
            for p_, (d_in, d_sc, d_out, d_bp) in zip(plans, bufs):
                p_.execute(d_in, d_sc, d_out, None, d_bp, stream)
        ev.record(e1, stream)
        seq_ms = ev.ms(e0, e1) / reps
        batch = dsaxbf.Batch(plans, *[[bb[k] for bb in bufs] for k in range(4)])

        def replay_ms():
            batch.launch(stream)
            ev.sync()
            a0, a1 = ev.create(2)
            ev.record(a0, stream)
            for _ in range(reps):
                batch.launch(stream)
            ev.record(a1, stream)
            return ev.ms(a0, a1) / reps
        default_ms = replay_ms()
        t0 = time.time()
        batch.autotune(3)
        tune_s = time.time() - t0
        tuned_ms = replay_ms()
        picks = [dsaxbf.VARIANT_NAMES[p_.variant] for p_ in plans]
        res[prec_name] = {
            "graph_ms": tuned_ms, "graph_ms_default_tiles": default_ms, "sequential_ms": seq_ms,
            "samples_per_s": samples / (tuned_ms / 1e3), "tflops": flop / (tuned_ms / 1e3) / 1e12,
            "mfma_frac": flop / (tuned_ms / 1e3) / 1e12 / PEAK_TFLOPS[prec_name],
            "speedup_vs_sequential": seq_ms / tuned_ms, "kernels_per_launch": batch.kernels,
            "autotune_s": tune_s,
            "tile_picks": {v: picks.count(v) for v in sorted(set(picks))}}
        batch.close()
        for p_ in plans:
            p_.close()
        for bb in bufs:
            for x in bb:
                x.free()
    res["flop_per_batch"] = flop
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", choices=list(PRECISIONS), default="fp8split")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-check", action="store_true", help="skip the oracle error report")
    ap.add_argument("--check-groups", type=int, default=8)
    ap.add_argument("--variant", default="auto",
                    choices=["auto", "classic256", "classic128s", "classic128",
                             "classic64", "shared256"],
                    help="fused-kernel variant (auto: the library's autotuner picks)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched uneven-blocks measurement (one graph launch)")
    ap.add_argument("--no-stream", action="store_true",
                    help="skip the streamed scaled-array measurement (256 antennas, 1024 beams)")
    ap.add_argument("--no-xengine", action="store_true",
                    help="skip the correlator X-engine measurement (96 antennas, full block)")
    ap.add_argument("--no-group", action="store_true",
                    help="skip the group-weights measurement (FP8/FP6 splits, full block)")
    ap.add_argument("--no-flagger", action="store_true",
                    help="skip the RFI flagger measurement (64 beams x 16384 x 1024 block)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy, three times: the "
                         "known-size kernel that scales the HBM byte counters")
    ap.add_argument("--copy-rate", action="store_true",
                    help="time a plain 1 GiB device-to-device copy (the practical read + write "
                         "rate that the block passes are compared with)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(channels shard across GPUs with no exchange; see the sharding note)", flush=True)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        n = 1 << 30
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n, "launches": 3}), flush=True)
        return
    if args.copy_rate:
        import torch
        n = 1 << 30
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        marks = [torch.cuda.Event(enable_timing=True) for _ in range(2)]   # torch's stream runs the copy
        ms = []
        for _ in range(20):
            marks[0].record()
            dst.copy_(src)
            marks[1].record()
            torch.cuda.synchronize()
            ms.append(marks[0].elapsed_time(marks[1]))
        med = float(np.median(ms))
        print(json.dumps({"copy_bytes": n, "copies": 20, "copy_ms_median": med, "copy_ms_min": float(min(ms)),
                          "read_plus_write_TB_s": 2 * n / (med / 1e3) / 1e12}), flush=True)
        return

    import dsaxbf
    ev = HipEvents()
    t0 = time.time()
    cfg, antpos, cal, theta, freqs, blocks, tone_beam = workload(dsaxbf, 3)
    print(f"# workload ready in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    res = {}
    for prec in PRECISIONS:
        res[prec] = time_path(dsaxbf, ev, prec, cfg, antpos, cal, theta, freqs, blocks,
                              args.steps, args.warmup, args.variant)
        print(f"# {prec}: {res[prec]['ms_per_step']:.4f} ms/step, kernel "
              f"{res[prec]['kernel_ms']:.4f} ms, variant {res[prec]['variant']} "
              f"(autotune ms {res[prec]['autotune_ms']})", file=sys.stderr, flush=True)

    errs = {}
    W = None
    if not args.no_check or not args.no_cpu:
        import orc
        d = orc.dims(NANT, NCHAN, NPKT, NBEAM, TAVG)
        W = orc.weights(d, 1, 1, antpos, cal, freqs, theta)
    if not args.no_check:
        g = (0, min(args.check_groups, cfg.ngroup))
        for prec in res:
            errs[prec] = errors(orc, cfg, blocks[0], W, res[prec], g)
            print(f"# {prec} errors: {errs[prec]}", file=sys.stderr, flush=True)
        peak = res["fp16"]["power"].astype(np.float64).sum(axis=(0, 2)).argmax()
        errs["tone_beam"] = tone_beam
        errs["peak_beam_fp16"] = int(peak)
    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(orc, cfg, blocks[0], W, res["fp16"]["scale"], threads)
        print(f"# cpu: {cpu}", file=sys.stderr, flush=True)

    stream = None
    if not args.no_stream:
        try:
            stream = stream_bench(dsaxbf, ev)
        except Exception as exc:             # reported, never fatal to the headline line
            stream = {"error": f"{type(exc).__name__}: {exc}"}
        print(f"# stream: {stream}", file=sys.stderr, flush=True)

    xeng = None
    if not args.no_xengine:
        try:
            xeng = xengine_bench(dsaxbf, ev)
        except Exception as exc:             # reported, never fatal to the headline line
            xeng = {"error": f"{type(exc).__name__}: {exc}"}
        print(f"# xengine: {xeng}", file=sys.stderr, flush=True)

    group = None
    if not args.no_group:
        try:
            group = group_weights_bench(dsaxbf, ev, freqs, blocks)
        except Exception as exc:             # reported, never fatal to the headline line
            group = {"error": f"{type(exc).__name__}: {exc}"}
        print(f"# group weights: {group}", file=sys.stderr, flush=True)

    flag = None
    if not args.no_flagger:
        try:
            flag = flagger_bench(dsaxbf, ev)
        except Exception as exc:             # reported, never fatal to the headline line
            flag = {"error": f"{type(exc).__name__}: {exc}"}
        print(f"# flagger: {flag}", file=sys.stderr, flush=True)

    batch = None
    if not args.no_batch:
        try:
            batch = batch_bench(dsaxbf, ev)
        except Exception as exc:             # reported, never fatal to the headline line
            batch = {"error": f"{type(exc).__name__}: {exc}"}
        print(f"# batch: { {k: v for k, v in batch.items() if k != 'blocks'} }", file=sys.stderr,
              flush=True)

    main_p = args.precision
    r = res[main_p]
    value = SAMPLES_PER_BLOCK / (r["ms_per_step"] / 1e3)
    achieved = FLOP_PER_BLOCK / (r["kernel_ms"] / 1e3) / 1e12
    kernel = {"shared256": "bf_shared_kernel"}.get(r["variant"], "bf_classic_kernel")
    line = {
        "metric": METRIC, "value": value, "unit": "samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPES[main_p],
        "data": "synthetic: seeded 4-bit complex Gaussian sigma 2 clipped +-7, tone SNR 10 in 4 "
                "channels; geometry-derived FP64 weights (random-init equivalent)",
        "config": {"workload": "full DSA block: 64 ant x 2 pol x 384 fine ch x 4096 samples "
                               "(packed 4-bit, capture order) -> 256 beams, fine-channel "
                               "weights, tavg 32 -> [128][256][48] uint8",
                   "precision": main_p, "blocks_in_rotation": len(blocks),
                   "realtime_fraction": r["ms_per_step"] / 134.217728},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_TFLOPS[main_p],
                     "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS[main_p],
                     "traffic": traffic_from_profiles(main_p, kernel),
                     "kernel": kernel, "variant": r["variant"], "kernel_ms": r["kernel_ms"],
                     "flop_per_launch": FLOP_PER_BLOCK,
                     "hbm_GB_s_algorithmic": ALG_BYTES_PER_BLOCK / (r["kernel_ms"] / 1e3) / 1e9,
                     "hbm_GB_s_measured": None if traffic_from_profiles(main_p, kernel) is None
                     else traffic_from_profiles(main_p, kernel) / (r["kernel_ms"] / 1e3) / 1e9},
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in
                                                  ("value", "unit", "cores", "kind", "sample")},
        "paths": {p: {"samples_per_s": SAMPLES_PER_BLOCK / (res[p]["ms_per_step"] / 1e3),
                      "ms_per_step": res[p]["ms_per_step"], "kernel_ms": res[p]["kernel_ms"],
                      "mfma_frac_of_own_peak": FLOP_PER_BLOCK / (res[p]["kernel_ms"] / 1e3) / 1e12
                      / PEAK_TFLOPS[p],
                      "executed_mfma_frac": WEIGHT_PARTS[p] * FLOP_PER_BLOCK
                      / (res[p]["kernel_ms"] / 1e3) / 1e12 / PEAK_TFLOPS[p],
                      "variant": res[p]["variant"], "autotune_ms": res[p]["autotune_ms"],
                      "errors_vs_fp64_oracle": errs.get(p)} for p in res},
        "fp8split_speedup_vs_fp16": res["fp16"]["ms_per_step"] / res["fp8split"]["ms_per_step"],
        "fp6split_speedup_vs_fp16": res["fp16"]["ms_per_step"] / res["fp6split"]["ms_per_step"],
        "fp6split3_speedup_vs_fp16": res["fp16"]["ms_per_step"] / res["fp6split3"]["ms_per_step"],
        "tone_check": {k: errs[k] for k in ("tone_beam", "peak_beam_fp16") if k in errs},
        "scaled_array_stream": stream,
        "uneven_batch_graph": batch,
        "group_weights": group,
        "correlator_xengine": xeng,
        "rfi_flagger": flag,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

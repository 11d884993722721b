"""ctypes binding of libdsaxbf (include/dsaxbf.h) for Python callers.

This mirrors the reference beamformer's per-block contract
(src/dsaX_beamformer.cu:939-1090): a block of packed 4-bit capture-order
voltages in, a [nsamp/tavg][nbeam][nchan/8] uint8 block out, with the
per-beam scale coming from the AGC.  Errors from the library raise
BeamformerError carrying the library's code and message, the way the
reference logs to syslog and exits on failure.

The library is loaded from the package's lib/ directory; if it is missing the
import fails loudly (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DSAXBF_LIB") or os.path.join(os.path.dirname(_HERE), "lib",
                                                         "libdsaxbf.so")

PREC_FP16 = 0
PREC_FP8SPLIT = 1
PREC_FP6SPLIT = 2
PREC_FP6SPLIT3 = 3
WEIGHTS_GROUP = 0
WEIGHTS_FINE = 1
ANGLE_THETA = 0
ANGLE_SIN = 1
STUFF_NONE = 0
STUFF_ANTS = 1
STUFF_IB = 2
ORDER_CAPTURE = 0
ORDER_SPLIT_R = 1
LAYOUT_TIME_MAJOR = 0        # [t][beam][ch], BF_BLOCK_KEY
LAYOUT_BEAM_MAJOR = 1        # [beam][t][ch], the corner turn's transmit order
VARIANT_CLASSIC_256 = 0
VARIANT_RETIRED = 1          # a persistent schedule that never won: no shape runs it
VARIANT_CLASSIC_128S = 2
VARIANT_CLASSIC_128 = 3
VARIANT_CLASSIC_64 = 4
VARIANT_SHARED_256 = 5
NVARIANT = 6
VARIANT_NAMES = {0: "classic256", 1: "retired", 2: "classic128s", 3: "classic128", 4: "classic64",
                 5: "shared256"}

ERRORS = {0: "OK", -1: "EINVAL", -2: "ENOMEM", -3: "EHIP", -4: "EIO", -5: "ESTATE", -6: "ENODEV"}


class BeamformerError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "nant", "nchan", "npkt", "nbeam", "tavg", "precision", "weight_mode",
        "angle_mode", "stuffants", "input_order", "device", "output_layout")]

    @property
    def nsamp(self):
        return 2 * self.npkt

    @property
    def ngroup(self):
        return self.nchan // 8

    @property
    def nout(self):
        return self.nsamp // self.tavg

    @property
    def ncal(self):
        return self.ngroup if self.weight_mode == WEIGHTS_GROUP else self.nchan


class AGC(C.Structure):
    _fields_ = [("nbeam", C.c_int), ("ngroup", C.c_int), ("nints", C.c_int),
                ("enabled", C.c_int), ("started", C.c_int), ("bpctr", C.c_int),
                ("medbp", C.c_float), ("many_bp", C.POINTER(C.c_float)),
                ("bp", C.POINTER(C.c_float))]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libdsaxbf not built: {LIB_PATH} (run __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "dsaxbf_config_default": (None, [C.POINTER(Config), C.c_int]),
        "dsaxbf_input_bytes": (C.c_size_t, [C.POINTER(Config)]),
        "dsaxbf_output_bytes": (C.c_size_t, [C.POINTER(Config)]),
        "dsaxbf_last_error": (C.c_char_p, []),
        "dsaxbf_create": (C.c_int, [C.POINTER(P), C.POINTER(Config)]),
        "dsaxbf_destroy": (None, [P]),
        "dsaxbf_get_config": (C.c_int, [P, C.POINTER(Config)]),
        "dsaxbf_get_stream": (P, [P]),
        "dsaxbf_read_weights_file": (C.c_int, [C.c_char_p, C.c_char_p, P, P]),
        "dsaxbf_set_weights": (C.c_int, [P, P, P, P, P, C.c_double]),
        "dsaxbf_get_weights": (C.c_int, [P, P]),
        "dsaxbf_reference_freqs": (None, [C.c_double, C.c_int, P]),
        "dsaxbf_execute": (C.c_int, [P, P, P, P, P, P, P]),
        "dsaxbf_execute_noalloc": (C.c_int, [P, P, P, P, P, P, P]),
        "dsaxbf_process_block": (C.c_int, [P, P, P, P, P, P]),
        "dsaxbf_beam_power": (C.c_int, [P, P, P]),
        "dsaxbf_set_variant": (C.c_int, [P, C.c_int]),
        "dsaxbf_get_variant": (C.c_int, [P]),
        "dsaxbf_autotune": (C.c_int, [P, C.c_int, P]),
        "dsaxbf_batch_create": (C.c_int, [C.POINTER(P), P, C.c_int, P, P, P, P]),
        "dsaxbf_batch_autotune": (C.c_int, [P, C.c_int, P]),
        "dsaxbf_batch_launch": (C.c_int, [P, P]),
        "dsaxbf_batch_kernels": (C.c_int, [P]),
        "dsaxbf_batch_destroy": (None, [P]),
        "dsaxbf_pipe_create": (C.c_int, [C.POINTER(P), P]),
        "dsaxbf_pipe_destroy": (None, [P]),
        "dsaxbf_pipe_input": (P, [P, C.c_int]),
        "dsaxbf_pipe_upload": (C.c_int, [P, C.c_int]),
        "dsaxbf_pipe_run": (C.c_int, [P, C.c_int, P, C.POINTER(P), C.POINTER(P)]),
        "dsaxbf_agc_init": (C.c_int, [C.POINTER(AGC), C.c_int, C.c_int, C.c_int, C.c_int]),
        "dsaxbf_agc_free": (None, [C.POINTER(AGC)]),
        "dsaxbf_agc_update": (C.c_int, [C.POINTER(AGC), P, P, P]),
        "dsaxbf_requant": (C.c_uint8, [C.c_float, C.c_float]),
        "dsaxbf_synth_block": (C.c_int, [C.POINTER(Config), C.c_uint64, C.c_double, C.c_int,
                                         C.c_double, C.c_double, P, C.c_int, P, P, P, P]),
        "dsaxbf_synth_geometry": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_double,
                                            C.c_double, C.c_double, P, P, P]),
        "dsaxbf_dev_alloc": (C.c_int, [C.POINTER(P), C.c_size_t, C.c_int]),
        "dsaxbf_dev_free": (C.c_int, [P]),
        "dsaxbf_copy_h2d": (C.c_int, [P, P, C.c_size_t, P]),
        "dsaxbf_copy_d2h": (C.c_int, [P, P, C.c_size_t, P]),
        "dsaxbf_sync": (C.c_int, [P]),
        "dsaxbf_antpol_rms": (C.c_int, [C.POINTER(Config), P, C.c_int, C.c_int, P]),
        "dsaxbf_xengine_create": (C.c_int, [C.POINTER(P), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
        "dsaxbf_xengine_destroy": (None, [P]),
        "dsaxbf_xengine_vis_bytes": (C.c_size_t, [C.c_int, C.c_int]),
        "dsaxbf_xengine_stream": (P, [P]),
        "dsaxbf_xengine_execute": (C.c_int, [P, P, P, P]),
        "dsaxbf_xengine_process_block": (C.c_int, [P, P, P]),
        "dsaxbf_xengine_set_slices": (C.c_int, [P, C.c_int]),
        "dsaxbf_xengine_get_slices": (C.c_int, [P]),
        "dsaxbf_version": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()

# Every symbol the C header declares (checked by the CPU test suite).
EXPORTED = [
    "dsaxbf_config_default", "dsaxbf_input_bytes", "dsaxbf_output_bytes", "dsaxbf_last_error",
    "dsaxbf_create", "dsaxbf_destroy", "dsaxbf_get_config", "dsaxbf_get_stream",
    "dsaxbf_read_weights_file", "dsaxbf_set_weights", "dsaxbf_get_weights", "dsaxbf_reference_freqs",
    "dsaxbf_execute", "dsaxbf_execute_noalloc", "dsaxbf_beam_power", "dsaxbf_set_variant", "dsaxbf_get_variant",
    "dsaxbf_autotune", "dsaxbf_batch_create", "dsaxbf_batch_autotune", "dsaxbf_batch_launch",
    "dsaxbf_batch_kernels", "dsaxbf_batch_destroy", "dsaxbf_pipe_create", "dsaxbf_pipe_destroy", "dsaxbf_pipe_input",
    "dsaxbf_pipe_upload", "dsaxbf_pipe_run", "dsaxbf_process_block", "dsaxbf_agc_init", "dsaxbf_agc_free",
    "dsaxbf_agc_update", "dsaxbf_requant", "dsaxbf_synth_block", "dsaxbf_synth_geometry",
    "dsaxbf_dev_alloc", "dsaxbf_dev_free", "dsaxbf_copy_h2d", "dsaxbf_copy_d2h", "dsaxbf_sync",
    "dsaxbf_antpol_rms", "dsaxbf_version", "dsaxbf_xengine_create", "dsaxbf_xengine_destroy",
    "dsaxbf_xengine_vis_bytes", "dsaxbf_xengine_stream", "dsaxbf_xengine_execute",
    "dsaxbf_xengine_process_block", "dsaxbf_xengine_set_slices", "dsaxbf_xengine_get_slices",
]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc: int):
    if rc != 0:
        raise BeamformerError(rc, lib.dsaxbf_last_error().decode())


def config(nant=64, nchan=384, npkt=2048, nbeam=256, tavg=8, precision=PREC_FP16,
           weight_mode=WEIGHTS_GROUP, angle_mode=ANGLE_THETA, stuffants=STUFF_NONE,
           input_order=ORDER_CAPTURE, device=0, output_layout=0) -> Config:
    c = Config()
    lib.dsaxbf_config_default(C.byref(c), nant)
    c.nchan, c.npkt, c.nbeam, c.tavg = nchan, npkt, nbeam, tavg
    c.precision, c.weight_mode, c.angle_mode = precision, weight_mode, angle_mode
    c.stuffants, c.input_order, c.device = stuffants, input_order, device
    c.output_layout = output_layout
    return c


def input_bytes(cfg: Config) -> int:
    return int(lib.dsaxbf_input_bytes(C.byref(cfg)))


def output_bytes(cfg: Config) -> int:
    return int(lib.dsaxbf_output_bytes(C.byref(cfg)))


def reference_freqs(fch1_mhz: float, nchan: int) -> np.ndarray:
    f = np.empty(nchan, np.float64)
    lib.dsaxbf_reference_freqs(fch1_mhz, nchan, _ptr(f))
    return f


def requant(x: float, scale: float) -> int:
    return int(lib.dsaxbf_requant(x, scale))


def read_weights_file(fnam: str, flagfile: str | None = None):
    antpos = np.zeros(64, np.float32)
    cal = np.zeros((64, 48, 2, 2), np.float32)
    _check(lib.dsaxbf_read_weights_file(fnam.encode(), (flagfile or "nofile").encode(),
                                        _ptr(antpos), _ptr(cal)))
    return antpos, cal


def synth_geometry(nant, ncal, nbeam, seed, length_m=1200.0, amp_jitter=0.1, half_span_deg=1.75):
    antpos = np.zeros(nant, np.float32)
    cal = np.zeros((nant, ncal, 2, 2), np.float32)
    theta = np.zeros(nbeam, np.float64)
    _check(lib.dsaxbf_synth_geometry(nant, ncal, nbeam, seed, length_m, amp_jitter,
                                     half_span_deg, _ptr(antpos), _ptr(cal), _ptr(theta)))
    return antpos, cal, theta


def synth_block(cfg: Config, seed: int, sigma=2.0, clip=7, tone_snr=0.0, tone_theta=0.0,
                tone_chans=(), antpos=None, gain=None, freqs=None) -> np.ndarray:
    blk = np.empty(input_bytes(cfg), np.uint8)
    tc = np.asarray(tone_chans, np.int32)
    g = None if gain is None else np.ascontiguousarray(gain, np.float32)
    _check(lib.dsaxbf_synth_block(C.byref(cfg), seed, sigma, clip, tone_snr, tone_theta,
                                  _ptr(tc) if len(tc) else None, len(tc),
                                  _ptr(antpos), _ptr(g), _ptr(freqs), _ptr(blk)))
    return blk


def uneven_batch_shapes(seed=2025, nblocks=24):
    """The uneven batch workload: nblocks blocks of 64 antennas whose channel
    counts come from {48, 96, 192, 384}, beam counts from {64, 128, 256, 512}
    and time lengths from {1024, 2048, 4096} samples, drawn from seed.
    Returns a list of (nchan, nbeam, nsamp)."""
    rng = np.random.default_rng(seed)
    return [(int(rng.choice([48, 96, 192, 384])), int(rng.choice([64, 128, 256, 512])),
             int(rng.choice([1024, 2048, 4096]))) for _ in range(nblocks)]


def antpol_rms(cfg: Config, block: np.ndarray, pkt: int, group: int) -> np.ndarray:
    rms = np.zeros(128, np.float32)
    _check(lib.dsaxbf_antpol_rms(C.byref(cfg), _ptr(block), pkt, group, _ptr(rms)))
    return rms.reshape(2, 64)


class DeviceBuffer:
    """Device allocation owned by Python (for bindings without torch)."""

    def __init__(self, nbytes: int, device: int = 0):
        self.ptr = C.c_void_p()
        self.nbytes = nbytes
        _check(lib.dsaxbf_dev_alloc(C.byref(self.ptr), nbytes, device))

    def upload(self, arr: np.ndarray, stream=None):
        assert arr.nbytes <= self.nbytes
        _check(lib.dsaxbf_copy_h2d(self.ptr, _ptr(np.ascontiguousarray(arr)), arr.nbytes, stream))

    def download(self, arr: np.ndarray, stream=None):
        assert arr.nbytes <= self.nbytes and arr.flags.c_contiguous
        _check(lib.dsaxbf_copy_d2h(_ptr(arr), self.ptr, arr.nbytes, stream))
        return arr

    def free(self):
        if self.ptr:
            lib.dsaxbf_dev_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Plan:
    """One beamformer plan (dsaxbf_create ... dsaxbf_destroy)."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self._h = C.c_void_p()
        _check(lib.dsaxbf_create(C.byref(self._h), C.byref(cfg)))

    @property
    def stream(self):
        return lib.dsaxbf_get_stream(self._h)

    def set_weights(self, antpos, cal, freqs, theta=None, sep_arcmin=1.0):
        a = np.ascontiguousarray(antpos, np.float32)
        c = np.ascontiguousarray(cal, np.float32)
        f = np.ascontiguousarray(freqs, np.float64)
        t = None if theta is None else np.ascontiguousarray(theta, np.float64)
        assert a.size == self.cfg.nant and c.size == self.cfg.nant * self.cfg.ncal * 4
        assert f.size == self.cfg.nchan and (t is None or t.size == self.cfg.nbeam)
        _check(lib.dsaxbf_set_weights(self._h, _ptr(a), _ptr(c), _ptr(f), _ptr(t), sep_arcmin))

    def get_weights(self) -> np.ndarray:
        """The weights as the kernels apply them: [ncal][2][nbeam][nant] complex."""
        cfg = self.cfg
        w = np.empty((cfg.ncal, 2, cfg.nbeam, cfg.nant, 2), np.float64)
        _check(lib.dsaxbf_get_weights(self._h, _ptr(w)))
        return w[..., 0] + 1j * w[..., 1]

    def process_block(self, block: np.ndarray, scale: np.ndarray, want_power=False,
                      want_beam_power=False):
        cfg = self.cfg
        block = np.ascontiguousarray(block, np.uint8)
        assert block.size == input_bytes(cfg)
        scale = np.ascontiguousarray(scale, np.float32)
        assert scale.size == cfg.nbeam
        shape = ((cfg.nout, cfg.nbeam, cfg.ngroup) if cfg.output_layout == LAYOUT_TIME_MAJOR
                 else (cfg.nbeam, cfg.nout, cfg.ngroup))
        out = np.empty(shape, np.uint8)
        power = np.empty(out.shape, np.float32) if want_power else None
        bp = np.empty(cfg.nbeam, np.float32) if want_beam_power else None
        _check(lib.dsaxbf_process_block(self._h, _ptr(block), _ptr(scale), _ptr(out),
                                        _ptr(power), _ptr(bp)))
        return out, power, bp

    def execute(self, d_in, d_scale, d_out, d_power=None, d_beam_power=None, stream=None):
        def p(x):
            return None if x is None else (x.ptr if isinstance(x, DeviceBuffer) else C.c_void_p(x))
        _check(lib.dsaxbf_execute(self._h, p(d_in), p(d_scale), p(d_out), p(d_power),
                                  p(d_beam_power), stream))

    def execute_noalloc(self, d_in, d_scale, d_out, d_power=None, d_beam_power=None, stream=None):
        """dsaxbf_execute_noalloc: the partial sums in the library's static scratch."""
        def p(x):
            return None if x is None else (x.ptr if isinstance(x, DeviceBuffer) else C.c_void_p(x))
        _check(lib.dsaxbf_execute_noalloc(self._h, p(d_in), p(d_scale), p(d_out), p(d_power),
                                          p(d_beam_power), stream))

    def beam_power(self, d_beam_power, stream=None):
        """Per-beam power sum of the last execute (device buffer of nbeam floats)."""
        p = d_beam_power.ptr if isinstance(d_beam_power, DeviceBuffer) else C.c_void_p(d_beam_power)
        _check(lib.dsaxbf_beam_power(self._h, p, stream))

    @property
    def variant(self) -> int:
        return int(lib.dsaxbf_get_variant(self._h))

    @variant.setter
    def variant(self, v: int):
        _check(lib.dsaxbf_set_variant(self._h, v))

    def autotune(self, reps: int = 5) -> np.ndarray:
        """Time every variant this plan's shape supports, keep the fastest;
        returns the mean launch time (ms) of each (-1 = cannot run)."""
        ms = np.zeros(NVARIANT, np.float32)
        _check(lib.dsaxbf_autotune(self._h, reps, _ptr(ms)))
        return ms

    def close(self):
        if self._h:
            lib.dsaxbf_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Pipe:
    """Two-slot streaming pipe over a plan (pinned, double-buffered uploads)."""

    def __init__(self, plan: "Plan"):
        self.plan = plan
        self._h = C.c_void_p()
        _check(lib.dsaxbf_pipe_create(C.byref(self._h), plan._h))
        cfg = plan.cfg
        self._nin, self._nout = input_bytes(cfg), output_bytes(cfg)
        self._shape = (cfg.nout, cfg.nbeam, cfg.ngroup)

    def input(self, slot: int) -> np.ndarray:
        """The slot's pinned host input buffer as a writable uint8 array."""
        addr = lib.dsaxbf_pipe_input(self._h, slot)
        return np.ctypeslib.as_array((C.c_uint8 * self._nin).from_address(addr))

    def upload(self, slot: int):
        _check(lib.dsaxbf_pipe_upload(self._h, slot))

    def run(self, slot: int, scale: np.ndarray):
        """Beamform the slot's uploaded block; returns copies of (out, beam_power)."""
        sc = np.ascontiguousarray(scale, np.float32)
        o, b = C.c_void_p(), C.c_void_p()
        _check(lib.dsaxbf_pipe_run(self._h, slot, _ptr(sc), C.byref(o), C.byref(b)))
        out = np.ctypeslib.as_array((C.c_uint8 * self._nout).from_address(o.value)).reshape(self._shape)
        bp = np.ctypeslib.as_array((C.c_float * self.plan.cfg.nbeam).from_address(b.value))
        return out.copy(), bp.copy()

    def close(self):
        if self._h:
            lib.dsaxbf_pipe_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Blocks of different shapes bound to device buffers, replayed as one HIP
    graph (dsaxbf_batch_*).  plans and buffers must outlive the batch."""

    def __init__(self, plans, d_in, d_scale, d_out, d_beam_power=None):
        n = len(plans)
        P = C.c_void_p

        def arr(bufs):
            return (P * n)(*[None if b is None else b.ptr for b in bufs])
        self.plans = list(plans)
        hs = (P * n)(*[p._h for p in plans])
        self._arrays = (hs, arr(d_in), arr(d_scale), arr(d_out),
                        None if d_beam_power is None else arr(d_beam_power))
        self._h = P()
        _check(lib.dsaxbf_batch_create(C.byref(self._h), hs, n, self._arrays[1], self._arrays[2],
                                       self._arrays[3], self._arrays[4]))
        self.n = n

    def autotune(self, reps: int = 3) -> np.ndarray:
        ms = np.zeros((self.n, NVARIANT), np.float32)
        _check(lib.dsaxbf_batch_autotune(self._h, reps, _ptr(ms)))
        return ms

    def launch(self, stream=None):
        _check(lib.dsaxbf_batch_launch(self._h, stream))

    @property
    def kernels(self) -> int:
        return int(lib.dsaxbf_batch_kernels(self._h))

    def close(self):
        if self._h:
            lib.dsaxbf_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class XEngine:
    """Correlator X-engine plan (dsaxbf_xengine_*): visibilities
    [nant(nant+1)/2 baselines][nchan][2 pols] of 4-bit blocks."""

    def __init__(self, nant: int, nchan: int, npkt: int, input_order=0, device=0):
        self.nant, self.nchan, self.npkt = nant, nchan, npkt
        self._h = C.c_void_p()
        _check(lib.dsaxbf_xengine_create(C.byref(self._h), nant, nchan, npkt, input_order, device))
        self.nbase = nant * (nant + 1) // 2
        self.vis_bytes = int(lib.dsaxbf_xengine_vis_bytes(nant, nchan))

    @property
    def stream(self):
        return lib.dsaxbf_xengine_stream(self._h)

    def process_block(self, block: np.ndarray) -> np.ndarray:
        block = np.ascontiguousarray(block, np.uint8)
        assert block.size == self.npkt * self.nant * self.nchan * 4
        vis = np.empty((self.nbase, self.nchan, 2, 2), np.float32)
        _check(lib.dsaxbf_xengine_process_block(self._h, _ptr(block), _ptr(vis)))
        return vis[..., 0].astype(np.float64) + 1j * vis[..., 1].astype(np.float64)

    def execute(self, d_in, d_vis, stream=None):
        _check(lib.dsaxbf_xengine_execute(self._h, d_in.ptr, d_vis.ptr, stream))

    @property
    def slices(self) -> int:
        return lib.dsaxbf_xengine_get_slices(self._h)

    @slices.setter
    def slices(self, n: int):
        _check(lib.dsaxbf_xengine_set_slices(self._h, n))

    def close(self):
        if self._h:
            lib.dsaxbf_xengine_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Agc:
    """AGC state machine (dsaX_beamformer.cu:1005-1066)."""

    def __init__(self, nbeam, ngroup, nints, enabled=True):
        self._s = AGC()
        self.nbeam = nbeam
        _check(lib.dsaxbf_agc_init(C.byref(self._s), nbeam, ngroup, nints, int(enabled)))

    def update(self, beam_power_raw):
        bp = np.ascontiguousarray(beam_power_raw, np.float32)
        scale = np.empty(self.nbeam, np.float32)
        z = C.c_int(0)
        _check(lib.dsaxbf_agc_update(C.byref(self._s), _ptr(bp), _ptr(scale), C.byref(z)))
        return scale, bool(z.value)

    @property
    def state(self):
        return dict(started=self._s.started, bpctr=self._s.bpctr, medbp=self._s.medbp)

    def __del__(self):
        try:
            lib.dsaxbf_agc_free(C.byref(self._s))
        except Exception:
            pass

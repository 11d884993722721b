"""ctypes binding of the flagger entry points of libdsaxbf (include/dsaxfl.h).

The search-side RFI flagger of the reference (src/gpu_flagger.cu) on a
[nbeam][ntime][nchan] uint8 filterbank block: device kernels taking device
buffers, the CPU mask functions taking numpy arrays, and a Plan that runs the
flagger loop's block step.  Errors raise dsaxbf.BeamformerError; there is no
CPU fallback for the device work.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from dsaxbf import DeviceBuffer, _check, _ptr, lib

P = C.c_void_p
I3 = [C.c_int] * 3


class Config(C.Structure):
    _fields_ = [("nbeam", C.c_int), ("ntime", C.c_int), ("nchan", C.c_int), ("thresh", C.c_float),
                ("tsthresh", C.c_float), ("mod_thresh", C.c_float), ("pwr", C.c_int), ("mkrand", C.c_int),
                ("tsflag", C.c_int), ("nperm", C.c_int), ("perm", C.c_int * 64), ("seed", C.c_uint32),
                ("device", C.c_int)]


class StepInfo(C.Structure):
    _fields_ = [("tpwr", C.c_float), ("replaced", C.c_int), ("n_mask", C.c_int), ("n_tsmask", C.c_int)]


_SIG = {
    "dsaxfl_stats": (C.c_int, [P] + I3 + [P, P, P, P, P, P]),
    "dsaxfl_scaley": (C.c_int, [P] + I3 + [P, P, P, P, P, P, P]),
    "dsaxfl_make_ts": (C.c_int, [P] + I3 + [P, P]),
    "dsaxfl_flag": (C.c_int, [P] + I3 + [P, C.c_int, P, P]),
    "dsaxfl_flag_mask": (C.c_int, [P] + I3 + [P, P, P, P]),
    "dsaxfl_flagts": (C.c_int, [P] + I3 + [P, C.c_int, P, P]),
    "dsaxfl_sumpulse": (C.c_int, [P, C.c_size_t, P, P]),
    "dsaxfl_genrand": (C.c_int, [P, C.c_size_t, C.c_uint32, P]),
    "dsaxfl_noise": (C.c_int, [P] + I3 + [P, P]),
    "dsaxfl_fix_zspec": (C.c_int, [P, P, C.c_int, C.c_int, P]),
    "dsaxfl_medval": (C.c_float, [P, C.c_int]),
    "dsaxfl_channflag": (None, [P, C.c_int, C.c_int, C.c_float, P, C.c_int, P]),
    "dsaxfl_simple_channflag": (None, [P, C.c_int, C.c_int, C.c_float, P]),
    "dsaxfl_simple_tsflag": (None, [P, C.c_int, C.c_int, C.c_float, P]),
    "dsaxfl_gather_mask": (C.c_int, [P, C.c_long, P]),
    "dsaxfl_calc_bpwr": (None, [P, C.c_int, C.c_int, P]),
    "dsaxfl_median_calc": (None, [P, C.c_int, C.c_int]),
    "dsaxfl_config_default": (None, [C.POINTER(Config)]),
    "dsaxfl_create": (C.c_int, [C.POINTER(P), C.POINTER(Config)]),
    "dsaxfl_destroy": (None, [P]),
    "dsaxfl_get_stream": (P, [P]),
    "dsaxfl_set_reference": (C.c_int, [P, P, P]),
    "dsaxfl_get_reference": (C.c_int, [P, P, P]),
    "dsaxfl_calibrate": (C.c_int, [P, P, C.c_int]),
    "dsaxfl_set_total_power": (C.c_int, [P, C.c_float]),
    "dsaxfl_step": (C.c_int, [P, P, C.POINTER(StepInfo)]),
}
for _n, (_r, _a) in _SIG.items():
    _f = getattr(lib, _n)
    _f.restype = _r
    _f.argtypes = _a

# Every function include/dsaxfl.h declares (checked by the CPU test suite).
EXPORTED = sorted(_SIG)


def _d(x):
    return None if x is None else (x.ptr if isinstance(x, DeviceBuffer) else C.c_void_p(x))


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def config(nbeam=64, ntime=16384, nchan=1024, **kw) -> Config:
    c = Config()
    lib.dsaxfl_config_default(C.byref(c))
    c.nbeam, c.ntime, c.nchan = nbeam, ntime, nchan
    perm = kw.pop("perm", None)
    for k, v in kw.items():
        setattr(c, k, v)
    if perm is not None:
        c.nperm = len(perm)
        for i, v in enumerate(perm):
            c.perm[i] = v
    return c


# ---- device kernels ----------------------------------------------------------

def stats(d_data, nb, nt, nc, d_spec=None, d_var=None, d_max=None, d_pp=None, d_spec_in=None, stream=None):
    _check(lib.dsaxfl_stats(_d(d_data), nb, nt, nc, _d(d_spec_in), _d(d_spec), _d(d_var), _d(d_max), _d(d_pp),
                            stream))


def scaley(d_data, nb, nt, nc, d_spec0, d_var0, d_spec=None, d_var=None, d_max=None, d_pp=None, stream=None):
    _check(lib.dsaxfl_scaley(_d(d_data), nb, nt, nc, _d(d_spec0), _d(d_var0), _d(d_spec), _d(d_var), _d(d_max),
                             _d(d_pp), stream))


def make_ts(d_data, nb, nt, nc, d_ts, stream=None):
    _check(lib.dsaxfl_make_ts(_d(d_data), nb, nt, nc, _d(d_ts), stream))


def flag(d_data, nb, nt, nc, d_midx, n, d_bpwr, stream=None):
    _check(lib.dsaxfl_flag(_d(d_data), nb, nt, nc, _d(d_midx), n, _d(d_bpwr), stream))


def flag_mask(d_data, nb, nt, nc, d_mask, d_bpwr, d_ts=None, stream=None):
    _check(lib.dsaxfl_flag_mask(_d(d_data), nb, nt, nc, _d(d_mask), _d(d_bpwr), _d(d_ts), stream))


def flagts(d_data, nb, nt, nc, d_midx, n, d_bpwr, stream=None):
    _check(lib.dsaxfl_flagts(_d(d_data), nb, nt, nc, _d(d_midx), n, _d(d_bpwr), stream))


def sumpulse(d_data, n, d_summand, stream=None):
    _check(lib.dsaxfl_sumpulse(_d(d_data), n, _d(d_summand), stream))


def genrand(d_rep, n, seed, stream=None):
    _check(lib.dsaxfl_genrand(_d(d_rep), n, seed, stream))


def noise(d_data, nb, nt, nc, d_rep, stream=None):
    _check(lib.dsaxfl_noise(_d(d_data), nb, nt, nc, _d(d_rep), stream))


def fix_zspec(d_s0, d_v0, nbc, naver, stream=None):
    _check(lib.dsaxfl_fix_zspec(_d(d_s0), _d(d_v0), nbc, naver, stream))


# ---- CPU functions -----------------------------------------------------------

def medval(a) -> float:
    a = _f32(a)
    return float(lib.dsaxfl_medval(_ptr(a), len(a)))


def mask(kind: str, spec, thr: float, perm=(), out=None):
    """channflag / simple_channflag / simple_tsflag ORed into an int32 mask."""
    s = _f32(spec)
    nb, n = s.shape
    m = np.zeros((nb, n), np.int32) if out is None else np.ascontiguousarray(out, np.int32).copy()
    if kind == "channflag":
        pm = np.ascontiguousarray(perm, np.int32)
        lib.dsaxfl_channflag(_ptr(s), nb, n, thr, _ptr(pm), len(pm), _ptr(m))
    elif kind == "simple_channflag":
        lib.dsaxfl_simple_channflag(_ptr(s), nb, n, thr, _ptr(m))
    else:
        lib.dsaxfl_simple_tsflag(_ptr(s), nb, n, thr, _ptr(m))
    return m


def gather_mask(m) -> np.ndarray:
    m = np.ascontiguousarray(m, np.int32).ravel()
    idx = np.empty(m.size, np.int32)
    k = lib.dsaxfl_gather_mask(_ptr(m), m.size, _ptr(idx))
    return idx[:k].copy()


def calc_bpwr(spec) -> np.ndarray:
    s = _f32(spec)
    b = np.empty(s.shape[0], np.float32)
    lib.dsaxfl_calc_bpwr(_ptr(s), s.shape[0], s.shape[1], _ptr(b))
    return b


def median_calc(arr) -> np.ndarray:
    a = _f32(arr).copy()
    lib.dsaxfl_median_calc(_ptr(a), a.shape[0], a.shape[1])
    return a


# ---- plan ----------------------------------------------------------------------

class Plan:
    """The flagger loop's state (dsaxfl_create ... dsaxfl_destroy)."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self._h = C.c_void_p()
        _check(lib.dsaxfl_create(C.byref(self._h), C.byref(cfg)))

    @property
    def stream(self):
        return lib.dsaxfl_get_stream(self._h)

    def set_reference(self, spec0, var0):
        _check(lib.dsaxfl_set_reference(self._h, _ptr(_f32(spec0)), _ptr(_f32(var0))))

    def reference(self):
        shape = (self.cfg.nbeam, self.cfg.nchan)
        s, v = np.empty(shape, np.float32), np.empty(shape, np.float32)
        _check(lib.dsaxfl_get_reference(self._h, _ptr(s), _ptr(v)))
        return s, v

    def calibrate(self, d_blocks):
        arr = (C.c_void_p * len(d_blocks))(*[_d(b) for b in d_blocks])
        _check(lib.dsaxfl_calibrate(self._h, arr, len(d_blocks)))

    def set_total_power(self, tpwr: float):
        _check(lib.dsaxfl_set_total_power(self._h, tpwr))

    def step(self, d_data) -> StepInfo:
        info = StepInfo()
        _check(lib.dsaxfl_step(self._h, _d(d_data), C.byref(info)))
        return info

    def close(self):
        if self._h:
            lib.dsaxfl_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

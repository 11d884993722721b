"""ctypes wrapper of the FP64 oracle and the FP32 CPU baseline (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: the checker, never the product.  Imported only by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.

The C restatement is oracle/bf_oracle.c (FP64) and oracle/bf_cpu.c (FP32);
each function there cites the reference file:line it follows.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liborc.so")


class Dims(C.Structure):
    _fields_ = [("nant", C.c_int), ("nchan", C.c_int), ("npkt", C.c_int), ("nbeam", C.c_int),
                ("tavg", C.c_int), ("input_order", C.c_int)]


class AgcState(C.Structure):
    _fields_ = [("nbeam", C.c_int), ("ngroup", C.c_int), ("nints", C.c_int), ("enabled", C.c_int),
                ("started", C.c_int), ("bpctr", C.c_int), ("medbp", C.c_float),
                ("many_bp", C.c_float * (8 * 1024))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"oracle not built: {LIB_PATH}")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_unpack_byte.argtypes = [C.c_uint8, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_in_index.restype = C.c_long
        L.orc_in_index.argtypes = [C.POINTER(Dims), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_weights.argtypes = [C.POINTER(Dims), C.c_int, C.c_int, P, P, P, P, P]
        L.orc_ref_theta.argtypes = [C.c_int, C.c_double, P]
        L.orc_power.argtypes = [C.POINTER(Dims), C.c_int, P, P, C.c_int, C.c_int, C.c_int, P]
        L.orc_requant.restype = C.c_uint8
        L.orc_requant.argtypes = [C.c_float, C.c_float]
        L.orc_beam_power.argtypes = [C.POINTER(Dims), P, P]
        L.orc_agc_init.argtypes = [C.POINTER(AgcState), C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_agc_update.restype = C.c_int
        L.orc_agc_update.argtypes = [C.POINTER(AgcState), P, P]
        for fmt in ("e4m3", "e2m3", "f16"):
            getattr(L, f"orc_{fmt}_decode").restype = C.c_double
            getattr(L, f"orc_{fmt}_decode").argtypes = [C.c_uint32]
            getattr(L, f"orc_{fmt}_encode").restype = C.c_uint32
            getattr(L, f"orc_{fmt}_encode").argtypes = [C.c_double]
        L.orc_dbnic_op.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P]
        L.orc_baseline.restype = C.c_long
        L.orc_baseline.argtypes = [C.c_int, C.c_int]
        L.orc_xcorr.argtypes = [C.POINTER(Dims), P, P]
        L.orc_quant_weights.restype = C.c_int
        L.orc_quant_weights.argtypes = [C.POINTER(Dims), C.c_int, C.c_int, P, P]
        L.cpu_beamform.restype = C.c_double
        L.cpu_beamform.argtypes = [C.POINTER(Dims), C.c_int, P, P, P, C.c_int, C.c_int, P, P,
                                   C.c_int]
        i3 = [C.c_int] * 3
        for f in ("spectrum", "maxspec", "ppspec", "make_ts"):
            getattr(L, f"orc_fl_{f}").argtypes = [P] + i3 + [P]
        for f in ("varspec", "varspec_exact", "scaley"):
            getattr(L, f"orc_fl_{f}").argtypes = [P] + i3 + [P, P]
        L.orc_fl_sumpulse.argtypes = [P, C.c_long, P]
        L.orc_fl_flag.argtypes = [P] + i3 + [P, C.c_int, P]
        L.orc_fl_flagts.argtypes = [P] + i3 + [P, C.c_int, P]
        L.orc_fl_mix.restype = C.c_uint64
        L.orc_fl_mix.argtypes = [C.c_uint64]
        L.orc_fl_genrand.argtypes = [P, C.c_long, C.c_uint32]
        L.orc_fl_fix_zspec.argtypes = [P, P, C.c_int, C.c_int]
        L.orc_fl_medval.restype = C.c_float
        L.orc_fl_medval.argtypes = [P, C.c_int]
        L.orc_fl_simple_channflag.argtypes = [P, C.c_int, C.c_int, C.c_float, P]
        L.orc_fl_simple_tsflag.argtypes = [P, C.c_int, C.c_int, C.c_float, P]
        L.orc_fl_channflag.argtypes = [P, C.c_int, C.c_int, C.c_float, P, C.c_int, P]
        L.orc_fl_gather.restype = C.c_int
        L.orc_fl_gather.argtypes = [P, C.c_long, P]
        L.orc_fl_calc_bpwr.argtypes = [P, C.c_int, C.c_int, P]
        L.orc_fl_median_calc.argtypes = [P, C.c_int, C.c_int]
        L.orc_fl_step.restype = C.c_float
        L.orc_fl_step.argtypes = [P] + i3 + [P, P, P] + [C.c_float] * 4 + [C.c_int] * 3 + \
            [P, C.c_int, C.c_int, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def dims(nant, nchan, npkt, nbeam, tavg, input_order=0) -> Dims:
    return Dims(nant, nchan, npkt, nbeam, tavg, input_order)


def unpack_byte(v: int):
    re, im = C.c_int(), C.c_int()
    lib().orc_unpack_byte(v, C.byref(re), C.byref(im))
    return re.value, im.value


def in_index(d: Dims, p, a, c, t, q) -> int:
    return int(lib().orc_in_index(C.byref(d), p, a, c, t, q))


def ref_theta(nbeam: int, sep_arcmin: float = 1.0) -> np.ndarray:
    t = np.empty(nbeam, np.float64)
    lib().orc_ref_theta(nbeam, sep_arcmin, _p(t))
    return t


def weights(d: Dims, wmode: int, amode: int, antpos, cal, freqs, theta) -> np.ndarray:
    """FP64 weights, shape [ncal][2][nbeam][nant] complex128."""
    ncal = d.nchan // 8 if wmode == 0 else d.nchan
    a = np.ascontiguousarray(antpos, np.float64)
    c = np.ascontiguousarray(cal, np.float64)
    f = np.ascontiguousarray(freqs, np.float64)
    t = np.ascontiguousarray(theta, np.float64)
    W = np.zeros((ncal, 2, d.nbeam, d.nant, 2), np.float64)
    lib().orc_weights(C.byref(d), wmode, amode, _p(a), _p(c), _p(f), _p(t), _p(W))
    return W[..., 0] + 1j * W[..., 1]


def power(d: Dims, wmode: int, block, W, stuff=0, g0=0, g1=None) -> np.ndarray:
    """FP64 averaged power [nout][nbeam][ngroup]; groups outside [g0,g1) are NaN."""
    ng = d.nchan // 8
    g1 = ng if g1 is None else min(g1, ng)
    g0 = max(g0, 0)
    Wr = np.ascontiguousarray(np.stack([W.real, W.imag], -1), np.float64)
    out = np.full((2 * d.npkt // d.tavg, d.nbeam, ng), np.nan, np.float64)
    blk = np.ascontiguousarray(block, np.uint8)
    lib().orc_power(C.byref(d), wmode, _p(blk), _p(Wr), stuff, g0, g1, _p(out))
    return out


def baseline(i: int, j: int) -> int:
    """Baseline number of antenna pair (i, j <= i)."""
    return int(lib().orc_baseline(i, j))


def xcorr(nant: int, nchan: int, npkt: int, block, input_order=0) -> np.ndarray:
    """Exact visibilities [nbase][nchan][2 pols] complex128 of a 4-bit block:
    V = sum over time of conj(d_j) d_i for baseline (i, j <= i)."""
    d = Dims(nant, nchan, npkt, 1, 2, input_order)
    nbase = nant * (nant + 1) // 2
    vis = np.zeros((nbase, nchan, 2, 2), np.float64)
    blk = np.ascontiguousarray(block, np.uint8)
    if blk.size != npkt * nant * nchan * 4:
        raise ValueError("block size does not match the shape")
    lib().orc_xcorr(C.byref(d), _p(blk), _p(vis))
    return vis[..., 0] + 1j * vis[..., 1]


def dbnic_op(out, slice_, nbps=64) -> np.ndarray:
    """dsaX_dbnic's transmit buffer of beam slice slice_ from a BF_BLOCK_KEY
    block out[nout][nbeam][nw]: [nbps][nout][nw] uint8."""
    out = np.ascontiguousarray(out, np.uint8)
    nout, nbeam, nw = out.shape
    if (slice_ + 1) * nbps > nbeam:
        raise ValueError("beam slice past the block")
    op = np.empty((nbps, nout, nw), np.uint8)
    lib().orc_dbnic_op(_p(out), nout, nbeam, nw, nbps, slice_, _p(op))
    return op


def decode(fmt: str, code: int) -> float:
    """Value of a code of format fmt ("e4m3", "e2m3", "f16")."""
    return float(getattr(lib(), f"orc_{fmt}_decode")(code))


def encode(fmt: str, v: float) -> int:
    """Round-to-nearest-even code of v in format fmt ("e4m3", "e2m3", "f16")."""
    return int(getattr(lib(), f"orc_{fmt}_encode")(v))


def quant_weights(d: Dims, wmode: int, prec: int, W) -> np.ndarray:
    """Weights W ([ncal][2][nbeam][nant] complex) through the GPU path's
    encoding of precision prec (0 FP16, 1 FP8 split, 2 FP6 split, 3 FP6
    three-part split), as the values its matrix cores multiply."""
    ncal = d.nchan // 8 if wmode == 0 else d.nchan
    if W.shape != (ncal, 2, d.nbeam, d.nant):
        raise ValueError(f"weights of shape {W.shape}, want {(ncal, 2, d.nbeam, d.nant)}")
    Wr = np.ascontiguousarray(np.stack([W.real, W.imag], -1), np.float64)
    Wq = np.empty_like(Wr)
    rc = lib().orc_quant_weights(C.byref(d), wmode, prec, _p(Wr), _p(Wq))
    if rc:
        raise ValueError(f"orc_quant_weights: error {rc}")
    return Wq[..., 0] + 1j * Wq[..., 1]


def requant(x: float, scale: float) -> int:
    return int(lib().orc_requant(x, scale))


def requant_block(power, scale) -> np.ndarray:
    """Vectorised restatement of orc_requant over [nout][nbeam][ngroup]."""
    v = (power.astype(np.float32) * np.asarray(scale, np.float32)[None, :, None]).astype(np.float32)
    r = np.rint(v.astype(np.float64))
    r = np.where(np.isnan(r), 0, r)
    r = np.clip(r, -2147483648.0, 2147483647.0).astype(np.int64)
    h = np.trunc(r / 2).astype(np.int64)   # C division truncates toward zero
    return (h & 0xFF).astype(np.uint8)


def beam_power(d: Dims, pw) -> np.ndarray:
    bp = np.zeros(d.nbeam, np.float64)
    lib().orc_beam_power(C.byref(d), _p(np.ascontiguousarray(pw, np.float64)), _p(bp))
    return bp


class Agc:
    def __init__(self, nbeam, ngroup, nints, enabled=True):
        self.s = AgcState()
        self.nbeam = nbeam
        lib().orc_agc_init(C.byref(self.s), nbeam, ngroup, nints, int(enabled))

    def update(self, bp_raw):
        bp = np.ascontiguousarray(bp_raw, np.float32)
        sc = np.empty(self.nbeam, np.float32)
        z = lib().orc_agc_update(C.byref(self.s), _p(bp), _p(sc))
        return sc, bool(z)


def native_cpu_lib(outdir: str):
    """The CPU baseline (bf_cpu.c with the oracle it calls) compiled with
    -O3 -march=native for the host this runs on, as a separate library in
    outdir; None when no C compiler is available."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        return None
    here = os.path.dirname(LIB_PATH)
    so = os.path.join(outdir, "liborc_native.so")
    cmd = [cc, "-O3", "-march=native", "-fPIC", "-std=gnu11", "-pthread", "-shared", "-o", so,
           os.path.join(here, "bf_oracle.c"), os.path.join(here, "bf_quant.c"),
           os.path.join(here, "bf_cpu.c"), "-lm", "-lpthread"]
    if subprocess.run(cmd, capture_output=True).returncode != 0:
        return None
    L = C.CDLL(so)
    L.cpu_beamform.restype = C.c_double
    L.cpu_beamform.argtypes = lib().cpu_beamform.argtypes
    return L


def cpu_beamform(d: Dims, wmode: int, block, W, scale, g0=0, g1=None, nthreads=0, native=None):
    """FP32 multithreaded CPU baseline; returns (seconds, power, out).
    native: a library from native_cpu_lib() to run instead of liborc."""
    ng = d.nchan // 8
    g1 = ng if g1 is None else g1
    Wf = np.ascontiguousarray(np.stack([W.real, W.imag], -1), np.float32)
    shape = (2 * d.npkt // d.tavg, d.nbeam, ng)
    pw = np.zeros(shape, np.float32)
    out = np.zeros(shape, np.uint8)
    sc = np.ascontiguousarray(scale, np.float32)
    secs = (native or lib()).cpu_beamform(C.byref(d), wmode, _p(np.ascontiguousarray(block, np.uint8)),
                              _p(Wf), _p(sc), g0, g1, _p(pw), _p(out), nthreads)
    return secs, pw, out


# ---- search-side RFI flagger (oracle/fl_oracle.c) ---------------------------

def _u8(a):
    return np.ascontiguousarray(a, np.uint8)


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def fl_stat(name: str, data, spec=None):
    """One of the per-(beam, channel) statistics of a [nbeam][ntime][nchan]
    block: spectrum, varspec (needs spec), varspec_exact (needs spec),
    maxspec, ppspec; or make_ts (per (beam, time))."""
    nb, nt, nc = data.shape
    out = np.empty((nb, nt) if name == "make_ts" else (nb, nc), np.float32)
    f = getattr(lib(), f"orc_fl_{name}")
    if spec is None:
        f(_p(_u8(data)), nb, nt, nc, _p(out))
    else:
        f(_p(_u8(data)), nb, nt, nc, _p(_f32(spec)), _p(out))
    return out


def fl_scaley(data, spec0, var0):
    d = _u8(data).copy()
    nb, nt, nc = d.shape
    lib().orc_fl_scaley(_p(d), nb, nt, nc, _p(_f32(spec0)), _p(_f32(var0)))
    return d


def fl_sumpulse(data, summand):
    d = _u8(data).copy()
    lib().orc_fl_sumpulse(_p(d), d.size, _p(_f32(summand)))
    return d


def fl_flag(data, midx, bpwr, ts=False):
    d = _u8(data).copy()
    nb, nt, nc = d.shape
    m = np.ascontiguousarray(midx, np.int32)
    (lib().orc_fl_flagts if ts else lib().orc_fl_flag)(_p(d), nb, nt, nc, _p(m), len(m), _p(_f32(bpwr)))
    return d


def fl_genrand(n: int, seed: int):
    r = np.empty(n, np.uint8)
    lib().orc_fl_genrand(_p(r), n, seed)
    return r


def fl_fix_zspec(s0, v0, naver: int):
    """s0, v0: [naver][...] stacks; returns the fixed copies (entry 0 holds the medians)."""
    s, v = _f32(s0).copy(), _f32(v0).copy()
    lib().orc_fl_fix_zspec(_p(s), _p(v), s[0].size, naver)
    return s, v


def fl_medval(a) -> float:
    a = _f32(a)
    return float(lib().orc_fl_medval(_p(a), len(a)))


def fl_mask(kind: str, spec, thr: float, perm=(), mask=None):
    """channflag / simple_channflag / simple_tsflag ORed into mask (int32)."""
    s = _f32(spec)
    nb, n = s.shape
    m = np.zeros((nb, n), np.int32) if mask is None else np.ascontiguousarray(mask, np.int32).copy()
    if kind == "channflag":
        pm = np.ascontiguousarray(perm, np.int32)
        lib().orc_fl_channflag(_p(s), nb, n, thr, _p(pm), len(pm), _p(m))
    elif kind == "simple_channflag":
        lib().orc_fl_simple_channflag(_p(s), nb, n, thr, _p(m))
    else:
        lib().orc_fl_simple_tsflag(_p(s), nb, n, thr, _p(m))
    return m


def fl_calc_bpwr(spec):
    s = _f32(spec)
    b = np.empty(s.shape[0], np.float32)
    lib().orc_fl_calc_bpwr(_p(s), s.shape[0], s.shape[1], _p(b))
    return b


def fl_median_calc(arr):
    a = _f32(arr).copy()
    lib().orc_fl_median_calc(_p(a), a.shape[0], a.shape[1])
    return a


def fl_step(data, spec0, var0, repval, prev_tpwr, mod_thresh, thresh, tsthresh, pwr=1, mkrand=0,
            tsflag=1, perm=(10,), bpwr=None, exact_var=False):
    """The flagger's steady-state block step; returns (block, tpwr, bpwr, n_mask, n_tsmask).
    exact_var: varspec from exact integer sums instead of the reference's float order."""
    d = _u8(data).copy()
    nb, nt, nc = d.shape
    b = np.ones(nb, np.float32) if bpwr is None else _f32(bpwr).copy()
    pm = np.ascontiguousarray(perm, np.int32)
    nm, nts = C.c_int(), C.c_int()
    t = lib().orc_fl_step(_p(d), nb, nt, nc, _p(_f32(spec0)), _p(_f32(var0)), _p(_u8(repval)),
                          prev_tpwr, mod_thresh, thresh, tsthresh, pwr, mkrand, tsflag, _p(pm), len(pm),
                          int(exact_var), _p(b), C.byref(nm), C.byref(nts))
    return d, float(t), b, nm.value, nts.value

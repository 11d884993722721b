"""GPU parity of the search-side RFI flagger (include/dsaxfl.h) against the
oracle's restatement of gpu_flagger.cu (oracle/fl_oracle.c), through the C
ABI.

Bit-exact: spectrum, maxspec, ppspec, make_ts (integer sums), the exact
variance (one rounding of the integer formula), scaley, sumpulse, flag,
flagts, the noise copy, fix_zspec and the whole block step against the oracle
step run with the exact variance.  The reference's float-order varspec is
within 1e-5 relative of the device's.  genrand: the device's FP64 log/cos may
differ from the host's in the last bit, so at most 1e-5 of the bytes may be
1 apart."""
import numpy as np
import pytest

import dsaxfl as fl
from dsaxbf import DeviceBuffer

pytestmark = pytest.mark.gpu


def _dev(a):
    a = np.ascontiguousarray(a)
    d = DeviceBuffer(max(a.nbytes, 16))
    d.upload(a)
    return d


def _get(d, shape, dtype):
    return d.download(np.empty(shape, dtype))


def _block(seed, nb, nt, nc, mean=40.0, sd=9.0):
    rng = np.random.default_rng(seed)
    x = np.clip(np.rint(rng.normal(mean, sd, (nb, nt, nc))), 0, 255).astype(np.uint8)
    x[0, :, nc // 2] = np.minimum(255, x[0, :, nc // 2].astype(int) + 60)    # a hot channel
    x[nb - 1, 1000:1100, :] = 250                                             # a burst
    return x


SHAPES = [(4, 8192, 1024), (3, 4096, 400), (2, 12288, 272)]


@pytest.mark.parametrize("nb,nt,nc", SHAPES)
def test_stats_match_oracle(orc, nb, nt, nc):
    print(f"flagger stats {nb}x{nt}x{nc}", flush=True)
    x = _block(nb + nc, nb, nt, nc)
    d = _dev(x)
    outs = {k: DeviceBuffer(nb * nc * 4) for k in ("spec", "var", "max", "pp")}
    fl.stats(d, nb, nt, nc, d_spec=outs["spec"], d_var=outs["var"], d_max=outs["max"], d_pp=outs["pp"])
    g = {k: _get(v, (nb, nc), np.float32) for k, v in outs.items()}
    spec = orc.fl_stat("spectrum", x)
    assert np.array_equal(g["spec"], spec)
    assert np.array_equal(g["max"], orc.fl_stat("maxspec", x))
    assert np.array_equal(g["pp"], orc.fl_stat("ppspec", x))
    assert np.array_equal(g["var"], orc.fl_stat("varspec_exact", x, spec))
    ref = orc.fl_stat("varspec", x, spec).astype(np.float64)
    assert (np.abs(g["var"] - ref) / ref).max() <= 1e-5
    # varspec with a spectrum argument that is not the mean
    s2 = (spec + np.float32(0.375)).astype(np.float32)
    d_s2 = _dev(s2)
    fl.stats(d, nb, nt, nc, d_var=outs["var"], d_spec_in=d_s2)
    assert np.array_equal(_get(outs["var"], (nb, nc), np.float32), orc.fl_stat("varspec_exact", x, s2))
    assert np.array_equal(_get(d, x.shape, np.uint8), x)            # the block is only read


@pytest.mark.parametrize("nc", [512, 400, 272])
def test_scaley_fused_with_stats_matches_oracle(orc, nc):
    """scaley and the statistics of the scaled block; 400 and 272 channels end
    in a part-filled 256-channel slab, whose lanes past the band must not
    write (the whole block is compared, so a stray byte anywhere fails)."""
    nb, nt = 3, 4096
    print(f"flagger scaley {nb}x{nt}x{nc}", flush=True)
    rng = np.random.default_rng(5)
    x = _block(5, nb, nt, nc)
    spec0 = rng.uniform(20, 60, (nb, nc)).astype(np.float32)
    var0 = rng.uniform(1, 200, (nb, nc)).astype(np.float32)
    var0[0, 0] = 0.0                                     # infinite scale: NaN / saturated bytes
    spec0[1, 7] = float(x[1, :, 7].mean())
    d = _dev(x)
    d_s0, d_v0 = _dev(spec0), _dev(var0)
    outs = {k: DeviceBuffer(nb * nc * 4) for k in ("spec", "var", "pp")}
    fl.scaley(d, nb, nt, nc, d_s0, d_v0, d_spec=outs["spec"], d_var=outs["var"], d_pp=outs["pp"])
    want = orc.fl_scaley(x, spec0, var0)
    got = _get(d, x.shape, np.uint8)
    assert np.array_equal(got, want)
    spec = orc.fl_stat("spectrum", want)
    assert np.array_equal(_get(outs["spec"], (nb, nc), np.float32), spec)
    assert np.array_equal(_get(outs["var"], (nb, nc), np.float32), orc.fl_stat("varspec_exact", want, spec))
    assert np.array_equal(_get(outs["pp"], (nb, nc), np.float32), orc.fl_stat("ppspec", want))
    # without statistics outputs: the scaling alone
    d2 = _dev(x)
    fl.scaley(d2, nb, nt, nc, d_s0, d_v0)
    assert np.array_equal(_get(d2, x.shape, np.uint8), want)


@pytest.mark.parametrize("nc", [1024, 400, 2048])
def test_flag_mask_with_make_ts_matches_oracle(orc, nc):
    nb, nt = 4, 4096
    print(f"flagger flag_mask + make_ts nc={nc}", flush=True)
    rng = np.random.default_rng(6)
    x = _block(6, nb, nt, nc)
    # per beam: no mask, sparse (masked chunks stored alone), 5 % and 50 %
    # (a quarter or more of a row's chunks: whole rows rewritten)
    dens = np.array([0.0, 0.005, 0.05, 0.5])[:, None]
    m = (rng.random((nb, nc)) < dens).astype(np.uint8)
    m[1, 5] = 1
    m[2, 17] = 3                                         # any nonzero byte flags
    bp = np.array([1.0, 1.3, 0.7, 4.9], np.float32)      # 48 * 4.9 = 235.2
    d, d_m, d_bp = _dev(x), _dev(m), _dev(bp)
    d_ts = DeviceBuffer(nb * nt * 4)
    fl.flag_mask(d, nb, nt, nc, d_m, d_bp, d_ts)
    midx = np.flatnonzero(m.ravel()).astype(np.int32)
    want = orc.fl_flag(x, midx, bp)
    assert np.array_equal(_get(d, x.shape, np.uint8), want)
    assert np.array_equal(_get(d_ts, (nb, nt), np.float32), _row_mean(want))
    if nc % 32 == 0:
        assert np.array_equal(_get(d_ts, (nb, nt), np.float32), orc.fl_stat("make_ts", want))
    # make_ts alone reads the block
    d2 = _dev(x)
    fl.make_ts(d2, nb, nt, nc, d_ts)
    assert np.array_equal(_get(d_ts, (nb, nt), np.float32), _row_mean(x))
    if nc % 32 == 0:
        assert np.array_equal(_get(d_ts, (nb, nt), np.float32), orc.fl_stat("make_ts", x))
    assert np.array_equal(_get(d2, x.shape, np.uint8), x)


def _row_mean(x):
    """Every row's exact integer sum over all its channels / nchan, rounded
    once to float.  The reference's make_ts (gpu_flagger.cu:419-429) gives
    each of its 32 threads nchan/32 channels, so it equals this (and the
    oracle) when nchan is a multiple of 32 and leaves the last nchan % 32
    channels out otherwise; the library sums them all."""
    return (x.sum(axis=2, dtype=np.int64) / x.shape[2]).astype(np.float32)


def test_flag_and_flagts_by_index_match_oracle(orc):
    nb, nt, nc = 3, 4096, 512
    print("flagger flag / flagts by index", flush=True)
    rng = np.random.default_rng(7)
    x = _block(7, nb, nt, nc)
    bp = np.array([1.0, 0.5, 2.0], np.float32)
    cols = np.sort(rng.choice(nb * nc, 40, replace=False)).astype(np.int32)
    rows = np.sort(rng.choice(nb * nt, 70, replace=False)).astype(np.int32)
    d, d_bp = _dev(x), _dev(bp)
    fl.flag(d, nb, nt, nc, _dev(np.append(cols, [-1, nb * nc]).astype(np.int32)), len(cols) + 2,
            d_bp)                                               # bad entries skipped
    want = orc.fl_flag(x, cols, bp)
    assert np.array_equal(_get(d, x.shape, np.uint8), want)
    fl.flagts(d, nb, nt, nc, _dev(rows), len(rows), d_bp)
    want = orc.fl_flag(want, rows, bp, ts=True)
    assert np.array_equal(_get(d, x.shape, np.uint8), want)


def test_sumpulse_matches_oracle(orc):
    print("flagger sumpulse", flush=True)
    rng = np.random.default_rng(8)
    n = 1_000_003                                        # a tail past the last 4-byte group
    x = rng.integers(0, 256, n).astype(np.uint8)
    s = rng.normal(0, 80, n).astype(np.float32)
    s[:8] = [0.25, 0.75, 1e12, -1e12, np.inf, -np.inf, np.nan, 127.5]
    d = _dev(x)
    fl.sumpulse(d, n, _dev(s))
    assert np.array_equal(_get(d, n, np.uint8), orc.fl_sumpulse(x, s))


def test_genrand_and_noise_match_oracle(orc):
    print("flagger genrand / noise", flush=True)
    nb, nt, nc = 3, 4096, 512
    n = nt * nc
    d_r = DeviceBuffer(n)
    fl.genrand(d_r, n, 12345)
    got = _get(d_r, n, np.uint8)
    want = orc.fl_genrand(n, 12345)
    diff = np.abs(got.astype(int) - want.astype(int))
    assert diff.max() <= 1 and (diff > 0).mean() <= 1e-5
    x = _block(9, nb, nt, nc)
    d = _dev(x)
    fl.noise(d, nb, nt, nc, d_r)
    assert np.array_equal(_get(d, x.shape, np.uint8), np.broadcast_to(got.reshape(nt, nc), x.shape))


@pytest.mark.parametrize("naver", [1, 3, 4, 5, 16])
def test_fix_zspec_matches_oracle(orc, naver):
    print(f"flagger fix_zspec naver={naver}", flush=True)
    rng = np.random.default_rng(naver)
    nbc = 5000
    s = rng.normal(40, 2, (naver, nbc)).astype(np.float32)
    v = rng.normal(80, 5, (naver, nbc)).astype(np.float32)
    s[:, :100] = np.round(s[:, :100])                    # ties
    d_s, d_v = _dev(s), _dev(v)
    fl.fix_zspec(d_s, d_v, nbc, naver)
    ws, wv = orc.fl_fix_zspec(s, v, naver)
    assert np.array_equal(_get(d_s, (naver, nbc), np.float32), ws)
    assert np.array_equal(_get(d_v, (naver, nbc), np.float32), wv)


def _oracle_reference(orc, blocks):
    """The start-up average: spectrum and exact varspec per block, fix_zspec,
    median_calc."""
    s = np.stack([orc.fl_stat("spectrum", b) for b in blocks])
    v = np.stack([orc.fl_stat("varspec_exact", b, sb) for b, sb in zip(blocks, s)])
    if len(blocks) > 1:
        s, v = orc.fl_fix_zspec(s, v, len(blocks))
    return orc.fl_median_calc(s[0]), orc.fl_median_calc(v[0])


@pytest.mark.parametrize("pwr,mkrand,tsflag,nc", [(1, 0, 1, 1024), (0, 0, 1, 1024), (1, 0, 0, 1024),
                                                  (0, 1, 1, 1024), (1, 0, 1, 512)])
def test_step_matches_oracle(orc, pwr, mkrand, tsflag, nc):
    nb, nt = 4, 8192
    print(f"flagger step pwr={pwr} mkrand={mkrand} tsflag={tsflag} nc={nc}", flush=True)
    cfg = fl.config(nb, nt, nc, pwr=pwr, mkrand=mkrand, tsflag=tsflag, tsthresh=6.0, seed=77)
    plan = fl.Plan(cfg)
    cal = [_block(20 + k, nb, nt, nc) for k in range(3)]
    d_cal = [_dev(b) for b in cal]
    plan.calibrate(d_cal)
    s0, v0 = plan.reference()
    ws0, wv0 = _oracle_reference(orc, cal)
    assert np.array_equal(s0, ws0) and np.array_equal(v0, wv0)
    rep = orc.fl_genrand(nt * nc, 77)
    # first block after calibration: previous total power 0, so noise (as the reference)
    x = _block(30, nb, nt, nc)
    d = _dev(x)
    info = plan.step(d)
    got = _get(d, x.shape, np.uint8)
    assert info.replaced == 1 and info.n_mask == -1
    assert (np.abs(got.astype(int) - np.broadcast_to(rep.reshape(nt, nc), x.shape).astype(int)) <= 1).all()
    # later blocks against the oracle step (exact variance)
    prev = info.tpwr
    bpwr = np.ones(nb, np.float32)
    for k in range(2):
        x = _block(40 + k, nb, nt, nc)
        x[1, :, 300 + k] += 15                          # a channel to mask (small enough to
        #                                                  stay under the power-jump test)
        d = _dev(x)
        info = plan.step(d)
        want, tp, bpwr, nm, nts = orc.fl_step(x, s0, v0, rep, prev, cfg.mod_thresh, cfg.thresh, cfg.tsthresh,
                                              pwr=pwr, mkrand=mkrand, tsflag=tsflag, perm=(10,), bpwr=bpwr,
                                              exact_var=True)
        assert info.tpwr == np.float32(tp) and info.replaced == 0
        assert (info.n_mask, info.n_tsmask) == (nm, nts), k
        got = _get(d, x.shape, np.uint8)
        if mkrand:
            diff = np.abs(got.astype(int) - want.astype(int))
            assert diff.max() <= 1 and (diff > 0).mean() <= 1e-5
        else:
            assert np.array_equal(got, want), (k, int((got != want).sum()))
            assert nm > 0
        prev = info.tpwr
    plan.close()


def test_full_reference_block_subset(orc):
    """The reference's block, 64 beams x 16384 times x 1024 channels (1 GiB),
    through the statistics pass and the fused scaley: beams 0 and 63 cut out
    and checked against the oracle bit for bit, and over the whole block the
    size-independent property that the time series of the scaled block sums
    to the spectrum's total (both integer sums of the same bytes)."""
    nb, nt, nc = 64, 16384, 1024
    print("flagger full block", flush=True)
    n = nb * nt * nc
    d = DeviceBuffer(n)
    fl.genrand(d, n, 99)                                     # N(48, 6) bytes on the device
    outs = {k: DeviceBuffer(nb * nc * 4) for k in ("spec", "var", "max", "pp")}
    fl.stats(d, nb, nt, nc, d_spec=outs["spec"], d_var=outs["var"], d_max=outs["max"], d_pp=outs["pp"])
    g = {k: _get(v, (nb, nc), np.float32) for k, v in outs.items()}
    full = _get(d, n, np.uint8).reshape(nb, nt, nc)
    for bm in (0, nb - 1):
        x = np.ascontiguousarray(full[bm:bm + 1])
        spec = orc.fl_stat("spectrum", x)
        assert np.array_equal(g["spec"][bm:bm + 1], spec), bm
        assert np.array_equal(g["var"][bm:bm + 1], orc.fl_stat("varspec_exact", x, spec)), bm
        assert np.array_equal(g["max"][bm:bm + 1], orc.fl_stat("maxspec", x)), bm
        assert np.array_equal(g["pp"][bm:bm + 1], orc.fl_stat("ppspec", x)), bm
    # scaley with the block's own spectra, then the time series of the result
    d_s0, d_v0 = _dev(g["spec"]), _dev(np.maximum(g["var"], 1.0))
    fl.scaley(d, nb, nt, nc, d_s0, d_v0, d_spec=outs["spec"])
    spec2 = _get(outs["spec"], (nb, nc), np.float32).astype(np.float64)
    d_ts = DeviceBuffer(nb * nt * 4)
    fl.make_ts(d, nb, nt, nc, d_ts)
    ts = _get(d_ts, (nb, nt), np.float32).astype(np.float64)
    assert np.allclose((ts * nc).sum(axis=1), (spec2 * nt).sum(axis=1), rtol=0, atol=0.5)
    scaled0 = _get(d, nt * nc, np.uint8).reshape(1, nt, nc)
    assert np.array_equal(scaled0, orc.fl_scaley(full[:1], g["spec"][:1], np.maximum(g["var"], 1.0)[:1]))
    for b in list(outs.values()) + [d, d_s0, d_v0, d_ts]:
        b.free()

"""Pins of the flagger oracle (oracle/fl_oracle.c, a restatement of
gpu_flagger.cu) against independent numpy / pure-Python restatements,
integer closed forms and hand cases.  CPU only."""
from fractions import Fraction

import numpy as np
import pytest

import orc

NB, NT, NC = 3, 4096, 384


def _block(seed=1, nb=NB, nt=NT, nc=NC, mean=40.0, sd=9.0):
    rng = np.random.default_rng(seed)
    x = np.clip(np.rint(rng.normal(mean, sd, (nb, nt, nc))), 0, 255).astype(np.uint8)
    # a few hot channels and bursts so the masks have something to find
    if nc > 200 and nb > 1:
        x[0, :, 200] = np.minimum(255, x[0, :, 200].astype(int) + 60)
        x[1, 1000:1100, :] = 250
    return x


def _u8_ref(val):
    """(unsigned char)(__float2uint_rn(2 * val) / 2) with the saturating cvt."""
    v2 = np.float32(2.0 * np.asarray(val, np.float64))
    r = np.rint(v2).astype(np.float64)
    u = np.where(np.isnan(r) | (r <= 0), 0, np.where(r >= 2.0 ** 32, 2 ** 32 - 1, np.nan_to_num(r)))
    return (u.astype(np.uint64) // 2 % 256).astype(np.uint8)


def _warp(c, op):
    """Pure-Python warp simulation of a 32-slot shared-memory tree, threads
    0..15 active in each step, all reads of a step before its writes."""
    c = [float(v) for v in c]
    a = 16
    while a > 0:
        new = []
        for t in range(16):
            x, y = c[t], c[t + a]
            new.append(np.float32(x + y) if op == "sum" else (y if (x < y if op == "max" else x > y) else x))
        c[:16] = new
        a //= 2
    return c


def test_spectrum_and_make_ts_are_exact_integer_means():
    x = _block()
    s = orc.fl_stat("spectrum", x)
    assert np.array_equal(s, (x.astype(np.int64).sum(axis=1) / NT).astype(np.float32))
    ts = orc.fl_stat("make_ts", x)
    assert np.array_equal(ts, (x.astype(np.int64).sum(axis=2) / NC).astype(np.float32))
    # the reference's 32 threads of nchan/32 channels each (gpu_flagger.cu:419-429):
    # at 400 channels the last 16 are left out of the sum, which is still
    # divided by 400
    y = _block(3, nc=400)
    ts = orc.fl_stat("make_ts", y)
    assert np.array_equal(ts, (y[:, :, :384].astype(np.int64).sum(axis=2) / 400).astype(np.float32))


def test_varspec_exact_is_the_rounded_integer_variance():
    x = _block(2)
    s = orc.fl_stat("spectrum", x)
    v = orc.fl_stat("varspec_exact", x, s)
    S = x.astype(np.int64).sum(axis=1)
    Q = (x.astype(np.int64) ** 2).sum(axis=1)
    want = ((NT * Q - S * S) / NT / NT).astype(np.float32)       # power-of-two n: one rounding
    assert np.array_equal(v, want)
    # with a spectrum that is not the mean: (nQ - S^2 + (S - s n)^2) / n^2
    s2 = (s + np.float32(0.375)).astype(np.float32)
    v2 = orc.fl_stat("varspec_exact", x, s2)
    d = S - s2.astype(np.float64) * NT
    assert np.allclose(v2, (NT * Q - S * S + d * d) / NT / NT, rtol=2e-7, atol=0)


def test_varspec_reference_order_is_within_fp32_accumulation_of_exact():
    x = _block(3)
    s = orc.fl_stat("spectrum", x)
    v = orc.fl_stat("varspec", x, s).astype(np.float64)
    ve = orc.fl_stat("varspec_exact", x, s).astype(np.float64)
    rel = np.abs(v - ve) / ve
    assert rel.max() < 1e-5 and rel.max() > 0       # FP32 sums of 128 squares, then a 32-leaf tree
    # one (beam, channel) restated thread by thread in numpy float32 with fma
    # done exactly (float64 product of floats, then one rounding to float32)
    col = x[1, :, 17].astype(np.float32)
    parts = []
    for t in range(32):
        acc = np.float32(0)
        for xv in col[t * 128:(t + 1) * 128]:
            val = np.float32(xv - s[1, 17])
            acc = np.float32(np.float64(val) * np.float64(val) + np.float64(acc))
        parts.append(acc)
    tot = _warp(parts, "sum")[0]
    assert np.float32(np.float64(tot) / NT) == np.float32(v[1, 17])


def test_maxspec_is_numpy_max():
    x = _block(4)
    assert np.array_equal(orc.fl_stat("maxspec", x), x.max(axis=1).astype(np.float32))


def test_ppspec_is_the_lockstep_warp_result():
    x = _block(5)
    pp = orc.fl_stat("ppspec", x)
    per = NT // 32
    binm = x.astype(np.int64).reshape(NB, 32, per // 128, 128, NC).sum(axis=3) / 128.0
    c = binm.max(axis=2)                                   # [nb][32 threads][nc]
    for bm, ch in [(0, 0), (0, 200), (1, 5), (2, 383)]:
        after_max = _warp(c[bm, :, ch], "max")
        v1 = after_max[0]
        want = np.float32(v1 - _warp(after_max, "min")[0])
        assert pp[bm, ch] == want, (bm, ch)
    # the closed form the tree fixes: max over threads - min over threads 16..31
    assert np.array_equal(pp, (c.max(axis=1) - c[:, 16:, :].min(axis=1)).astype(np.float32))


def test_ppspec_hand_cases():
    nt, nc = 4096, 32
    x = np.full((1, nt, nc), 10, np.uint8)
    assert (orc.fl_stat("ppspec", x) == 0).all()
    # channel 0: one 128-sample bin at 50 in thread 3's run, thread 20's bins at 4
    x[0, 3 * 128:4 * 128, 0] = 50
    x[0, 20 * 128:21 * 128, 0] = 4
    # channel 1: a low bin in thread 2 (lower half) is not seen by the min tree
    x[0, 2 * 128:3 * 128, 1] = 1
    pp = orc.fl_stat("ppspec", x)[0]
    assert pp[0] == 46.0 and pp[1] == 0.0 and pp[2] == 0.0


def test_scaley_matches_exact_fma_restatement():
    rng = np.random.default_rng(6)
    x = _block(6, nb=2, nt=4096, nc=64)
    spec0 = rng.uniform(20, 60, (2, 64)).astype(np.float32)
    var0 = rng.uniform(1, 200, (2, 64)).astype(np.float32)
    var0[0, 0] = 0.0                                        # inf scale: NaN / +inf / -inf bytes
    out = orc.fl_scaley(x, spec0, var0)
    idx = rng.integers(0, x.size, 3000)
    for i in idx:
        b, t, c = np.unravel_index(i, x.shape)
        a = np.float32(np.float32(x[b, t, c]) - spec0[b, c])
        with np.errstate(divide="ignore"):
            k = 6.0 / np.float64(np.sqrt(var0[b, c]))
        if np.isinf(k):
            val = np.float32(np.nan) if a == 0 else np.float32(np.inf if a > 0 else -np.inf)
        else:
            val = np.float32(float(Fraction(float(a)) * Fraction(k) + 48))
        assert out[b, t, c] == _u8_ref(val), (b, t, c)


def test_byte_conversion_edges():
    # ties go to even, wrap past 255, saturation, NaN
    data = np.zeros(9, np.uint8)
    summ = np.array([47.25, 47.75, 200, 300, -3, np.inf, np.nan, 127.25, 1e12], np.float32)
    got = orc.fl_sumpulse(data, summ)
    assert list(got) == [47, 48, 200, 44, 0, 255, 0, 127, 255]
    assert np.array_equal(got, _u8_ref(summ))


def test_flag_and_flagts():
    x = _block(7)
    bp = np.array([1.0, 1.5, 0.5], np.float32)
    midx = np.array([0 * NC + 7, 2 * NC + 300], np.int32)
    y = orc.fl_flag(x, midx, bp)
    want = x.copy()
    want[0, :, 7] = 48
    want[2, :, 300] = 24
    assert np.array_equal(y, want)
    tidx = np.array([1 * NT + 5, 2 * NT + 4095], np.int32)
    z = orc.fl_flag(x, tidx, bp, ts=True)
    want = x.copy()
    want[1, 5, 128:NC - 128] = 72
    want[2, 4095, 128:NC - 128] = 24
    assert np.array_equal(z, want)
    # the replacement byte truncates 48 * bpwr
    assert orc.fl_flag(x, [0], np.array([1.01, 1, 1], np.float32))[0, 0, 0] == 48


def _splitmix(z):
    M = (1 << 64) - 1
    z = (z + 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def test_genrand_generator_and_distribution():
    for z in (0, 1, 12345, (7 << 32) + 99):
        assert orc.lib().orc_fl_mix(z) == _splitmix(z)
    n, seed = 1 << 18, 7
    r = orc.fl_genrand(n, seed)
    assert np.array_equal(r, orc.fl_genrand(n, seed)) and not np.array_equal(r, orc.fl_genrand(n, 8))
    raw = np.array([_splitmix((seed << 32) + i) for i in range(4096)], np.uint64)
    u1 = (raw & 0xFFFFFFFF) % 16384 / 16384.0
    u2 = ((raw >> 32) & 0xFFFFFFFF) % 16384 / 16384.0
    with np.errstate(divide="ignore", invalid="ignore"):
        va = np.sqrt(-2.0 * np.log(u1)) * np.cos(2 * np.pi * u2)
        want = _u8_ref(np.float32(va * 6.0 + 48.0))
    assert (np.abs(r[:4096].astype(int) - want.astype(int)) <= 1).all()
    assert (r[:4096] == want).mean() > 0.999
    # N(48, 6) through rne(2v)/2 with a truncating /2: mean 48 - 1/4
    body = r[(r > 0) & (r < 255)].astype(np.float64)
    assert abs(body.mean() - 47.75) < 0.05 and abs(body.std() - 6) < 0.1


def test_fix_zspec_medians_and_tie_quirk():
    rng = np.random.default_rng(8)
    for naver in (3, 4, 5):
        s = rng.permutation(naver * 50).astype(np.float32).reshape(naver, 50)
        v = rng.permutation(naver * 50).astype(np.float32).reshape(naver, 50) * 0.5
        fs, fv = orc.fl_fix_zspec(s, v, naver)
        assert np.array_equal(fs[0], np.sort(s, axis=0)[naver // 2])
        assert np.array_equal(fv[0], np.sort(v, axis=0)[naver // 2])
        assert np.array_equal(fs[1:], s[1:])
    # ties: [1, 2, 2] has no entry with exactly one other <= it, so entry 0 stays
    s = np.array([[1.0], [2.0], [2.0]], np.float32)
    fs, _ = orc.fl_fix_zspec(s, s.copy(), 3)
    assert fs[0, 0] == 1.0
    s = np.array([[2.0], [1.0], [1.0]], np.float32)
    fs, _ = orc.fl_fix_zspec(s, s.copy(), 3)
    assert fs[0, 0] == 1.0


def test_medval_and_median_calc():
    rng = np.random.default_rng(9)
    for n in (1, 2, 5, 21, 128):
        a = rng.normal(size=n).astype(np.float32)
        assert orc.fl_medval(a) == np.sort(a)[(n + 1) // 2 - 1]
    arr = rng.normal(size=(6, 40)).astype(np.float32)
    assert np.array_equal(orc.fl_median_calc(arr), np.sort(arr, axis=0))


def _med(a):
    return np.sort(a)[(len(a) + 1) // 2 - 1]


def test_channflag_matches_numpy_restatement():
    rng = np.random.default_rng(10)
    spec = rng.normal(40, 1, (3, NC)).astype(np.float32)
    spec[0, 150] += 30
    spec[2, 230] -= 25
    m = orc.fl_mask("channflag", spec, 5.0, perm=(10, 131))
    want = np.zeros((3, NC), np.int32)
    for i in range(3):
        row = spec[i]
        cs = np.zeros(NC, np.float32)
        for j in range(128, NC - 128):
            w = row[j:j + 21] if NC - 128 - j >= 21 else row[NC - 128 - 21:NC - 128]
            cs[j] = np.float32(row[j] - _med(w))
        md = _med(cs[128:NC - 128])
        mad = _med(np.abs(np.float32(cs[128:NC - 128] - md)))
        thr = np.float32(5.0) * mad
        for j in range(128, NC - 128):
            if cs[j] > thr or cs[j] < -thr or j in (10, 131):
                want[i, j] = 1
    assert np.array_equal(m, want)
    assert m[0, 150] == 1 and m[2, 230] == 1 and m[:, 131].all() and not m[:, 10].any()


def test_simple_channflag_and_tsflag_match_numpy_restatement():
    rng = np.random.default_rng(11)
    spec = rng.normal(10, 1, (2, NC)).astype(np.float32)
    spec[1, 222] += 12
    m = orc.fl_mask("simple_channflag", spec, 4.0)
    want = np.zeros_like(m)
    mythr = np.float32(4.0 / np.sqrt(8.0))
    for i in range(2):
        md = _med(spec[i, 128:NC - 128])
        mad = _med(np.abs(np.float32(spec[i, 128:NC - 128] - md)))
        for j in range(128, NC - 128 - 8):
            vv = np.float32(0)
            for _ in range(8):
                vv = np.float32(vv + spec[i, j])
            vv = np.float32(np.float64(vv) / 8.0 - np.float64(md))
            if vv > np.float32(mythr * mad):
                want[i, j] = 1
    assert np.array_equal(m, want) and m[1, 222] == 1
    ts = rng.normal(5, 1, (2, NT)).astype(np.float32)
    ts[0, 3000] += 20
    ts[1, 10] += 20                                          # inside the first nt/16: raises MAD a bit
    mt = orc.fl_mask("simple_tsflag", ts, 6.0)
    want = np.zeros_like(mt)
    for i in range(2):
        md = _med(ts[i, :NT // 16])
        mad = _med(np.abs(np.float32(ts[i, :NT // 16] - md)))
        want[i] = (np.float32(ts[i] - md) > np.float32(6.0) * mad).astype(np.int32)
    assert np.array_equal(mt, want) and mt[0, 3000] == 1


def test_calc_bpwr_is_a_sequential_float_sum():
    rng = np.random.default_rng(12)
    spec = rng.uniform(30, 60, (4, NC)).astype(np.float32)
    b = orc.fl_calc_bpwr(spec)
    want = [np.float32(np.float64(np.cumsum(spec[i], dtype=np.float32)[-1]) / (NC - 256) / 48.0)
            for i in range(4)]
    assert np.array_equal(b, np.array(want, np.float32))


def test_step_noise_replacement_and_quiet_block():
    x = _block(13)
    spec0 = orc.fl_stat("spectrum", x)
    var0 = orc.fl_stat("varspec", x, spec0)
    rep = orc.fl_genrand(NT * NC, 3)
    # total power jump beyond mod_thresh: every beam becomes the random block
    y, tp, _, nm, nts = orc.fl_step(x, spec0, var0, rep, prev_tpwr=1.0, mod_thresh=0.01, thresh=5.0,
                                    tsthresh=6.0)
    assert nm == -1 and (y == rep.reshape(NT, NC)[None]).all()
    # a quiet block with masks off (huge thresholds): scaley output, then the
    # permanent channel 10 is not in 128..nchan-129, so nothing is flagged
    sc = orc.fl_scaley(x, spec0, var0)
    tp0 = float(np.cumsum(orc.fl_stat("spectrum", sc).ravel(), dtype=np.float32)[-1])
    y, tp, bp, nm, nts = orc.fl_step(x, spec0, var0, rep, prev_tpwr=tp0, mod_thresh=0.5, thresh=1e9,
                                     tsthresh=1e9)
    assert tp == tp0 and nm == 0 and nts == 0 and np.array_equal(y, sc)
    assert np.allclose(bp, orc.fl_calc_bpwr(orc.fl_stat("spectrum", sc)))
    # normal thresholds: the hot channel and the burst rows are flagged
    y, tp, bp, nm, nts = orc.fl_step(x, spec0, var0, rep, prev_tpwr=tp0, mod_thresh=0.5, thresh=5.0,
                                     tsthresh=6.0)
    assert nm > 0 and nts > 0
    assert (y[1, 1000:1100, 128:NC - 128] == np.uint8(48 * bp[1])).all()

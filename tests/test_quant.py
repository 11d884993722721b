"""Pins of the oracle's weight encodings (oracle/bf_quant.c), CPU only.

The GPU path multiplies voltages by weights encoded as binary16 (the
reference's own __float2half weights, dsaX_beamformer.cu:442-446) or as sums
of narrow floating-point parts (E4M3 or E2M3).  The oracle restates those
encodings so that its power can be computed with the very weights the matrix
cores see.  Here the restatement is pinned:
  - E4M3 against torch's float8_e4m3fn conversion (a library routine) on every
    code, every midpoint and random values;
  - E2M3 against the OCP MX format table written out from the specification,
    and its encoder against a brute-force nearest-code search;
  - binary16 against numpy's float16 conversion;
  - the split weights against error bounds fixed by the formats, chosen so that
    a dropped part, a wrong boost or a wrong row scale fails.
"""
import numpy as np
import pytest
import torch

# OCP Microscaling Formats v1.0, FP6 E2M3: the 32 non-negative values.
E2M3_TABLE = [0.0, 0.125, 0.25, 0.375, 0.5, 0.625, 0.75, 0.875,
              1.0, 1.125, 1.25, 1.375, 1.5, 1.625, 1.75, 1.875,
              2.0, 2.25, 2.5, 2.75, 3.0, 3.25, 3.5, 3.75,
              4.0, 4.5, 5.0, 5.5, 6.0, 6.5, 7.0, 7.5]


def _torch_e4m3(v):
    t = torch.tensor(np.asarray(v, np.float32)).to(torch.float8_e4m3fn)
    return t.view(torch.uint8).numpy().astype(np.int64)


def test_e4m3_decode_every_code_matches_torch(orc):
    codes = [c for c in range(256) if c & 0x7F != 0x7F]           # 0x7F / 0xFF are NaN
    want = torch.tensor(codes, dtype=torch.uint8).view(torch.float8_e4m3fn).to(torch.float64).numpy()
    got = np.array([orc.decode("e4m3", c) for c in codes])
    assert np.array_equal(got, want)
    assert np.isnan(orc.decode("e4m3", 0x7F))
    assert orc.decode("e4m3", 0x7E) == 448.0 and orc.decode("e4m3", 1) == 2.0 ** -9


def test_e4m3_encode_roundtrip_midpoints_and_random(orc):
    for c in range(256):
        if c & 0x7F != 0x7F:
            assert orc.encode("e4m3", orc.decode("e4m3", c)) == c, c
    # every midpoint between neighbouring codes goes to the even code
    pos = [orc.decode("e4m3", c) for c in range(0x7F)]
    mids = [(pos[c] + pos[c + 1]) / 2 for c in range(0x7E)]
    for c, m in enumerate(mids):
        assert orc.encode("e4m3", m) == (c if c % 2 == 0 else c + 1), (c, m)
        assert orc.encode("e4m3", -m) == 0x80 | (c if c % 2 == 0 else c + 1)
    assert np.array_equal([orc.encode("e4m3", m) for m in mids], _torch_e4m3(mids))
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.uniform(-460, 460, 3000), rng.normal(0, 1e-2, 2000),
                        rng.normal(0, 3, 2000)]).astype(np.float32)
    assert np.array_equal([orc.encode("e4m3", float(x)) for x in v], _torch_e4m3(v))
    # saturating (torch's float8_e4m3fn gives NaN above 464)
    assert orc.encode("e4m3", 1000.0) == 0x7E and orc.encode("e4m3", -1e30) == 0xFE
    assert orc.encode("e4m3", 463.9) == 0x7E


def test_e2m3_decode_matches_ocp_table(orc):
    assert [orc.decode("e2m3", c) for c in range(32)] == E2M3_TABLE
    assert [orc.decode("e2m3", 32 + c) for c in range(32)] == [-x for x in E2M3_TABLE]


def _brute_e2m3(v):
    """Nearest of the 32 magnitudes, ties to the even code, saturating."""
    a = abs(v)
    best = min(range(32), key=lambda c: (abs(E2M3_TABLE[c] - a), c % 2))
    return (32 if np.signbit(v) else 0) | best


def test_e2m3_encode_matches_brute_force(orc):
    for c in range(64):
        assert orc.encode("e2m3", orc.decode("e2m3", c)) == c
    mids = [(E2M3_TABLE[c] + E2M3_TABLE[c + 1]) / 2 for c in range(31)]
    for c, m in enumerate(mids):
        assert orc.encode("e2m3", m) == (c if c % 2 == 0 else c + 1)
    rng = np.random.default_rng(1)
    for v in np.concatenate([rng.uniform(-8, 8, 4000), rng.normal(0, 0.2, 1000), mids,
                             np.nextafter(mids, 0), np.nextafter(mids, 9)]):
        assert orc.encode("e2m3", float(v)) == _brute_e2m3(float(v)), v
    assert orc.encode("e2m3", 100.0) == 0x1F and orc.encode("e2m3", -7.9) == 0x3F


def test_f16_matches_numpy(orc):
    codes = np.arange(0, 0x10000, 37, dtype=np.uint32)
    codes = codes[(codes & 0x7C00) != 0x7C00]
    want = codes.astype(np.uint16).view(np.float16).astype(np.float64)
    assert np.array_equal([orc.decode("f16", int(c)) for c in codes], want)
    rng = np.random.default_rng(2)
    v = np.concatenate([rng.uniform(-6.5e4, 6.5e4, 2000), rng.normal(0, 1e-5, 2000),
                        rng.uniform(16384, 32768, 2000)])
    want = v.astype(np.float16).view(np.uint16).astype(np.int64)
    assert np.array_equal([orc.encode("f16", float(x)) for x in v], want)
    # ties: 2049 sits between 2048 and 2050 -> even mantissa 2048
    assert orc.decode("f16", orc.encode("f16", 2049.0)) == 2048.0
    assert orc.decode("f16", orc.encode("f16", 2051.0)) == 2052.0


def _weights(orc, bf, nant=48, nchan=16, nbeam=24, spread=0.1, seed=5):
    antpos, cal, theta = bf.synth_geometry(nant, nchan, nbeam, seed, 1200.0, spread, 1.75)
    d = orc.dims(nant, nchan, 64, nbeam, 8)
    return d, orc.weights(d, 1, 1, antpos, cal, bf.reference_freqs(1498.75, nchan), theta)


def test_quant_fp16_is_rowscaled_binary16(orc, bf):
    d, W = _weights(orc, bf)
    Wq = orc.quant_weights(d, 1, 0, W)
    comp = np.stack([W.real, W.imag], -1)
    mx = np.abs(comp).max(axis=(-1, -2), keepdims=True)
    sc = 2.0 ** (14 - np.floor(np.log2(mx)))                     # max * sc in [2^14, 2^15)
    want = (comp * sc).astype(np.float16).astype(np.float64) / sc
    assert np.array_equal(np.stack([Wq.real, Wq.imag], -1), want)


def test_quant_fp8_split_error_bound(orc, bf):
    """hi + lo E4M3 with the row's largest component in [128, 256): each
    component is within 2^-8 of itself plus half a subnormal step, and the
    bound is reached (a dropped lo part would be off by up to 8 of 256)."""
    d, W = _weights(orc, bf)
    Wq = orc.quant_weights(d, 1, 1, W)
    comp, q = np.stack([W.real, W.imag], -1), np.stack([Wq.real, Wq.imag], -1)
    mx = np.abs(comp).max(axis=(-1, -2), keepdims=True)
    sc = 2.0 ** (7 - np.floor(np.log2(mx)))
    x, e = comp * sc, np.abs(comp - q) * sc
    assert (e <= 2.0 ** -8 * np.abs(x) + 2.0 ** -10).all()
    # the largest components (|x| in [128, 256): hi step 16, |lo| < 8, lo step
    # <= 0.5) are off by at most 0.25, and that is reached
    assert 0.2 < e.max() <= 0.25
    assert (np.abs(x).max(axis=(-1, -2)) >= 128).all() and (np.abs(x).max(axis=(-1, -2)) < 256).all()


@pytest.mark.parametrize("spread", [0.1, 3.0])
def test_quant_fp6_split_lane_bounds(orc, bf, spread):
    """FP6 splits: x = w * 7.5/max, lanes of 16 antennas.  With the lane's
    largest |x| = m, the hi part's boost k puts m * 2^k in (3.75, 7.5], so the
    hi rounding error is at most 2^-k/4 <= m/15; each further part cuts the
    remaining error the same way.  Two parts: <= m/15 * 1/15; three parts:
    <= m/15^3.  Exact E2M3 values pass through unchanged."""
    d, W = _weights(orc, bf, spread=spread)
    comp = np.stack([W.real, W.imag], -1)
    mx = np.abs(comp).max(axis=(-1, -2), keepdims=True)
    x = comp * 7.5 / mx
    lanes = x.reshape(*x.shape[:-2], -1, 16, 2)            # [..., lane, 16 ants, re/im]
    m = np.abs(lanes).max(axis=(-1, -2), keepdims=True)
    errs = {}
    for prec, nparts in ((2, 2), (3, 3)):
        q = orc.quant_weights(d, 1, prec, W)
        xq = (np.stack([q.real, q.imag], -1) * 7.5 / mx).reshape(lanes.shape)
        err = np.abs(xq - lanes)
        assert (err <= m / 15.0 ** nparts * (1 + 1e-12)).all(), prec
        errs[prec] = err.max()
    assert errs[3] < errs[2] / 8
    # exact E2M3 codes (the row maximum is 7.5 already): reproduced bit for bit
    codes = np.array(E2M3_TABLE)[np.random.default_rng(3).integers(0, 32, size=(1, 2, 3, 32, 2))]
    codes[..., 0, 0] = 7.5
    Wc = codes[..., 0] + 1j * codes[..., 1]
    dc = orc.dims(32, 8, 64, 3, 8)
    for prec in (2, 3):
        q = orc.quant_weights(dc, 0, prec, Wc)      # group mode: one weight channel
        assert np.array_equal(q, Wc)


def test_quant_power_error_ordering(orc, bf):
    """Through the oracle's power: FP16 and three-part FP6 weights stay within
    1e-3 of the exact weights' power on a small block, the two-part splits do
    not reach 1e-2, and each split part shrinks the error."""
    nant, nchan, npkt, nbeam, tavg = 64, 16, 64, 64, 16
    cfg = bf.config(nant=nant, nchan=nchan, npkt=npkt, nbeam=nbeam, tavg=tavg)
    antpos, cal, theta = bf.synth_geometry(nant, nchan, nbeam, 9, 1200.0, 0.1, 1.75)
    d = orc.dims(nant, nchan, npkt, nbeam, tavg)
    W = orc.weights(d, 1, 1, antpos, cal, bf.reference_freqs(1498.75, nchan), theta)
    block = bf.synth_block(cfg, 4)
    P = orc.power(d, 1, block, W)
    rel = {}
    for prec in range(4):
        Pq = orc.power(d, 1, block, orc.quant_weights(d, 1, prec, W))
        rel[prec] = (np.abs(Pq - P) / P).max()
    assert rel[0] < 1e-3 and rel[3] < 1e-3
    assert rel[1] < 1e-2 and rel[2] < 1e-2
    assert rel[3] < rel[2] / 4 and rel[3] < rel[0]

"""Pins for the CPU oracle (oracle/bf_oracle.c) -- no GPU needed.

Each oracle function is checked against something other than its own formula:
reference fixtures (utils/packet.out decode statistics), the reference's own
index arithmetic restated from a different kernel, an independent numpy
restatement, closed forms, invariants and hand-computed values.
"""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN


# ---- unpack (dsaX_beamformer.cu:190-191) ---------------------------------------

def test_unpack_exhaustive_matches_int8_shift(orc):
    # get_rms.py:33-36 decodes with int8 reinterpretation: ((d&15)<<4) as int8 / 16
    v = np.arange(256, dtype=np.uint8)
    re_ref = (((v & 15) << 4).astype(np.uint8).view(np.int8) // 16).astype(int)
    im_ref = ((v & 240).astype(np.uint8).view(np.int8) // 16).astype(int)
    got = np.array([orc.unpack_byte(int(x)) for x in v])
    assert (got[:, 0] == re_ref).all() and (got[:, 1] == im_ref).all()


def test_unpack_is_bijective_on_4bit_pairs(orc):
    seen = set()
    for re in range(-8, 8):
        for im in range(-8, 8):
            b = ((im & 15) << 4) | (re & 15)      # gen_packet.py:196-206 packing
            assert orc.unpack_byte(b) == (re, im)
            seen.add(b)
    assert len(seen) == 256


def test_packet_out_decode_statistics(orc):
    """utils/packet.out: the reference's test packet (3 ants x 384 ch x 2 t x 2 pol)."""
    d = np.fromfile(os.path.join(GOLDEN, "packet.out"), np.uint8)
    assert d.size == 4608
    pairs = np.array([orc.unpack_byte(int(x)) for x in d])
    hre = np.bincount(pairs[:, 0] + 8, minlength=16).tolist()
    him = np.bincount(pairs[:, 1] + 8, minlength=16).tolist()
    assert hre == [0, 0, 1, 20, 79, 311, 756, 1106, 1144, 761, 308, 107, 12, 3, 0, 0]
    assert him == [0, 0, 3, 11, 78, 293, 732, 1111, 1157, 767, 344, 89, 19, 4, 0, 0]
    assert np.sqrt((pairs[:, 0] ** 2).mean()) == pytest.approx(1.6105, abs=1e-4)
    assert np.sqrt((pairs[:, 1] ** 2).mean()) == pytest.approx(1.5904, abs=1e-4)


# ---- input indexing ----------------------------------------------------------------

def test_capture_index_matches_promoter_launch(orc):
    """promoter (dsaX_beamformer.cu:170-193) reads byte iidx = bidx*32 + tidx of an
    int with bidx = (tim*NANT + ant)*48 + chan and tidx = chunnel*2 + pol,
    chunnel = (channel%8)*2 + t.  Restate that launch and compare."""
    nant = 5
    d = orc.dims(nant, 384, 32, 32, 8)
    rng = np.random.default_rng(1)
    for _ in range(300):
        intn = int(rng.integers(0, 2))          # which 16-packet int
        tim, ant, chan = (int(rng.integers(0, 16)), int(rng.integers(0, nant)),
                          int(rng.integers(0, 48)))
        tidx = int(rng.integers(0, 32))
        bidx = (tim * nant + ant) * 48 + chan
        iidx = intn * 16 * nant * 1536 + bidx * 32 + tidx
        chunnel, pol = tidx // 2, tidx % 2
        c, t = chan * 8 + chunnel // 2, chunnel % 2
        assert orc.in_index(d, intn * 16 + tim, ant, c, t, pol) == iidx


def test_split_r_index_matches_split_memcpy(orc):
    """dsaX_split.c:94-95: out + i*1536*NANT + g*NANT*32 + j*32 <- in + i*1536*NANT + j*1536 + g*32."""
    nant = 7
    dc = orc.dims(nant, 384, 4, 32, 8, 0)
    dr = orc.dims(nant, 384, 4, 32, 8, 1)
    for p in range(4):
        for a in range(nant):
            for c in (0, 9, 100, 383):
                for t in (0, 1):
                    for q in (0, 1):
                        src = orc.in_index(dc, p, a, c, t, q)
                        k = src - (p * 1536 * nant + a * 1536 + (c // 8) * 32)
                        dst = p * 1536 * nant + (c // 8) * nant * 32 + a * 32 + k
                        assert orc.in_index(dr, p, a, c, t, q) == dst


# ---- weights (dsaX_beamformer.cu:420-449) ------------------------------------------

def _geom(bf, nant, ncal, nbeam, seed=3):
    return bf.synth_geometry(nant, ncal, nbeam, seed)


def test_weights_invariants(orc, bf):
    nant, nchan, nbeam = 16, 16, 9
    antpos, cal, theta = _geom(bf, nant, nchan, nbeam)
    theta[4] = 0.0
    freqs = bf.reference_freqs(1498.75, nchan)
    d = orc.dims(nant, nchan, 32, nbeam, 8)
    W = orc.weights(d, 1, 1, antpos, cal, freqs, theta)          # [c][q][b][a]
    cal = cal.astype(np.float64)
    calc = cal[..., 0] + 1j * cal[..., 1]                        # [a][c][q]
    # beam at angle 0: W = cal
    assert np.allclose(W[:, :, 4, :], calc.transpose(1, 2, 0), atol=1e-6)
    # |W| = |cal| everywhere
    assert np.allclose(np.abs(W), np.abs(calc.transpose(1, 2, 0))[:, :, None, :], rtol=1e-12)
    # phase is linear in x: W/cal for antenna pair sums multiplies
    ph = W / calc.transpose(1, 2, 0)[:, :, None, :]
    expect = np.exp(-2j * np.pi * freqs[:, None, None, None] * np.sin(theta)[None, None, :, None]
                    * antpos.astype(np.float64)[None, None, None, :] / 299792458.0)
    assert np.allclose(ph, np.broadcast_to(expect, ph.shape), atol=1e-9)


def test_weights_group_mode_uses_centre_channel(orc, bf):
    """Group mode phases use freqs[8g+4] for all 8 channels (dsaX_beamformer.cu:441)."""
    nant, nchan, nbeam = 8, 24, 4
    antpos, cal_g, theta = _geom(bf, nant, nchan // 8, nbeam)
    freqs = bf.reference_freqs(1530.0, nchan)
    d = orc.dims(nant, nchan, 32, nbeam, 8)
    Wg = orc.weights(d, 0, 0, antpos, cal_g, freqs, theta)
    cal_f = np.repeat(cal_g, 8, axis=1)
    f_centre = np.repeat(freqs[4::8], 8)
    Wf = orc.weights(d, 1, 0, antpos, cal_f, f_centre, theta)
    assert np.allclose(Wg, Wf[::8], atol=0)


def test_weights_match_reference_float_formula(orc):
    """calc_weights in float32 with the reference's own theta and PI constant
    (dsaX_beamformer.cu:436-446, dsaX_def.h:28) agrees with the FP64 oracle to
    float precision; a wrong sign or a dropped term would not."""
    nant, nbeam = 64, 256
    rng = np.random.default_rng(5)
    antpos = rng.uniform(-600, 600, nant).astype(np.float32)
    cal = rng.normal(size=(nant, 48, 2, 2)).astype(np.float32)
    freqs = ((1530.0 - np.arange(384) * 250.0 / 8192.0) * 1e6).astype(np.float32)
    theta = orc.ref_theta(nbeam, 1.0)
    d = orc.dims(nant, 384, 32, nbeam, 8)
    W = orc.weights(d, 0, 0, antpos, cal, freqs.astype(np.float64), theta)
    PI = 3.14159265359
    for f in (0, 17, 47):
        for pol in (0, 1):
            for beam in (0, 127, 200, 255):
                th = np.float32(1.0 * (127.0 - beam) * PI / 10800.0)
                afac = np.float32(-2.0 * PI * float(freqs[f * 8 + 4]) * float(th) / 299792458.0)
                ang = (afac * antpos).astype(np.float32)
                twr, twi = np.cos(ang).astype(np.float32), np.sin(ang).astype(np.float32)
                wr = twr * cal[:, f, pol, 0] - twi * cal[:, f, pol, 1]
                wi = twi * cal[:, f, pol, 0] + twr * cal[:, f, pol, 1]
                got = W[f, pol, beam]
                scale = np.abs(cal[:, f, pol, 0] + 1j * cal[:, f, pol, 1])
                assert np.abs(got - (wr + 1j * wi)).max() <= 2e-4 * max(1.0, scale.max())


def test_ref_theta(orc):
    t = orc.ref_theta(256, 1.0)
    assert t[127] == 0.0
    assert t[0] == pytest.approx(127 * math.pi / 10800)
    assert t[255] == pytest.approx(-128 * math.pi / 10800)


# ---- power (GEMM + detect + averaging) ------------------------------------------------

def _numpy_power(block, W, nant, nchan, npkt, nbeam, tavg, wmode, order=0):
    """Independent numpy restatement (einsum over complex128)."""
    if order == 0:
        raw = block.reshape(npkt, nant, nchan, 2, 2)
    else:
        raw = block.reshape(npkt, nchan // 8, nant, 8, 2, 2).transpose(0, 2, 1, 3, 4, 5) \
                   .reshape(npkt, nant, nchan, 2, 2)
    lo = (raw & 15).astype(np.int64)
    hi = (raw >> 4).astype(np.int64)
    D = (np.where(lo >= 8, lo - 16, lo) + 1j * np.where(hi >= 8, hi - 16, hi))
    D = D.transpose(2, 4, 0, 3, 1).reshape(nchan, 2, 2 * npkt, nant)   # c q n a
    Wc = W if wmode == 1 else np.repeat(W, 8, axis=0)
    B = np.einsum("cqba,cqna->cqnb", Wc, D)
    P = np.abs(B) ** 2
    ng, nout = nchan // 8, 2 * npkt // tavg
    P = P.reshape(ng, 8, 2, nout, tavg, nbeam).sum(axis=(1, 2, 4))  # g o b
    return P.transpose(1, 2, 0)


@pytest.mark.parametrize("wmode,amode,tavg,order,nant", [
    (1, 1, 8, 0, 64), (0, 0, 16, 0, 63), (1, 0, 32, 1, 30), (0, 1, 64, 1, 24), (1, 1, 16, 0, 256),
    (0, 0, 32, 1, 130)])
def test_power_matches_numpy(orc, bf, wmode, amode, tavg, order, nant):
    nchan, npkt, nbeam = 16, 64, 40
    cfg = bf.config(nant=nant, nchan=nchan, npkt=npkt, nbeam=nbeam, tavg=tavg, input_order=order)
    ncal = nchan // 8 if wmode == 0 else nchan
    antpos, cal, theta = _geom(bf, nant, ncal, nbeam)
    freqs = bf.reference_freqs(1450.0, nchan)
    block = bf.synth_block(cfg, seed=5)
    d = orc.dims(nant, nchan, npkt, nbeam, tavg, order)
    W = orc.weights(d, wmode, amode, antpos, cal, freqs, theta)
    got = orc.power(d, wmode, block, W)
    ref = _numpy_power(block, W, nant, nchan, npkt, nbeam, tavg, wmode, order)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-9)


def test_power_single_antenna_and_linearity(orc, bf):
    nant, nchan, npkt, nbeam, tavg = 16, 8, 32, 8, 8
    cfg = bf.config(nant=nant, nchan=nchan, npkt=npkt, nbeam=nbeam, tavg=tavg)
    antpos, cal, theta = _geom(bf, nant, nchan, nbeam)
    freqs = bf.reference_freqs(1450.0, nchan)
    d = orc.dims(nant, nchan, npkt, nbeam, tavg)
    W = orc.weights(d, 1, 1, antpos, cal, freqs, theta)
    block = bf.synth_block(cfg, seed=9, sigma=1.0, clip=3).reshape(npkt, nant, nchan, 2, 2)
    one = np.zeros_like(block)
    one[:, 5] = block[:, 5]                              # only antenna 5 live
    P = orc.power(d, 1, one.ravel(), W)
    raw = one[:, 5].astype(np.int64)
    lo, hi = raw & 15, raw >> 4
    e = (np.where(lo >= 8, lo - 16, lo) ** 2 + np.where(hi >= 8, hi - 16, hi) ** 2)  # p c t q
    w2 = np.abs(W[:, :, :, 5]) ** 2                                                    # c q b
    expect = np.einsum("pctq,cqb->pb", e, w2).reshape(npkt * 2 // tavg, tavg // 2, nbeam).sum(1)
    assert np.allclose(P[:, :, 0], expect, rtol=1e-12)
    # doubling every sample (values stay in 4 bits) quadruples the power
    lo, hi = block & 15, block >> 4
    re, im = np.where(lo >= 8, lo - 16, lo), np.where(hi >= 8, hi - 16, hi)
    dbl = (((2 * im) & 15) << 4 | ((2 * re) & 15)).astype(np.uint8)
    P1 = orc.power(d, 1, block.ravel(), W)
    P2 = orc.power(d, 1, dbl.ravel(), W)
    assert np.allclose(P2, 4 * P1, rtol=1e-12)
    # zero input, zero power
    assert (orc.power(d, 1, np.zeros_like(block).ravel(), W) == 0).all()


def test_beam_power_sums_outputs_and_groups(orc, bf):
    """orc_beam_power (badder + calc_allbp, dsaX_beamformer.cu:78-112,482-495)
    on a shape where nout (8), nbeam (5) and ngroup (3) all differ: the sum of
    the power over outputs and groups, and for one live antenna the closed
    form sum_{t,c,q} |W[c,q,b,a]|^2 |d[t,c,q]|^2."""
    nant, nchan, npkt, nbeam, tavg = 12, 24, 64, 5, 16
    cfg = bf.config(nant=nant, nchan=nchan, npkt=npkt, nbeam=nbeam, tavg=tavg)
    antpos, cal, theta = _geom(bf, nant, nchan, nbeam)
    d = orc.dims(nant, nchan, npkt, nbeam, tavg)
    W = orc.weights(d, 1, 1, antpos, cal, bf.reference_freqs(1450.0, nchan), theta)
    block = bf.synth_block(cfg, seed=5).reshape(npkt, nant, nchan, 2, 2)
    P = orc.power(d, 1, block.ravel(), W)
    assert P.shape == (8, nbeam, 3)
    assert np.allclose(orc.beam_power(d, P), P.sum(axis=(0, 2)), rtol=1e-13)
    one = np.zeros_like(block)
    one[:, 7] = block[:, 7]
    P1 = orc.power(d, 1, one.ravel(), W)
    raw = one[:, 7].astype(np.int64)
    lo, hi = raw & 15, raw >> 4
    e2 = np.where(lo >= 8, lo - 16, lo) ** 2 + np.where(hi >= 8, hi - 16, hi) ** 2   # p c t q
    w2 = np.abs(W[:, :, :, 7]) ** 2                                                  # c q b
    expect = np.einsum("pctq,cqb->b", e2, w2)
    assert np.allclose(orc.beam_power(d, P1), expect, rtol=1e-12)
    # constant voltage (re = 1, im = 0 on every antenna): |B|^2 = |sum_a W|^2
    ones = np.full_like(block, 0x01)
    Pc = orc.power(d, 1, ones.ravel(), W)
    expect_c = 2 * npkt * np.einsum("cqb->b", np.abs(W.sum(axis=3)) ** 2)
    assert np.allclose(orc.beam_power(d, Pc), expect_c, rtol=1e-12)


def test_dbnic_transpose(orc):
    """orc_dbnic_op (dsaX_dbnic.c:111-116) at the corner turn's own sizes
    (dsaX_def.h:87-94: NSAMPS_PER_TRANSMIT 512, NBMS 256, NBEAMS_PER_BLOCK 64,
    NW 48): on a block of distinct values, slice s is beams 64s..64s+63 moved
    to the front, i.e. the beam-major block cut into 64-beam slices; each
    slice is 384 packets of P_SIZE - 12 = 4096 data bytes."""
    rng = np.random.default_rng(0)
    out = rng.integers(0, 256, size=(512, 256, 48), dtype=np.uint8)
    beam_major = np.ascontiguousarray(out.transpose(1, 0, 2))
    for s in range(4):
        op = orc.dbnic_op(out, s)
        assert op.shape == (64, 512, 48) and op.nbytes == 384 * 4096
        assert np.array_equal(op, beam_major[64 * s: 64 * s + 64])
        # spot values: op[j][i][k] = out[i][64 s + j][k]
        for j, i, k in ((0, 0, 0), (63, 511, 47), (5, 100, 7)):
            assert op[j, i, k] == out[i, 64 * s + j, k]
    with pytest.raises(ValueError):
        orc.dbnic_op(out, 4)


def test_power_steered_tone_peaks_at_its_beam(orc, bf):
    nant, nchan, npkt, nbeam = 32, 8, 64, 33
    cfg = bf.config(nant=nant, nchan=nchan, npkt=npkt, nbeam=nbeam, tavg=32)
    antpos, cal, theta = bf.synth_geometry(nant, nchan, nbeam, 4, 1200.0, 0.1, 1.75)
    freqs = bf.reference_freqs(1498.75, nchan)
    block = bf.synth_block(cfg, 12, 2.0, 7, 10.0, theta[21], tuple(range(8)), antpos,
                           cal[:, 0, 0, :], freqs)
    d = orc.dims(nant, nchan, npkt, nbeam, 32)
    W = orc.weights(d, 1, 1, antpos, cal, freqs, theta)
    P = orc.power(d, 1, block, W)
    assert P.sum(axis=(0, 2)).argmax() == 21


def test_power_stuff_modes(orc, bf):
    nant, nchan, npkt, nbeam, tavg = 20, 8, 32, 70, 16
    cfg = bf.config(nant=nant, nchan=nchan, npkt=npkt, nbeam=nbeam, tavg=tavg)
    antpos, cal, theta = _geom(bf, nant, nchan, nbeam)
    freqs = bf.reference_freqs(1450.0, nchan)
    d = orc.dims(nant, nchan, npkt, nbeam, tavg)
    W = orc.weights(d, 1, 0, antpos, cal, freqs, theta)
    block = bf.synth_block(cfg, seed=21)
    raw = block.reshape(npkt, nant, nchan, 2, 2).astype(np.int64)
    lo, hi = raw & 15, raw >> 4
    e = (np.where(lo >= 8, lo - 16, lo) ** 2 + np.where(hi >= 8, hi - 16, hi) ** 2)   # p a c t q
    per_ant = e.sum(axis=(2, 3, 4)).reshape(npkt * 2 // tavg, tavg // 2, nant).sum(1)  # o a
    P1 = orc.power(d, 1, block, W, stuff=1)
    for b in range(nbeam):
        ant = ((b // 16) % 4) * 16 + b % 16
        want = per_ant[:, ant] if ant < nant else 0.0
        assert np.allclose(P1[:, b, 0], want)
    P2 = orc.power(d, 1, block, W, stuff=2)
    P0 = orc.power(d, 1, block, W, stuff=0)
    assert np.allclose(P2[:, 0, 0], per_ant.sum(axis=1))
    assert np.allclose(P2[:, 1:], P0[:, 1:])


# ---- requantisation (dsaX_beamformer.cu:150-156) ---------------------------------------

@pytest.mark.parametrize("v,want", [
    (0.0, 0), (0.5, 0), (1.5, 1), (2.5, 1), (3.5, 2), (4.5, 2), (320.0, 160), (511.0, 255),
    (512.0, 0), (513.0, 0), (514.0, 1), (1023.0, 255), (1024.0, 0), (-3.0, 255), (1e10, 255)])
def test_requant_hand_values(orc, bf, v, want):
    # rint half-even, C truncating /2, wrap modulo 256; 1e10 saturates to INT_MAX
    assert orc.requant(v, 1.0) == want
    assert bf.requant(v, 1.0) == want


def test_requant_block_matches_scalar(orc):
    rng = np.random.default_rng(2)
    pw = rng.uniform(0, 2000, size=(3, 5, 2))
    sc = rng.uniform(0.1, 0.5, size=5).astype(np.float32)
    blk = orc.requant_block(pw, sc)
    for idx in np.ndindex(pw.shape):
        assert blk[idx] == orc.requant(float(np.float32(pw[idx])), float(sc[idx[1]]))


# ---- AGC (dsaX_beamformer.cu:1005-1066) ---------------------------------------------------

def test_agc_hand_sequence(orc, bf):
    nb, ng, nints = 256, 48, 128
    o, p = orc.Agc(nb, ng, nints), bf.Agc(nb, ng, nints)
    base = np.arange(1, nb + 1, dtype=np.float32) * 1000.0
    # block 0: median (0.5*(128000+129000)), output zeroed
    s_o, z_o = o.update(base)
    s_p, z_p = p.update(base)
    med = 0.5 * (128000.0 + 129000.0)
    want = np.float32(320.0 / (np.float32(med) / (48.0 * 128)))
    assert z_o and z_p
    assert np.allclose(s_o, want) and np.array_equal(s_o, s_p)
    # blocks 1..7: median of each block
    for k in range(1, 8):
        bp = base * (1 + 0.01 * k)
        s_o, z_o = o.update(bp)
        s_p, z_p = p.update(bp)
        assert not z_o and not z_p and np.array_equal(s_o, s_p)
        m = np.float32(0.5 * (np.float32(bp[127]) + np.float32(bp[128])))
        assert np.allclose(s_o, 320.0 / (m / (48.0 * 128)), rtol=1e-6)
    # block 8: mean of the last 8 per beam, clamped to medbp beyond 10 %
    bp = np.full(nb, med, np.float32)
    s_o, _ = o.update(bp)
    s_p, _ = p.update(bp)
    assert np.array_equal(s_o, s_p)
    hist = np.stack([base * (1 + 0.01 * k) for k in range(1, 8)] + [bp]).astype(np.float32)
    mean = hist.mean(axis=0)
    mean = np.where(np.abs(mean - med) / med > 0.1, med, mean)
    assert np.allclose(s_o, 320.0 / (mean / (48.0 * 128)), rtol=1e-5)
    assert p.state["started"] == 2 and p.state["bpctr"] == 9


def test_agc_zero_power_and_disabled(orc, bf):
    o = orc.Agc(256, 48, 128)
    s, z = o.update(np.zeros(256, np.float32))
    assert z and (s == 0).all()
    p = bf.Agc(256, 48, 128, enabled=False)
    base = np.linspace(1, 2, 256, dtype=np.float32)
    s0, _ = p.update(base)
    for k in range(12):
        s, _ = p.update(base * (3 + k))       # held at the first block's median
        assert np.array_equal(s, s0)


def test_power_group_range_is_clamped(orc, bf):
    """A group range past the block's last group computes only the groups that
    exist (the C restatement neither reads nor writes past the block)."""
    nant, nchan, npkt, nbeam, tavg = 16, 8, 32, 8, 8
    cfg = bf.config(nant=nant, nchan=nchan, npkt=npkt, nbeam=nbeam, tavg=tavg)
    antpos, cal, theta = _geom(bf, nant, nchan, nbeam)
    d = orc.dims(nant, nchan, npkt, nbeam, tavg)
    W = orc.weights(d, 1, 1, antpos, cal, bf.reference_freqs(1450.0, nchan), theta)
    block = bf.synth_block(cfg, seed=3)
    full = orc.power(d, 1, block, W)
    assert full.shape == (2 * npkt // tavg, nbeam, 1)
    assert np.array_equal(orc.power(d, 1, block, W, 0, 0, 3), full)
    assert np.array_equal(orc.power(d, 1, block, W, 0, -2, 1), full)

"""The correlator X-engine on the GPU against the exact oracle
(oracle/xc_oracle.c, dcorrelator of dsaX_bfCorr.cu:472-569).  Visibilities
are integer sums of 4-bit products, so the int8 matrix-core path (biased
unsigned operands, int32 sums, exact corrections) must match the oracle bit
for bit: every
baseline, channel and pol, both input orders, 1 to 96 antennas (the 32-antenna
tile edges), the full 384-channel x 4096-sample block with 96 antennas on a
channel subset, and the reference's packet.out fixture."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _run(bf, nant, nchan, npkt, block, order=0):
    x = bf.XEngine(nant, nchan, npkt, order)
    vis = x.process_block(block)
    x.close()
    return vis


@pytest.mark.parametrize("nant,nchan,npkt,order", [
    (1, 8, 32, 0), (24, 48, 64, 0), (30, 16, 32, 1), (31, 8, 64, 0), (32, 16, 32, 0),
    (33, 8, 32, 1), (63, 24, 64, 0), (64, 16, 96, 1), (65, 8, 32, 0), (96, 48, 64, 0),
    (96, 16, 128, 1)])
def test_xengine_matches_oracle_exactly(bf, orc, nant, nchan, npkt, order):
    cfg = bf.config(nant=nant, nchan=nchan, npkt=npkt, nbeam=8, input_order=order)
    block = bf.synth_block(cfg, nant + nchan + npkt)
    vis = _run(bf, nant, nchan, npkt, block, order)
    ref = orc.xcorr(nant, nchan, npkt, block, order)
    assert vis.shape == ref.shape
    assert np.array_equal(vis, ref)


@pytest.mark.parametrize("order", [0, 1])
def test_xengine_every_nibble_value_and_extremes(bf, orc, order):
    """Uniformly random bytes, so every 4-bit value including -8 (the unsigned
    operand 0) occurs in both components, against the oracle; and a block of
    -8 - 8j everywhere, the largest products: Re V = 128 * nsamp, Im V = 0 for
    every baseline, channel and pol."""
    nant, nchan, npkt = 96, 16, 256
    rng = np.random.default_rng(31 + order)
    block = rng.integers(0, 256, npkt * nant * nchan * 4, dtype=np.uint8)
    assert np.array_equal(_run(bf, nant, nchan, npkt, block, order), orc.xcorr(nant, nchan, npkt, block, order))
    nant, nchan, npkt = 40, 8, 512
    block = np.full(npkt * nant * nchan * 4, 0x88, np.uint8)
    vis = _run(bf, nant, nchan, npkt, block, order)
    assert np.array_equal(vis, orc.xcorr(nant, nchan, npkt, block, order))
    assert (vis.real == 128 * 2 * npkt).all() and (vis.imag == 0).all()


def test_xengine_full_block_96_antennas(bf, orc):
    """384 channels x 2048 packets x 96 antennas (the reference's NANTS):
    exact on channel group 7 (cut out of the block) and, over the whole block,
    autocorrelations real and equal to the summed power of each antenna."""
    nant, nchan, npkt = 96, 384, 2048
    cfg = bf.config(nant=nant, nchan=nchan, npkt=npkt, nbeam=8)
    block = bf.synth_block(cfg, 77)
    vis = _run(bf, nant, nchan, npkt, block)
    g = 7
    sub = np.ascontiguousarray(block.reshape(npkt, nant, nchan, 4)[:, :, 8 * g: 8 * g + 8]).ravel()
    assert np.array_equal(vis[:, 8 * g: 8 * g + 8], orc.xcorr(nant, 8, npkt, sub))
    b = block.reshape(npkt, nant, nchan, 2, 2)
    lo, hi = (b & 15).astype(np.int64), (b >> 4).astype(np.int64)
    p = np.where(lo >= 8, lo - 16, lo) ** 2 + np.where(hi >= 8, hi - 16, hi) ** 2
    power = p.sum(axis=(0, 3))                                     # [ant][chan][pol]
    for a in range(0, nant, 7):
        auto = vis[orc.baseline(a, a)]
        assert (auto.imag == 0).all() and np.array_equal(auto.real, power[a])


def test_xengine_packet_fixture_and_device_execute(bf, orc):
    """The reference's utils/packet.out tiled over a 63-antenna block (as
    dsaX_fake builds the test_bf.py block), and the device-pointer entry point
    giving the same bytes as the host one."""
    pkt = np.fromfile(os.path.join(GOLDEN, "packet.out"), np.uint8)
    nant, nchan, npkt = 63, 384, 64
    block = np.tile(pkt, npkt * nant // 3)
    vis = _run(bf, nant, nchan, npkt, block)
    assert np.array_equal(vis, orc.xcorr(nant, nchan, npkt, block))
    x = bf.XEngine(nant, nchan, npkt)
    d_in, d_vis = bf.DeviceBuffer(block.nbytes), bf.DeviceBuffer(x.vis_bytes)
    d_in.upload(block)
    x.execute(d_in, d_vis)
    out = d_vis.download(np.empty((x.nbase, nchan, 2, 2), np.float32), x.stream)
    assert np.array_equal(out[..., 0] + 1j * out[..., 1].astype(np.float64), vis)
    x.close()


@pytest.mark.parametrize("nslice", [1, 2, 3, 5, 8])
def test_xengine_time_slices_give_the_same_visibilities(bf, orc, nslice):
    """The correlation split into time slices whose partial visibilities are
    summed (uneven slices too): identical to the oracle."""
    nant, nchan, npkt = 96, 48, 256
    cfg = bf.config(nant=nant, nchan=nchan, npkt=npkt, nbeam=8)
    block = bf.synth_block(cfg, 500 + nslice)
    x = bf.XEngine(nant, nchan, npkt)
    x.slices = nslice
    assert x.slices == nslice
    vis = x.process_block(block)
    x.close()
    assert np.array_equal(vis, orc.xcorr(nant, nchan, npkt, block))


def test_xengine_rejects_bad_shapes(bf):
    for args in ((97, 8, 32), (0, 8, 32), (32, 12, 32), (32, 8, 48)):
        with pytest.raises(bf.BeamformerError):
            bf.XEngine(*args)
    x = bf.XEngine(32, 8, 64)
    for n in (-1, 3):                        # 64 packets are 2 units of 32
        with pytest.raises(bf.BeamformerError):
            x.slices = n
    x.close()

"""Pins of the correlator oracle (oracle/xc_oracle.c, dsaX_bfCorr.cu
dcorrelator), CPU only: brute force in numpy on small blocks in both input
orders, the reference's baseline numbering loop at NANTS = 96 (NBASE 4656),
the Hermitian symmetry under antenna reversal, real non-negative
autocorrelations equal to the summed power, linearity, and the reference's own
packet.out fixture tiled into a block as dsaX_fake does."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _voltages(block, npkt, nant, nchan, order=0):
    """Complex voltages [packet][time][antenna][channel][pol] of a block."""
    b = np.asarray(block, np.uint8)
    lo, hi = (b & 15).astype(np.int64), (b >> 4).astype(np.int64)
    v = np.where(lo >= 8, lo - 16, lo) + 1j * np.where(hi >= 8, hi - 16, hi)
    if order == 0:
        v = v.reshape(npkt, nant, nchan, 2, 2)                  # p a c t q
    else:
        v = v.reshape(npkt, nchan // 8, nant, 8, 2, 2)          # p g a c8 t q
        v = v.transpose(0, 2, 1, 3, 4, 5).reshape(npkt, nant, nchan, 2, 2)
    return v.transpose(0, 3, 1, 2, 4)                           # p t a c q


@pytest.mark.parametrize("order", [0, 1])
def test_xcorr_matches_numpy_brute_force(orc, bf, order):
    nant, nchan, npkt = 7, 16, 32
    cfg = bf.config(nant=nant, nchan=nchan, npkt=npkt, nbeam=8, input_order=order)
    block = bf.synth_block(cfg, 5)
    V = orc.xcorr(nant, nchan, npkt, block, order)
    d = _voltages(block, npkt, nant, nchan, order)
    full = np.einsum("ptjcq,pticq->ijcq", np.conj(d), d)      # [i][j] = sum conj(d_j) d_i
    for i in range(nant):
        for j in range(i + 1):
            assert np.array_equal(V[orc.baseline(i, j)], full[i, j]), (i, j)
    assert V.shape == (nant * (nant + 1) // 2, nchan, 2)


def test_baseline_numbering_is_the_reference_loop(orc):
    # reorder_output's h_idxs: for i in 0..NANTS-1, for j in 0..i: ii++
    ii = 0
    for i in range(96):
        for j in range(i + 1):
            assert orc.baseline(i, j) == ii
            ii += 1
    assert ii == 4656                        # NBASE, dsaX_def.h:40


def test_xcorr_hermitian_autos_and_linearity(orc, bf):
    nant, nchan, npkt = 9, 8, 64
    cfg = bf.config(nant=nant, nchan=nchan, npkt=npkt, nbeam=8)
    block = bf.synth_block(cfg, 11)
    V = orc.xcorr(nant, nchan, npkt, block)
    # reversing the antenna order conjugates every baseline: (i, j) -> (n-1-j, n-1-i)
    rev = np.ascontiguousarray(block.reshape(npkt, nant, nchan * 4)[:, ::-1]).ravel()
    Vr = orc.xcorr(nant, nchan, npkt, rev)
    for i in range(nant):
        for j in range(i + 1):
            assert np.array_equal(Vr[orc.baseline(nant - 1 - j, nant - 1 - i)], np.conj(V[orc.baseline(i, j)]))
    # autocorrelations: real, the summed |d|^2 of the antenna
    d = _voltages(block, npkt, nant, nchan)
    for a in range(nant):
        auto = V[orc.baseline(a, a)]
        assert (auto.imag == 0).all()
        assert np.array_equal(auto.real, (d[:, :, a].real ** 2 + d[:, :, a].imag ** 2).sum(axis=(0, 1)))
    # doubling every voltage (values stay in 4 bits) quadruples every visibility
    small = bf.synth_block(cfg, 3, 1.0, 3)
    lo, hi = small & 15, small >> 4
    re, im = np.where(lo >= 8, lo.astype(int) - 16, lo), np.where(hi >= 8, hi.astype(int) - 16, hi)
    dbl = ((((2 * im) & 15) << 4) | ((2 * re) & 15)).astype(np.uint8)
    assert np.array_equal(orc.xcorr(nant, nchan, npkt, dbl), 4 * orc.xcorr(nant, nchan, npkt, small))


def test_xcorr_on_the_reference_packet_fixture(orc):
    """utils/packet.out (3 antennas x 384 channels x 2 times x 2 pols) tiled
    over packets and antennas as dsaX_fake does: antenna a carries packet
    antenna a % 3, so every visibility is npkt times the packet's own
    correlation of antennas (i % 3, j % 3)."""
    pkt = np.fromfile(os.path.join(GOLDEN, "packet.out"), np.uint8)
    nant, nchan, npkt = 6, 384, 8
    block = np.tile(pkt, npkt * nant // 3)
    V = orc.xcorr(nant, nchan, npkt, block)
    x = _voltages(pkt, 1, 3, 384)[0]                      # t a c q
    for i in range(nant):
        for j in range(i + 1):
            want = npkt * (np.conj(x[:, j % 3]) * x[:, i % 3]).sum(axis=0)
            assert np.array_equal(V[orc.baseline(i, j)], want), (i, j)
    # the fixture's power: its re and im nibbles have rms 1.611 and 1.590
    auto = sum(V[orc.baseline(a, a)].real.sum() for a in range(3)) / npkt
    assert auto / (3 * 384 * 2 * 2) == pytest.approx(1.611 ** 2 + 1.590 ** 2, abs=0.01)

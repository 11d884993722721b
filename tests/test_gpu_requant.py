"""The device requantisation, exactly: (uint8)(rint(x*s)/2) of
dsaX_beamformer.cu:150-156 (round half to even, saturating conversion to int,
C division truncating toward zero, wrap modulo 256) as the fused kernel and the
-s kernel apply it.

Every antenna carries the constant voltage 1+0i and every beam points at
angle 0 with unit calibration, so every weight is exactly 1 in each precision
that can hold it (binary16, E4M3) and every output's power is exactly
8 channels x 2 pols x 8 samples = 128.  Each beam gets its own scale
t/128, so x*s lands exactly on a chosen value t: half-way ties, 511/512/513
around the wrap, negatives, huge values that saturate and NaN.  The bytes
must equal the oracle's requantisation (pinned against hand values in
tests/test_oracle.py) and the hand-written expectations below."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (x*s, expected byte)
HAND = [(0.5, 0), (1.5, 1), (2.5, 1), (3.5, 2), (4.5, 2), (510.5, 255), (511.0, 255),
        (511.5, 0), (512.0, 0), (513.0, 0), (514.0, 1), (1023.0, 255), (-0.5, 0), (-1.5, 255),
        (-2.5, 255), (-3.0, 255), (-4.0, 254), (1e10, 255), (-1e10, 0), (3e9, 255),
        (float("nan"), 0), (float("inf"), 255), (0.0, 0), (257.0, 128)]


def _targets(nbeam):
    rng = np.random.default_rng(5)
    t = [x for x, _ in HAND]
    t += list(np.round(rng.uniform(-600, 600, nbeam - len(t)) * 2) / 2)   # many more ties
    return np.array(t[:nbeam], np.float32)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("stuff", [0, 1])
@pytest.mark.parametrize("variant", ["default", "classic256"])
def test_requant_exact_on_device(bf, orc, prec, stuff, variant):
    nant = 64 if stuff else 1
    nbeam, tavg = 256, 8
    cfg = bf.config(nant=nant, nchan=16, npkt=64, nbeam=nbeam, tavg=tavg, precision=prec,
                    weight_mode=bf.WEIGHTS_GROUP, angle_mode=bf.ANGLE_THETA, stuffants=stuff)
    block = np.full(bf.input_bytes(cfg), 0x01, np.uint8)     # re = 1, im = 0 everywhere
    cal = np.zeros((nant, cfg.ncal, 2, 2), np.float32)
    cal[..., 0] = 1.0
    plan = bf.Plan(cfg)
    plan.set_weights(np.zeros(nant, np.float32), cal, bf.reference_freqs(1450.0, 16),
                     np.zeros(nbeam))
    if variant != "default":
        plan.variant = bf.VARIANT_CLASSIC_256
    t = _targets(nbeam)
    scale = (t / np.float32(128.0)).astype(np.float32)
    out, power, _ = plan.process_block(block, scale, want_power=True)
    plan.close()
    assert (power == 128.0).all()
    assert np.array_equal(out, orc.requant_block(power.astype(np.float64), scale))
    for b, (_, want) in enumerate(HAND):
        assert (out[:, b] == want).all(), (b, HAND[b], out[0, b])

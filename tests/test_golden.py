"""The committed golden fixtures (tests/golden/make_golden.py) against the
oracle, an independent numpy restatement, the FP32 CPU baseline and the host
AGC.  The GPU side of the small block is in test_gpu_parity.py."""
import hashlib
import os

import numpy as np

from conftest import GOLDEN
from test_oracle import _numpy_power


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def _small(orc):
    import sys
    sys.path.insert(0, GOLDEN)
    import make_golden
    return make_golden.SMALL, make_golden.small_inputs()


def test_small_block_regenerates_and_oracle_matches(orc):
    g = _load("small_block.npz")
    p, (cfg, block, antpos, cal, freqs, theta, d, W) = _small(orc)
    assert hashlib.sha256(block.tobytes()).hexdigest() == str(g["block_sha256"])
    assert np.array_equal(antpos, g["antpos"]) and np.array_equal(cal, g["cal"])
    assert np.array_equal(theta, g["theta"]) and np.array_equal(freqs, g["freqs"])
    power = orc.power(d, p["wmode"], block, W)
    np.testing.assert_allclose(power, g["power"], rtol=1e-12)
    assert np.array_equal(orc.requant_block(power, g["scale"]), g["out"])
    assert g["out"].shape == (32, 32, 6)
    # the fixed scale keeps the bytes mid-range: nothing saturates or wraps
    assert 32 < g["out"].mean() < 128 and int(g["out"].max()) < 255


def test_small_block_numpy_restatement(orc):
    g = _load("small_block.npz")
    p, (cfg, block, antpos, cal, freqs, theta, d, W) = _small(orc)
    ref = _numpy_power(block, W, p["nant"], p["nchan"], p["npkt"], p["nbeam"], p["tavg"], p["wmode"])
    np.testing.assert_allclose(ref, g["power"], rtol=1e-10)


def test_small_block_fp32_cpu_baseline(orc):
    g = _load("small_block.npz")
    p, (cfg, block, antpos, cal, freqs, theta, d, W) = _small(orc)
    _, pw, out = orc.cpu_beamform(d, p["wmode"], block, W, g["scale"], nthreads=4)
    rel = np.abs(pw - g["power"]) / g["power"]
    assert rel.max() < 1e-5
    diff = np.abs(out.astype(int) - g["out"].astype(int))
    assert diff.max() <= 1 and (diff > 0).mean() < 1e-3


def test_agc_trace_host_and_oracle(orc, bf):
    g = _load("agc_trace.npz")
    nb, ng, ni = int(g["nbeam"]), int(g["ngroup"]), int(g["nints"])
    host, ref = bf.Agc(nb, ng, ni), orc.Agc(nb, ng, ni)
    for k, bp in enumerate(g["bp_raw"]):
        s_host, z_host = host.update(bp)
        s_ref, z_ref = ref.update(bp)
        assert np.array_equal(s_host, g["scale"][k]), k
        assert np.array_equal(s_ref, g["scale"][k]), k
        assert z_host == z_ref == bool(g["zero"][k]), k
    # block 0 zeroes the output and the first 8 blocks use one median scale
    assert g["zero"][0] and not g["zero"][1:].any()
    assert all(np.unique(g["scale"][k]).size == 1 for k in range(8))
    # the clamp: beams whose 8-block mean leaves the +-10 % band around the
    # frozen median get the frozen median's scale (block 0's); the burst in
    # blocks 10-11 pushes more beams out of the band than block 9 had
    frozen = g["scale"][0][0]
    n_frozen = [(g["scale"][k] == frozen).sum() for k in range(8, 16)]
    assert n_frozen[2] > n_frozen[1] and n_frozen[3] > n_frozen[2]
    assert 0 < n_frozen[1] < nb

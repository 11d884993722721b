"""Host-side tests of libdsaxbf (no GPU): the C ABI exports, configuration
checks, the weights-file reader (init_weights semantics), the seeded
generator and ANTPOL_RMS.  GPU entry points are only checked for being
exported and for failing cleanly without a device."""
import ctypes as C
import os
import re
import struct
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

HEADER = os.path.join(ROOT, "include", "dsaxbf.h")


def _declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dsaxbf_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported(bf):
    names = _declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(bf.lib, n), n
    assert sorted(bf.EXPORTED) == names


def test_library_is_gfx950_build(bf, tmp_path):
    assert bf.lib.dsaxbf_version() == b"dsaxbf 1.0 gfx950"
    # llvm-objdump --offloading writes the extracted code objects next to its
    # input, so it reads a copy in a scratch folder
    so = tmp_path / "lib.so"
    shutil.copy(bf.LIB_PATH, so)
    out = subprocess.run(["/opt/rocm/llvm/bin/llvm-objdump", "--offloading", str(so)],
                         capture_output=True, text=True, cwd=tmp_path)
    if out.returncode == 0 and out.stdout:
        assert "gfx950" in out.stdout


def test_config_defaults_match_reference(bf):
    c = bf.Config()
    bf.lib.dsaxbf_config_default(C.byref(c), 63)
    assert (c.nant, c.nchan, c.npkt, c.nbeam, c.tavg) == (63, 384, 2048, 256, 8)
    assert (c.precision, c.weight_mode, c.angle_mode, c.stuffants, c.input_order) == (0, 0, 0, 0, 0)
    # block sizes: 2048 x 63 x 1536 in (test_bf.py:29), 512 x 256 x 48 out (:31)
    assert bf.input_bytes(c) == 198180864
    assert bf.output_bytes(c) == 6291456
    c.nant = 30
    assert bf.input_bytes(c) == 94371840          # config_corr.yaml ddda buffer
    c.tavg = 32
    assert bf.output_bytes(c) == 1572864          # README 128 x 256 x 48


@pytest.mark.parametrize("field,value", [
    ("nant", 0), ("nant", 257), ("nchan", 12), ("npkt", 48), ("nbeam", 0), ("nbeam", 1025),
    ("tavg", 4), ("tavg", 24), ("precision", 4), ("precision", -1), ("weight_mode", 3),
    ("angle_mode", -1), ("stuffants", 3), ("input_order", 2), ("nant", 129)])
def test_create_rejects_bad_config(bf, field, value):
    # nant 129 is rejected for the three-part FP6 split only (up to 128 antennas)
    cfg = bf.config(nant=16, nchan=16, npkt=64, nbeam=32,
                    precision=bf.PREC_FP6SPLIT3 if (field, value) == ("nant", 129) else bf.PREC_FP16)
    setattr(cfg, field, value)
    with pytest.raises(bf.BeamformerError) as e:
        bf.Plan(cfg)
    assert e.value.code == -1


def test_create_without_gpu_fails_cleanly(bf):
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(bf.BeamformerError) as e:
        bf.Plan(bf.config(nant=16, nchan=16, npkt=64, nbeam=32))
    assert e.value.code in (-6, -3)


def test_reference_freqs(bf):
    f = bf.reference_freqs(1530.0, 384)
    assert f[0] == 1530e6 and f[383] == pytest.approx((1530.0 - 383 * 250 / 8192) * 1e6)


# ---- weights file (init_weights, dsaX_beamformer.cu:559-602) ----------------------

def _write_weights(path, antpos, cal):
    with open(path, "wb") as fh:
        fh.write(np.asarray(antpos, np.float32).tobytes())
        fh.write(np.asarray(cal, np.float32).tobytes())


def test_weights_file_normalisation_and_flags(bf, tmp_path):
    rng = np.random.default_rng(4)
    antpos = rng.uniform(-500, 500, 64).astype(np.float32)
    cal = rng.normal(size=(64, 48, 2, 2)).astype(np.float32)
    cal[3, 5, 1] = 0.0                                   # zero stays zero
    fn = tmp_path / "w.bin"
    _write_weights(fn, antpos, cal)
    a, c = bf.read_weights_file(str(fn), os.path.join(GOLDEN, "flagants.dat"))
    assert np.array_equal(a, antpos)
    w = cal[..., 0] + 1j * cal[..., 1]
    with np.errstate(divide="ignore", invalid="ignore"):
        expect = np.where(np.abs(w) > 0, w / np.abs(w) ** 2, 0)   # w / |w|^2
    flags = [int(x) for x in open(os.path.join(GOLDEN, "flagants.dat")).read().split()]
    assert len(flags) == 18 and 63 in flags
    expect[flags] = 0
    got = c[..., 0] + 1j * c[..., 1]
    assert np.allclose(got, expect, rtol=2e-6, atol=1e-7)
    assert (c[flags] == 0).all()
    # unflagged antennas keep |w'| = 1/|w|
    keep = [i for i in range(64) if i not in flags]
    assert np.allclose(np.abs(got[keep]) * np.abs(w[keep]), np.where(np.abs(w[keep]) > 0, 1, 0),
                       rtol=1e-5)


def test_weights_file_errors(bf, tmp_path):
    with pytest.raises(bf.BeamformerError) as e:
        bf.read_weights_file(str(tmp_path / "missing.bin"))
    assert e.value.code == -4
    short = tmp_path / "short.bin"
    _write_weights(short, np.zeros(64), np.zeros(100))
    with pytest.raises(bf.BeamformerError) as e:
        bf.read_weights_file(str(short))
    assert e.value.code == -4
    good = tmp_path / "good.bin"
    _write_weights(good, np.zeros(64), np.ones((64, 48, 2, 2)))
    bad_flags = tmp_path / "flags.txt"
    bad_flags.write_text("3\n64\n")
    with pytest.raises(bf.BeamformerError) as e:
        bf.read_weights_file(str(good), str(bad_flags))
    assert e.value.code == -1
    junk = tmp_path / "junk.txt"
    junk.write_text("x\n")
    with pytest.raises(bf.BeamformerError) as e:
        bf.read_weights_file(str(good), str(junk))
    assert e.value.code == -4
    empty = tmp_path / "empty.txt"
    empty.write_text("")
    a, c = bf.read_weights_file(str(good), str(empty))      # no flags: nothing zeroed
    assert (c != 0).all()


# ---- seeded generator ----------------------------------------------------------------

def test_synth_block_deterministic_and_distributed(bf, orc):
    cfg = bf.config(nant=8, nchan=16, npkt=64, nbeam=8)
    b1 = bf.synth_block(cfg, seed=3)
    b2 = bf.synth_block(cfg, seed=3)
    b3 = bf.synth_block(cfg, seed=4)
    assert np.array_equal(b1, b2) and not np.array_equal(b1, b3)
    pairs = np.array([orc.unpack_byte(int(v)) for v in b1[:20000]])
    assert pairs.min() >= -7 and pairs.max() <= 7                 # clipped to +-7
    assert abs(pairs.mean()) < 0.05
    # sigma 2 Gaussian, rounded: variance ~ 4 + 1/12
    assert 3.8 < pairs.var() < 4.3


def test_synth_block_split_r_order_is_a_permutation(bf, orc):
    c0 = bf.config(nant=5, nchan=24, npkt=32, nbeam=8, input_order=0)
    c1 = bf.config(nant=5, nchan=24, npkt=32, nbeam=8, input_order=1)
    b0, b1 = bf.synth_block(c0, 9), bf.synth_block(c1, 9)
    d0, d1 = orc.dims(5, 24, 32, 8, 8, 0), orc.dims(5, 24, 32, 8, 8, 1)
    for p, a, c, t, q in [(0, 0, 0, 0, 0), (3, 4, 23, 1, 1), (31, 2, 9, 0, 1), (17, 1, 15, 1, 0)]:
        assert b0[orc.in_index(d0, p, a, c, t, q)] == b1[orc.in_index(d1, p, a, c, t, q)]


def test_synth_geometry(bf):
    antpos, cal, theta = bf.synth_geometry(64, 48, 32, 7, 1200.0, 0.1, 1.75)
    assert antpos.min() >= 0 and antpos.max() <= 1200
    amp = np.hypot(cal[..., 0], cal[..., 1])
    assert amp.min() >= 0.9 - 1e-6 and amp.max() <= 1.1 + 1e-6
    assert theta[0] == pytest.approx(-1.75 * np.pi / 180) and theta[-1] == pytest.approx(1.75 * np.pi / 180)


def test_antpol_rms_matches_reference_rms_printer(bf, orc):
    """rms_printer (dsaX_beamformer.cu:217-238): sqrt(sum over 16 chunnels of
    (re^2 + im^2) / 32) for one packet and group."""
    cfg = bf.config(nant=10, nchan=384, npkt=32, nbeam=8)
    blk = bf.synth_block(cfg, 2)
    rms = bf.antpol_rms(cfg, blk, 8, 24)
    d = orc.dims(10, 384, 32, 8, 8)
    for pol in (0, 1):
        for ant in range(64):
            s = 0.0
            if ant < 10:
                for ci in range(8):
                    for t in (0, 1):
                        re, im = orc.unpack_byte(int(blk[orc.in_index(d, 8, ant, 24 * 8 + ci, t, pol)]))
                        s += re * re + im * im
            assert rms[pol, ant] == pytest.approx(np.sqrt(s / 32.0), rel=1e-6)


def test_cli_usage_and_missing_io():
    exe = os.path.join(ROOT, "dsa110-xengine_amd", "bin", "dsaX_beamformer")
    r = subprocess.run([exe, "-h"], capture_output=True, text=True)
    assert r.returncode == 0 and "-z fch1 in MHz [default 1530]" in r.stdout
    r = subprocess.run([exe, "-f", "nofile"], capture_output=True, text=True)
    assert r.returncode != 0 and "psrdada" in r.stderr
    assert "fp16|fp8split|fp6split|fp6split3" in subprocess.run([exe, "-h"], capture_output=True,
                                                                text=True).stdout
    r = subprocess.run([exe, "-P", "fp4", "-I", "-", "-O", "-"], capture_output=True, text=True)
    assert r.returncode != 0 and "unknown precision fp4" in r.stderr


@pytest.mark.parametrize("nant", ["65", "256", "0"])
def test_cli_rejects_antenna_counts_beyond_the_weights_file(nant):
    # the weights file holds 64 antennas: larger -n is an error before any GPU call
    exe = os.path.join(ROOT, "dsa110-xengine_amd", "bin", "dsaX_beamformer")
    r = subprocess.run([exe, "-n", nant, "-I", "-", "-O", "-"], capture_output=True, text=True,
                       input="")
    assert r.returncode != 0 and "outside 1..64" in r.stderr

"""The beam-major output layout (DSAXBF_LAYOUT_BEAM_MAJOR): the epilogue
writes [beam][t][ch] instead of the BF_BLOCK_KEY [t][beam][ch], so each 64-beam
slice of the block is the buffer dsaX_dbnic's transmit thread builds with its
per-block transpose (dsaX_dbnic.c:111-116) and the corner turn sends it as it
is.  Bit-exact against the time-major output put through the oracle's
restatement of that transpose, for every kernel schedule and precision, the
-s/-q kernels, and the float power."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def _both(bf, geo, block, scale, variant=None, **kw):
    res = []
    for layout in (bf.LAYOUT_TIME_MAJOR, bf.LAYOUT_BEAM_MAJOR):
        cfg = bf.config(output_layout=layout, **kw)
        plan = bf.Plan(cfg)
        plan.set_weights(*geo)
        if variant is not None:
            plan.variant = variant
        out, pw, bp = plan.process_block(block, scale, want_power=True, want_beam_power=True)
        plan.close()
        res.append((out, pw, bp, cfg))
    return res


@pytest.mark.parametrize("prec,variant,nant,nbeam,tavg,stuff", [
    (0, "shared", 64, 256, 8, 0),       # the dbnic shape: 512 outputs x 256 beams x 48 ch
    (1, "shared", 64, 256, 32, 0),
    (2, "shared", 48, 256, 16, 0),
    (3, "shared", 64, 128, 64, 0),
    (0, "classic128s", 64, 128, 8, 0),
    (1, "classic64", 30, 64, 16, 0),
    (0, "shared", 200, 256, 32, 0),     # 8-wave schedule
    (0, "shared", 64, 256, 8, 1),       # -s
    (1, "shared", 64, 256, 8, 2),       # -q
])
def test_beam_major_is_dbnic_transpose(bf, orc, prec, variant, nant, nbeam, tavg, stuff):
    nchan, npkt = 384, 2048 if tavg == 8 else 256
    var = {"shared": bf.VARIANT_SHARED_256, "classic128s": bf.VARIANT_CLASSIC_128S,
           "classic64": bf.VARIANT_CLASSIC_64}[variant]
    kw = dict(nant=nant, nchan=nchan, npkt=npkt, nbeam=nbeam, tavg=tavg, precision=prec,
              weight_mode=bf.WEIGHTS_GROUP, stuffants=stuff)
    cfg0 = bf.config(**kw)
    antpos, cal, theta = bf.synth_geometry(nant, cfg0.ncal, nbeam, 3, 1200.0, 0.1, 1.75)
    geo = (antpos, cal, bf.reference_freqs(1498.75, nchan), theta)
    block = bf.synth_block(cfg0, 12)
    scale = np.full(nbeam, 1e-3, np.float32)
    (o0, p0, b0, _), (o1, p1, b1, c1) = _both(bf, geo, block, scale, var, **kw)
    assert o1.shape == (nbeam, c1.nout, c1.ngroup)
    for s in range(nbeam // 64):
        assert np.array_equal(o1[64 * s: 64 * s + 64], orc.dbnic_op(o0, s)), s
    assert np.array_equal(p1, p0.transpose(1, 0, 2))
    assert np.array_equal(b0, b1)


def test_cli_beam_major_flag(bf, orc, tmp_path):
    """dsaX_beamformer -B writes each block beam-major: the file-mode output
    of -B is the default output transposed block by block."""
    nant = 24
    cfg = bf.config(nant=nant, nchan=384, npkt=2048, nbeam=256, tavg=8)
    fin = tmp_path / "in.dat"
    with open(fin, "wb") as fh:
        for k in range(3):
            fh.write(bf.synth_block(cfg, 30 + k).tobytes())
    exe = os.path.join(ROOT, "dsa110-xengine_amd", "bin", "dsaX_beamformer")
    outs = []
    for extra in ([], ["-B"]):
        fo = tmp_path / f"out{len(extra)}.dat"
        r = subprocess.run([exe, "-z", "1498.75", "-n", str(nant), "-I", str(fin), "-O", str(fo)] + extra,
                           capture_output=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs.append(np.fromfile(fo, np.uint8).reshape(3, -1))
    nob = bf.output_bytes(cfg)
    for k in range(3):
        tm = outs[0][k].reshape(cfg.nout, 256, 48)
        assert np.array_equal(outs[1][k], np.ascontiguousarray(tm.transpose(1, 0, 2)).ravel()), k
    assert outs[0].shape[1] == nob

"""dsaxbf_execute_noalloc through the C ABI: the same bytes as dsaxbf_execute
(uint8 output, float power, per-beam power) on two calls in a row, the second
with a smaller block that reuses the static scratch array the first one
filled, then the first shape again; a block whose partial sums exceed the
array is rejected with DSAXBF_EINVAL."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(bf, nant, nchan, npkt, nbeam, tavg, prec, seed):
    cfg = bf.config(nant=nant, nchan=nchan, npkt=npkt, nbeam=nbeam, tavg=tavg, precision=prec,
                    weight_mode=bf.WEIGHTS_FINE, angle_mode=bf.ANGLE_SIN)
    antpos, cal, theta = bf.synth_geometry(nant, nchan, nbeam, seed)
    plan = bf.Plan(cfg)
    plan.set_weights(antpos, cal, bf.reference_freqs(1498.75, nchan), theta)
    blk = bf.synth_block(cfg, seed + 1)
    bufs = dict(d_in=bf.DeviceBuffer(blk.nbytes), d_sc=bf.DeviceBuffer(4 * nbeam))
    bufs["d_in"].upload(blk)
    bufs["d_sc"].upload(np.full(nbeam, 2.0 ** -12, np.float32))
    return cfg, plan, bufs


def _outputs(bf, cfg, plan, bufs, noalloc):
    nout = bf.output_bytes(cfg)
    d_out, d_pw, d_bp = bf.DeviceBuffer(nout), bf.DeviceBuffer(4 * nout), bf.DeviceBuffer(4 * cfg.nbeam)
    run = plan.execute_noalloc if noalloc else plan.execute
    run(bufs["d_in"], bufs["d_sc"], d_out, d_pw, d_bp)
    res = (d_out.download(np.empty(nout, np.uint8), plan.stream),
           d_pw.download(np.empty(nout, np.float32), plan.stream),
           d_bp.download(np.empty(cfg.nbeam, np.float32), plan.stream))
    for b in (d_out, d_pw, d_bp):
        b.free()
    return res


def test_noalloc_matches_execute_twice_then_smaller(bf):
    big = _setup(bf, 64, 96, 512, 256, 16, bf.PREC_FP8SPLIT, 3)
    small = _setup(bf, 40, 16, 128, 100, 32, bf.PREC_FP16, 9)
    for tag, (cfg, plan, bufs) in (("first", big), ("second, smaller", small), ("first again", big)):
        want = _outputs(bf, cfg, plan, bufs, noalloc=False)
        got = _outputs(bf, cfg, plan, bufs, noalloc=True)
        for w, g in zip(want, got):
            assert np.array_equal(w, g), tag
        assert want[2].min() > 0, tag                       # beam powers were produced
    for cfg, plan, bufs in (big, small):
        plan.close()
        for b in bufs.values():
            b.free()


def test_noalloc_rejects_blocks_beyond_the_scratch(bf):
    # 1024 beams x 128 units x 48 groups x 4 bytes = 25 MB of partial sums > 8 MiB
    cfg, plan, bufs = _setup(bf, 32, 384, 2048, 1024, 32, bf.PREC_FP16, 5)
    nout = bf.output_bytes(cfg)
    d_out = bf.DeviceBuffer(nout)
    with pytest.raises(bf.BeamformerError) as e:
        plan.execute_noalloc(bufs["d_in"], bufs["d_sc"], d_out)
    assert e.value.code == -1 and "static" in str(e.value)
    plan.execute(bufs["d_in"], bufs["d_sc"], d_out)         # the allocating entry point still runs it
    plan.close()
    d_out.free()
    for b in bufs.values():
        b.free()

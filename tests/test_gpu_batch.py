"""The uneven batch workload on the GPU: 24 blocks of 64 antennas with seeded
channel counts {48, 96, 192, 384}, beam counts {64, 128, 256, 512} and time
lengths {1024, 2048, 4096} samples (dsaxbf.uneven_batch_shapes, the draw the
benchmark uses), 4-bit Gaussian voltages, geometry-derived fine-channel
weights, tavg 32, on the FP16 and FP8-split paths together in ONE batch graph
(48 blocks).

Checked: every block's batch output and beam power equal its own execute bit
for bit (over two replays with eager work between them), and every block
matches the oracle on one channel group over its first 128 samples: the
kernel against the oracle fed the device's own weights within 1e-5, and end
to end against the exact FP64 weights (FP16 within 1e-3; FP8-split within its
weight encoding's own error).  Reference loop being replaced:
dsaX_beamformer.cu:955-1003 (one fixed-shape block per process there)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED, NBLOCKS, NANT, TAVG = 2025, 24, 64, 32


@pytest.mark.timeout(300)
def test_uneven_batch_both_precisions_one_graph(bf, orc):
    shapes = bf.uneven_batch_shapes(SEED, NBLOCKS)
    assert len(shapes) == NBLOCKS and {c for c, _, _ in shapes} <= {48, 96, 192, 384}
    plans, bufs, blocks, geos = [], [], [], []
    d_ins = []
    for i, (c, b, t) in enumerate(shapes):
        cfg = bf.config(nant=NANT, nchan=c, npkt=t // 2, nbeam=b, tavg=TAVG,
                        weight_mode=bf.WEIGHTS_FINE, angle_mode=bf.ANGLE_SIN)
        blk = bf.synth_block(cfg, SEED + 100 + i)
        d_in = bf.DeviceBuffer(blk.nbytes)
        d_in.upload(blk)
        d_ins.append(d_in)
        blocks.append(blk)
        geos.append(bf.synth_geometry(NANT, c, b, SEED + i, 1200.0, 0.1, 1.75))
    for prec in (bf.PREC_FP16, bf.PREC_FP8SPLIT):
        for i, (c, b, t) in enumerate(shapes):
            cfg = bf.config(nant=NANT, nchan=c, npkt=t // 2, nbeam=b, tavg=TAVG, precision=prec,
                            weight_mode=bf.WEIGHTS_FINE, angle_mode=bf.ANGLE_SIN)
            antpos, cal, theta = geos[i]
            plan = bf.Plan(cfg)
            plan.set_weights(antpos, cal, bf.reference_freqs(1498.75, c), theta)
            d_sc = bf.DeviceBuffer(4 * b)
            d_sc.upload(np.full(b, 1e-4, np.float32))
            d_out, d_bp = bf.DeviceBuffer(bf.output_bytes(cfg)), bf.DeviceBuffer(4 * b)
            plans.append(plan)
            bufs.append((d_ins[i], d_sc, d_out, d_bp, cfg, i))
    # each block alone first
    want = []
    for plan, (d_in, d_sc, d_out, d_bp, cfg, i) in zip(plans, bufs):
        plan.execute(d_in, d_sc, d_out, None, d_bp)
        want.append((d_out.download(np.empty((cfg.nout, cfg.nbeam, cfg.ngroup), np.uint8), plan.stream),
                     d_bp.download(np.empty(cfg.nbeam, np.float32), plan.stream)))
    batch = bf.Batch(plans, [x[0] for x in bufs], [x[1] for x in bufs], [x[2] for x in bufs],
                     [x[3] for x in bufs])
    # one beamforming launch per precision, one for all 48 beam-power sums
    assert batch.kernels == 3
    for rep in range(2):
        for d_in, d_sc, d_out, d_bp, cfg, i in bufs:
            d_out.upload(np.zeros(bf.output_bytes(cfg), np.uint8))
            d_bp.upload(np.zeros(cfg.nbeam, np.float32))
        batch.launch()
        for k, (d_in, d_sc, d_out, d_bp, cfg, i) in enumerate(bufs):
            out = d_out.download(np.empty((cfg.nout, cfg.nbeam, cfg.ngroup), np.uint8))
            bp = d_bp.download(np.empty(cfg.nbeam, np.float32))
            assert np.array_equal(out, want[k][0]) and np.array_equal(bp, want[k][1]), (rep, k)
        plans[0].execute(bufs[0][0], bufs[0][1], bufs[0][2], None, bufs[0][3])   # eager work
    batch.close()
    # the oracle: one channel group per block (its 8 channels cut out of the
    # capture-order block), the first 64 packets (128 samples, 4 outputs)
    npk = 64
    for k, (plan, (d_in, d_sc, d_out, d_bp, cfg, i)) in enumerate(zip(plans, bufs)):
        c, b, t = shapes[i]
        g = i % cfg.ngroup
        ch = slice(8 * g, 8 * g + 8)
        antpos, cal, theta = geos[i]
        freqs = bf.reference_freqs(1498.75, c)[ch]
        sub = np.ascontiguousarray(blocks[i].reshape(t // 2, NANT, c, 4)[:npk, :, ch]).ravel()
        d = orc.dims(NANT, 8, npk, b, TAVG)
        W = orc.weights(d, 1, 1, antpos, cal[:, ch], freqs, theta)
        Wd = plan.get_weights()[ch]
        # the block's float power from a plain execute (the batch does not keep it)
        d_pw = bf.DeviceBuffer(4 * bf.output_bytes(cfg))
        plan.execute(d_in, d_sc, d_out, d_pw, None)
        pw = d_pw.download(np.empty((cfg.nout, b, cfg.ngroup), np.float32), plan.stream)
        d_pw.free()
        assert np.array_equal(d_out.download(np.empty((cfg.nout, b, cfg.ngroup), np.uint8),
                                             plan.stream), want[k][0])
        p = pw[: 2 * npk // TAVG, :, g].astype(np.float64)
        ref = orc.power(d, 1, sub, W)[:, :, 0]
        ref_dev = orc.power(d, 1, sub, Wd)[:, :, 0]
        relk = np.abs(p - ref_dev) / ref_dev
        assert relk.max() <= 1e-5, (k, relk.max())
        rel = np.abs(p - ref) / ref
        if cfg.precision == bf.PREC_FP16:
            assert rel.max() <= 1e-3, (k, rel.max())
        else:
            ref_enc = orc.power(d, 1, sub, orc.quant_weights(d, 1, cfg.precision, W))[:, :, 0]
            assert (rel <= np.abs(ref_enc - ref) / ref + 1e-5).all(), k
        out_ref = orc.requant_block(ref[:, :, None], np.full(b, 1e-4, np.float32))[:, :, 0]
        du = np.abs(want[k][0][: 2 * npk // TAVG, :, g].astype(int) - out_ref.astype(int))
        assert np.minimum(du, 256 - du).max() <= 1, k
    for plan in plans:
        plan.close()
    for d_in, d_sc, d_out, d_bp, cfg, i in bufs:
        for x in (d_sc, d_out, d_bp):
            x.free()
    for x in d_ins:
        x.free()

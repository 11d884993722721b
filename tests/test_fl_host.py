"""The flagger's CPU functions in libdsaxbf (host/fl_host.c: selection-based
medians) against the oracle's statement-by-statement restatement of
gpu_flagger.cu:530-790 (bubble-sort medians), and the C ABI exports of
include/dsaxfl.h.  CPU only."""
import os
import re

import numpy as np
import pytest

import dsaxfl as fl
import orc
from conftest import ROOT
from dsaxbf import BeamformerError


def test_every_declared_flagger_symbol_is_exported():
    src = open(os.path.join(ROOT, "include", "dsaxfl.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(dsaxfl_[a-z0-9_]+)\s*\(", src)))
    assert len(names) >= 25
    for n in names:
        assert hasattr(fl.lib, n), n
    assert fl.EXPORTED == names


@pytest.mark.parametrize("n", [1, 2, 3, 20, 21, 256, 768, 1024, 1025])
def test_medval_selection_equals_bubble_sort(n):
    rng = np.random.default_rng(n)
    a = rng.normal(size=n).astype(np.float32)
    assert fl.medval(a) == orc.fl_medval(a)
    b = np.round(a * 3).astype(np.float32)          # many ties
    assert fl.medval(b) == orc.fl_medval(b)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_channel_masks_equal_the_oracle(seed):
    rng = np.random.default_rng(seed)
    nb, nc = 5, 1024
    spec = rng.normal(48, 0.5, (nb, nc)).astype(np.float32)
    spec[:, 300:310] += rng.uniform(2, 8, 10).astype(np.float32)
    spec[2, 700] -= 9
    spec[3] = np.round(spec[3] * 4) / 4                 # ties in the medians
    for thr in (3.0, 5.0, 5.5):
        a = fl.mask("channflag", spec, thr, perm=(10, 500))
        b = orc.fl_mask("channflag", spec, thr, perm=(10, 500))
        assert np.array_equal(a, b) and a.sum() > nb
        a = fl.mask("simple_channflag", spec, thr)
        b = orc.fl_mask("simple_channflag", spec, thr)
        assert np.array_equal(a, b)


@pytest.mark.parametrize("nc", [278, 300, 1024])
def test_channflag_window_medians_with_ties_and_flat_runs(nc):
    """channflag's 21-channel window medians (a sorted window sliding along
    the band, clamped at the top edge) against the oracle's per-window sort:
    heavy ties, a flat run longer than the window, a spike, the narrowest band
    (nchan 278: 22 channels between the 128-channel edges)."""
    rng = np.random.default_rng(nc)
    nb = 4
    spec = rng.normal(48, 2, (nb, nc)).astype(np.float32)
    spec[0] = np.round(spec[0])                          # few distinct values
    spec[1, 130:190] = 50.0                              # flat run
    spec[2, nc // 2] += 25.0                             # spike
    spec[3] = np.float32(47.5)                           # constant band: MAD 0
    for thr in (2.0, 4.0):
        a = fl.mask("channflag", spec, thr, perm=(10, 140))
        assert np.array_equal(a, orc.fl_mask("channflag", spec, thr, perm=(10, 140)))


def test_masks_over_many_beams_equal_the_oracle():
    """64 beams (the reference's block), which the host functions split over
    threads by beam ranges: every mask equals the oracle's single pass."""
    rng = np.random.default_rng(64)
    nb, nc, nt = 64, 1024, 4096
    spec = rng.normal(48, 1.5, (nb, nc)).astype(np.float32)
    spec[::7, 400:406] += 9
    spec[5] = np.round(spec[5])
    for kind in ("channflag", "simple_channflag"):
        a = fl.mask(kind, spec, 4.0, perm=(10,))
        assert np.array_equal(a, orc.fl_mask(kind, spec, 4.0, perm=(10,))) and a.sum() > nb
    ts = rng.normal(40, 1, (nb, nt)).astype(np.float32)
    ts[::5, 1000:1004] += 15
    a = fl.mask("simple_tsflag", ts, 6.0)
    assert np.array_equal(a, orc.fl_mask("simple_tsflag", ts, 6.0)) and a.sum() >= 4 * 13


def test_masks_set_flagged_entries_only_and_accumulate():
    """The mask functions write 1 where they flag and leave every other entry
    as it was (gpu_flagger.cu:597,642): over 64 beams (the threaded split) a
    mask pre-filled with 7 keeps its 7s, and the reference loop's four calls
    into one zeroed mask (gpu_flagger.cu:1339-1347) give the OR of the
    oracle's four masks."""
    rng = np.random.default_rng(77)
    nb, nc, nt = 64, 1024, 4096
    spec = rng.normal(48, 1.5, (nb, nc)).astype(np.float32)
    spec[::5, 600:604] += 10
    for kind, args in (("channflag", dict(perm=(10,))), ("simple_channflag", {})):
        want = orc.fl_mask(kind, spec, 4.0, **args)
        got = fl.mask(kind, spec, 4.0, out=np.full((nb, nc), 7, np.int32), **args)
        assert np.array_equal(got, np.where(want == 1, 1, 7)) and want.sum() > nb
    ts = rng.normal(40, 1, (nb, nt)).astype(np.float32)
    ts[::3, 2000:2003] += 15
    want = orc.fl_mask("simple_tsflag", ts, 6.0)
    got = fl.mask("simple_tsflag", ts, 6.0, out=np.full((nb, nt), 7, np.int32))
    assert np.array_equal(got, np.where(want == 1, 1, 7)) and want.sum() > 0
    var = rng.normal(4, 0.3, (nb, nc)).astype(np.float32)
    var[1::9, 300:302] += 3
    mx = rng.normal(70, 2, (nb, nc)).astype(np.float32)
    pp = rng.normal(5, 0.4, (nb, nc)).astype(np.float32)
    pp[2::11, 800] += 4
    m = np.zeros((nb, nc), np.int32)
    m = fl.mask("channflag", spec, 5.0, perm=(10,), out=m)
    m = fl.mask("channflag", var, 5.5, perm=(10,), out=m)
    m = fl.mask("channflag", mx, 5.0, perm=(10,), out=m)
    m = fl.mask("simple_channflag", pp, 5.0, out=m)
    ors = (orc.fl_mask("channflag", spec, 5.0, perm=(10,)) | orc.fl_mask("channflag", var, 5.5, perm=(10,)) |
           orc.fl_mask("channflag", mx, 5.0, perm=(10,)) | orc.fl_mask("simple_channflag", pp, 5.0))
    assert np.array_equal(m, ors) and m.sum() > nb


def test_time_mask_gather_bpwr_and_median_calc_equal_the_oracle():
    rng = np.random.default_rng(4)
    nb, nt = 4, 16384
    ts = rng.normal(40, 1, (nb, nt)).astype(np.float32)
    ts[1, 5000:5010] += 20
    ts[3, 100] += 30
    a = fl.mask("simple_tsflag", ts, 6.0)
    assert np.array_equal(a, orc.fl_mask("simple_tsflag", ts, 6.0)) and a[1, 5000:5010].all()
    assert np.array_equal(fl.gather_mask(a), np.flatnonzero(a == 1).astype(np.int32))
    spec = rng.uniform(30, 60, (nb, 1024)).astype(np.float32)
    assert np.array_equal(fl.calc_bpwr(spec), orc.fl_calc_bpwr(spec))
    arr = rng.normal(size=(64, 1024)).astype(np.float32)
    assert np.array_equal(fl.median_calc(arr), orc.fl_median_calc(arr))


def test_config_defaults_are_the_reference_options():
    c = fl.config()
    assert (c.nbeam, c.ntime, c.nchan) == (64, 16384, 1024)
    assert c.thresh == 5.0 and c.tsthresh == 10.0 and abs(c.mod_thresh - 5e-4) < 1e-9
    assert (c.pwr, c.mkrand, c.tsflag, c.nperm, c.perm[0]) == (0, 0, 0, 1, 10)


def test_device_entry_points_fail_cleanly_on_bad_arguments():
    with pytest.raises(BeamformerError) as e:
        fl.stats(None, 1, 4096, 512)
    assert e.value.code == -1
    with pytest.raises(BeamformerError) as e:
        fl.Plan(fl.config(ntime=4000))
    assert e.value.code == -1 and "4096" in str(e.value)

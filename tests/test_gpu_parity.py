"""Parity of the HIP path (through the C-ABI) against the CPU oracle.

Integer data: bit-exact planes and IQRM masks.  float32 data: within
|gpu - oracle| <= 1e-6 * sum_c |x_c| per output (the FP64 oracle)."""
import sys

import numpy as np
import pytest

from astroflow_amd import capi, synth
from oracle import oracle as O
from tests.helpers import make_spec, oracle_spec, small_band

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_both(x, spec, want_masks=True):
    plan = capi.Plan(spec)
    got, masks = plan.run_host(x, want_masks=want_masks)
    ref, rmask, t0, t1, _ = O.run_pipeline(x, oracle_spec(spec), O.IQRM_GPU)
    if spec.use_iqrm:
        w0, w1 = plan.windows()
        assert np.array_equal(w0, t0) and np.array_equal(w1, t1)
    plan.close()
    return got, ref, masks, rmask


def noise(rng, n, nc, dtype=np.uint8):
    if dtype == np.float32:
        return rng.standard_normal((n, nc)).astype(np.float32)
    vmax = np.iinfo(dtype).max
    return np.clip(np.round(rng.normal(vmax / 2, vmax / 16, (n, nc))), 0, vmax).astype(dtype)


def assert_f32_close(got, ref, x, spec):
    # tolerance per output: 1e-6 * sum over the contributing samples of |x|;
    # bounded above by the same plane over |x|
    absx = np.abs(x).astype(np.float32)
    s2 = oracle_spec(spec)
    s2.use_iqrm = s2.use_zero_dm = False
    s2.bad_channels = []
    bound, *_ = O.run_pipeline(absx + 1.0, s2)
    assert np.all(np.abs(got.astype(np.float64) - ref) <= 1e-6 * bound)


def test_direct_u8_plain(gpu_lib):
    say("case: direct u8 300ch x 5000, 37 DMs")
    rng = np.random.default_rng(1)
    b = small_band(300, 1200.0, 1500.0, 1e-4)
    x = noise(rng, 5000, 300)
    spec = make_spec(x, b, 0.0, 7.0, 37, method="direct", dd_c0=3, dd_c1=290, nout=4321)
    got, ref, *_ = run_both(x, spec)
    assert np.array_equal(got, ref)


def test_direct_u8_many_channels_flush(gpu_lib):
    say("case: direct u8 1000ch (>256-channel flush)")
    rng = np.random.default_rng(2)
    b = small_band(1000, 1000.0, 1500.0, 2e-4)
    x = np.full((3000, 1000), 255, np.uint8)   # worst case for the 16-bit lanes
    x[::3] = noise(rng, 1000, 1000)
    spec = make_spec(x, b, 10.0, 11.0, 9, method="direct")
    got, ref, *_ = run_both(x, spec)
    assert np.array_equal(got, ref)


def test_subband_u8_full_rfi(gpu_lib):
    say("case: subband u8 512ch, IQRM windows with hop + zero-DM + static mask")
    rng = np.random.default_rng(3)
    b = small_band(512, 1100.0, 1500.0, 2e-4)
    x = noise(rng, 6000, 512)
    synth.inject_narrowband_rfi(x, rng, frac=0.05, amp=40, n_stuck=5, pmax=512)
    synth.inject_pulse(x, b.freq, b.f_top, float(b.tsamp), 150.0, 500, 2e-3, 48.0)
    spec = make_spec(x, b, 0.0, 1.5, 130, method="subband", sbw=64, ndm0=16,
                     use_iqrm=True, iqrm_c0=0, iqrm_c1=511, iqrm_mode=1,
                     iqrm_radius_frac=0.02, iqrm_nsigma=3.0, iqrm_W=2500, iqrm_H=1500,
                     use_zero_dm=True, zdm_c0=0, zdm_c1=511, bad_channels=[7, 100, 511, 900])
    got, ref, masks, rmask = run_both(x, spec)
    assert masks.shape == rmask.shape and masks.sum() > 0
    assert np.array_equal(masks, rmask)
    assert np.array_equal(got, ref)


def test_iqrm_many_lags(gpu_lib):
    # radius 256 channels at geofactor 1.05: 164 lags, so the vote merge keeps
    # six 32-bit words of flags per channel; the IQRM range [5, 290) also ends
    # inside a 16-channel group of the statistics pass
    say("case: direct u8 300ch, IQRM with 164 lags")
    rng = np.random.default_rng(4)
    b = small_band(300, 1200.0, 1500.0, 1e-4)
    x = noise(rng, 4000, 300)
    synth.inject_narrowband_rfi(x, rng, frac=0.08, amp=40, n_stuck=4, pmax=300)
    spec = make_spec(x, b, 0.0, 5.0, 20, method="direct", use_iqrm=True, iqrm_c0=5,
                     iqrm_c1=290, iqrm_mode=1, iqrm_radius_frac=0.9, iqrm_geofactor=1.05,
                     iqrm_nsigma=2.0)
    assert len(O.iqrm_lags(O.iqrm_radius(0.9, 285), 1.05)) > 128
    got, ref, masks, rmask = run_both(x, spec)
    assert masks.sum() > 0
    assert np.array_equal(masks, rmask)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("radius_frac,W", [(0.2, 0), (0.45, 700)])
def test_iqrm_vote_blocks_with_wide_halo(gpu_lib, radius_frac, W):
    # 3000 IQRM channels: the vote runs in three 1024-channel blocks whose
    # lags reach 600 and 1350 channels, so each block gathers flags and
    # casting counts far across its neighbours (and across the range edges)
    say(f"case: IQRM vote blocks, 3000 ch, radius_frac {radius_frac}, W {W}")
    rng = np.random.default_rng(12)
    b = small_band(3001, 1100.0, 1500.0, 1e-4)
    x = noise(rng, 2000, 3001)
    synth.inject_narrowband_rfi(x, rng, frac=0.06, amp=40, n_stuck=6, pmax=300)
    spec = make_spec(x, b, 0.0, 10.0, 8, method="direct", use_iqrm=True, iqrm_c0=0,
                     iqrm_c1=3000, iqrm_mode=1, iqrm_radius_frac=radius_frac,
                     iqrm_nsigma=2.5, iqrm_W=W, iqrm_H=W // 2)
    got, ref, masks, rmask = run_both(x, spec)
    assert masks.sum() > 0
    assert np.array_equal(masks, rmask)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("ds", [2, 3])
def test_subband_u8_downsampled(gpu_lib, ds):
    say(f"case: subband u8 ds={ds}")
    rng = np.random.default_rng(10 + ds)
    b = small_band(256, 1200.0, 1500.0, 1e-4)
    x = noise(rng, 7001, 256)
    spec = make_spec(x, b, 5.0, 2.0, 70, method="subband", sbw=32, ndm0=8, ds=ds,
                     use_iqrm=True, iqrm_c0=1, iqrm_c1=255, iqrm_mode=0,
                     iqrm_radius_frac=0.1, use_zero_dm=True, zdm_c0=1, zdm_c1=255)
    got, ref, masks, rmask = run_both(x, spec)
    assert np.array_equal(masks, rmask)
    assert np.array_equal(got, ref)


def test_subband_u8_wide_subbands_generic(gpu_lib):
    say("case: subband u8 with 300-channel subbands (u32 partial sums)")
    rng = np.random.default_rng(5)
    b = small_band(600, 1000.0, 1500.0, 2e-4)
    x = noise(rng, 4000, 600)
    spec = make_spec(x, b, 0.0, 3.0, 40, method="subband", sbw=300, ndm0=4)
    got, ref, *_ = run_both(x, spec)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("ds", [1, 2, 4])
@pytest.mark.parametrize("dtype", [np.uint16, np.uint32])
@pytest.mark.parametrize("method", ["direct", "subband"])
def test_wide_integer_types(gpu_lib, dtype, method, ds):
    # ds > 1 runs the binning kernel on 16/32-bit data (sqrt compression with
    # the type's own vmax, src/core/dedispered.cu:337-376) before IQRM
    say(f"case: {np.dtype(dtype).name} {method} ds {ds}")
    rng = np.random.default_rng(7 + ds)
    b = small_band(200, 1200.0, 1500.0, 1e-4)
    x = noise(rng, 3001, 200, dtype)
    if dtype == np.uint32:
        # values near the top of the range: the binned sums exceed 2^32
        x[::7] = np.uint32(4294967295 - 5)
    spec = make_spec(x, b, 0.0, 5.0, 33, method=method, sbw=25, ndm0=8, ds=ds,
                     use_iqrm=True, iqrm_c0=0, iqrm_c1=199, iqrm_mode=1,
                     use_zero_dm=True, zdm_c0=0, zdm_c1=199)
    got, ref, masks, rmask = run_both(x, spec)
    assert np.array_equal(masks, rmask)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("method", ["direct", "subband"])
def test_float32_within_tolerance(gpu_lib, method):
    say(f"case: float32 {method}")
    rng = np.random.default_rng(8)
    b = small_band(256, 1200.0, 1500.0, 1e-4)
    x = noise(rng, 3000, 256, np.float32)
    spec = make_spec(x, b, 0.0, 4.0, 40, method=method, sbw=32, ndm0=8, ds=2,
                     use_zero_dm=True, zdm_c0=0, zdm_c1=255)
    got, ref, *_ = run_both(x, spec, want_masks=False)
    assert got.dtype == np.float32
    assert_f32_close(got, ref, x, spec)


def test_delay_tables_equal_oracle(gpu_lib):
    say("case: delay/residual tables")
    b = synth.band_1000_1500(4096)
    x = np.zeros((1, 4096), np.uint8)
    spec = make_spec(x, b, 0.0, 0.5, 6001, method="subband", sbw=128, ndm0=32)
    spec.nsamp = 160000
    plan = capi.Plan(spec)
    d1, res = plan.tables()
    f = b.freq
    nnom = (6001 + 31) // 32
    assert np.array_equal(d1, O.delay_table(f, 0, 4096, f[-1], float(b.tsamp), 1,
                                            np.arange(nnom) * 32 * 0.5))
    assert np.array_equal(res, O.residual_table(f, 0, 4096, 128, f[-1], float(b.tsamp), 1,
                                                0.0, 0.5, 6001, 32))
    plan.close()


@pytest.mark.parametrize("dtype,tile", [(np.uint8, 0), (np.uint8, 256), (np.uint16, 0),
                                        (np.float32, 0)])
def test_subband_time_super_tiles(gpu_lib, dtype, tile):
    # a small intermediate budget splits the output into several time
    # super-tiles (stage 1 + stage 2 per super-tile); the plane is unchanged
    say(f"case: subband {np.dtype(dtype).name} in time super-tiles (stage-2 tile {tile})")
    rng = np.random.default_rng(21)
    b = small_band(384, 1100.0, 1500.0, 1e-4)
    x = noise(rng, 20000, 384, dtype)
    spec = make_spec(x, b, 0.0, 1.0, 150, method="subband", sbw=64, ndm0=16, nout=17000,
                     use_zero_dm=True, zdm_c0=0, zdm_c1=384, tile_t=tile,
                     inter_budget=10 * 6 * 2 * 4096)
    plan = capi.Plan(spec)
    assert plan.info.inter_bytes <= 3 * 10 * 6 * 2 * 4096 + (1 << 20)
    plan.close()
    got, ref, *_ = run_both(x, spec, want_masks=False)
    if dtype == np.float32:
        assert_f32_close(got, ref, x, spec)
    else:
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("sbw,tile", [(16, 0), (16, 512), (32, 0), (32, 1024), (32, 256)])
def test_subband_u8_many_subbands_large_halo(gpu_lib, sbw, tile):
    # A large residual halo (~1400 samples at the lowest subband, a few at the
    # highest).  64 subbands fit LDS as packed rows, each row holding the tile
    # plus its own subband's halo; 128 subbands do not, and stage 2 stages them
    # in chunks.
    say(f"case: subband u8 2048ch, {sbw}-ch subbands, large halo (stage-2 tile {tile})")
    rng = np.random.default_rng(23)
    b = synth.band_1000_1500(2048)
    x = noise(rng, 6000, 2048)
    spec = make_spec(x, b, 100.0, 2.0, 64, method="subband", sbw=sbw, ndm0=16, tile_t=tile,
                     use_zero_dm=True, zdm_c0=0, zdm_c1=2048)
    got, ref, *_ = run_both(x, spec, want_masks=False)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("dtype,ds", [(np.uint8, 1), (np.uint8, 2), (np.uint16, 1),
                                      (np.uint32, 1), (np.float32, 1), (np.float32, 4)])
def test_clean_per_super_tile(gpu_lib, dtype, ds):
    # a clean budget below the whole channel-major buffer: each super-tile's
    # rows are cleaned just before its dedispersion (the overlap rows once per
    # super-tile, columns past the input zeroed); the plane and the masks equal
    # the oracle's and the whole-buffer plan's
    say(f"case: clean per super-tile, {np.dtype(dtype).name} ds {ds}")
    rng = np.random.default_rng(31)
    b = small_band(384, 1100.0, 1500.0, 1e-4)
    x = noise(rng, 24000, 384, dtype)
    if dtype == np.uint8:
        synth.inject_narrowband_rfi(x, rng, frac=0.05, amp=40.0, n_stuck=2, pmax=512)
    kw = dict(method="subband", sbw=64, ndm0=16, ds=ds, inter_budget=10 * 6 * 2 * 4096,
              use_iqrm=True, iqrm_c0=0, iqrm_c1=383, iqrm_mode=1, iqrm_W=7000 // ds,
              iqrm_H=5000 // ds, use_zero_dm=True, zdm_c0=0, zdm_c1=383, bad_channels=[5, 77])
    whole = make_spec(x, b, 0.0, 1.0, 150, **kw)
    pw = capi.Plan(whole)
    es = np.dtype(dtype).itemsize
    assert pw.info.clean_passes == 1
    budget = 384 * es * (2 * pw.info.nbin // 3)
    pw.close()
    tiled = make_spec(x, b, 0.0, 1.0, 150, clean_budget=budget, **kw)
    pt = capi.Plan(tiled)
    assert pt.info.clean_passes > 2, pt.info.clean_passes
    got, masks = pt.run_host(x, want_masks=True)
    pt.close()
    pw = capi.Plan(whole)
    got_w, masks_w = pw.run_host(x, want_masks=True)
    pw.close()
    assert np.array_equal(masks, masks_w)
    assert np.array_equal(got, got_w)
    ref, rmask, *_ = O.run_pipeline(x, oracle_spec(tiled), O.IQRM_GPU)
    assert np.array_equal(masks, rmask)
    if dtype == np.float32:
        assert_f32_close(got, ref, x, tiled)
    else:
        assert np.array_equal(got, ref)


def test_clean_budget_below_one_super_tile_is_rejected(gpu_lib):
    rng = np.random.default_rng(32)
    b = small_band(256, 1100.0, 1500.0, 1e-4)
    x = noise(rng, 20000, 256)
    spec = make_spec(x, b, 0.0, 1.0, 100, method="subband", sbw=64, ndm0=16,
                     inter_budget=10 * 6 * 2 * 4096, clean_budget=4096)
    with pytest.raises(MemoryError):
        capi.Plan(spec)

"""The `_astroflow_core` drop-in on the GPU: the raw plane and masks of
`_dedisperse_spec_raw` are bit-exact against the oracle for the parameters
the reference derives from the call (src/core/dedispered.cu:1003-1036,
src/core/iqrm.cu:368-369); `_dedisperse_spec` images are the imaging of that
plane; float32 input takes the float overload; `_dedispered_fil` equals the
array call on the flipped data."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import write_sigproc

core = pytest.importorskip("astroflow_amd._astroflow_core")
pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


NCH, NSAMP = 256, 6000
FCH1, FOFF, TSAMP = np.float32(1200.0), np.float32(300.0 / 256), np.float32(1e-4)


def header(ndata=NSAMP, nbits=8):
    return core.Header(60000.0, float(TSAMP), float(FCH1), float(FOFF), nbits, NCH, ndata, "synthetic")


def data_u8(seed=3):
    rng = np.random.default_rng(seed)
    x = np.clip(np.rint(rng.normal(128, 16, (NSAMP, NCH))), 0, 255).astype(np.uint8)
    x[:, 17] = np.clip(128 + 60 * np.sin(np.arange(NSAMP) * 0.9), 0, 255).astype(np.uint8)
    x[:, 200] = 255
    return x


def oracle_for_call(x, dm_low, dm_high, fs, fe, dm_step, ds, rfi, bad):
    """The oracle, set up the way the reference derives a call's parameters."""
    freq = O.freq_table_f32(float(FCH1), float(FOFF), NCH)
    cs, ce = O.chan_range(freq, fs, fe)
    nd = O.dm_steps(dm_low, dm_high, dm_step)
    tb = float(np.float32(TSAMP * np.float32(ds)))
    down = -(-NSAMP // ds)
    W = int(np.floor(rfi.iqrm_cfg.win_sec / tb)) if rfi.iqrm_cfg.win_sec > 0 else 0
    H = int(np.floor(rfi.iqrm_cfg.hop_sec / tb)) if rfi.iqrm_cfg.hop_sec > 0 else W
    spec = O.PlanSpec(
        nchan=NCH, freq=freq, tsamp=float(TSAMP), dm_low=float(np.float32(dm_low)),
        dm_step=float(np.float32(dm_step)), ndm=nd, dd_c0=cs, dd_c1=ce + 1, f_ref=freq[ce], ds=ds,
        method="subband", sbw=32, ndm0=32, use_iqrm=rfi.use_iqrm, iqrm_c0=cs, iqrm_c1=ce,
        iqrm_mode=rfi.iqrm_cfg.mode, iqrm_radius_frac=rfi.iqrm_cfg.radius_frac,
        iqrm_nsigma=rfi.iqrm_cfg.nsigma, iqrm_geofactor=rfi.iqrm_cfg.geofactor,
        iqrm_W=0 if W >= down else W, iqrm_H=H, iqrm_include_tail=rfi.iqrm_cfg.include_tail,
        use_zero_dm=rfi.use_zero_dm, zdm_c0=cs, zdm_c1=ce,
        bad_channels=list(bad) if rfi.use_mask else [])
    return O.run_pipeline(x, spec, O.IQRM_GPU)


def test_raw_plane_and_masks_match_oracle(gpu_lib, tmp_path, monkeypatch):
    say("case: _dedisperse_spec_raw u8 256ch ds2, IQRM windows + zero-DM + mask file")
    x = data_u8()
    mask = tmp_path / "mask.txt"
    mask.write_text("3 9\n77\n")
    q = core.IQRMconfig(1, 0.03, 3.0, 1.5, 0.12, 0.06, True)
    rfi = core.RFIconfig(True, True, True, q)
    args = (0.0, 120.0, 1210.0, 1490.0, 1.0, 2)
    plane, masks = core._dedisperse_spec_raw(x, header(), *args, 0, str(mask), rfi, True)
    ref, rmask, *_ = oracle_for_call(x, *args, rfi, [3, 9, 77])
    assert plane.dtype == np.uint32 and plane.shape == ref.shape
    assert np.array_equal(masks, rmask)
    assert rmask.any()
    assert np.array_equal(plane, ref)


def test_images_are_the_imaging_of_the_plane(gpu_lib):
    say("case: _dedisperse_spec images vs _image_slice of the raw plane")
    x = data_u8(4)
    rfi = core.RFIconfig(False, True, False, core.IQRMconfig())
    args = (0.0, 60.0, 1200.0, 1498.0, 0.5, 1)
    plane = core._dedisperse_spec_raw(x, header(), *args, 0, "no-such-mask.txt", rfi)
    res = core._dedisperse_spec(x, header(), *args, 0.25, 0, "no-such-mask.txt", rfi)
    sps = int(np.float32(0.25) / (TSAMP * np.float32(1)))
    n = -(-plane.shape[1] // sps)
    assert len(res.dm_times) == n and res.shape == [512, 512, 3]
    assert res.tsample == pytest.approx(0.25) and res.dm_low == 0.0
    for k in (0, n - 1):
        sl = np.ascontiguousarray(plane[:, k * sps:(k + 1) * sps].astype(np.float32))
        assert np.array_equal(res.dm_times[k], core._image_slice(sl, 512)), k


def test_float32_input_uses_the_float_overload(gpu_lib):
    say("case: float32 input through _dedisperse_spec_raw")
    rng = np.random.default_rng(5)
    x = rng.standard_normal((NSAMP, NCH)).astype(np.float32)
    rfi = core.RFIconfig()
    args = (0.0, 40.0, 1200.0, 1498.0, 1.0, 1)
    plane = core._dedisperse_spec_raw(x, header(nbits=32), *args, 0, "none", rfi)
    ref, *_ = oracle_for_call(x, *args, rfi, [])
    assert plane.dtype == np.float32
    bound, *_ = oracle_for_call(np.abs(x) + 1.0, *args, rfi, [])
    assert np.all(np.abs(plane.astype(np.float64) - ref) <= 1e-6 * bound)


def test_filterbank_call_equals_array_call(gpu_lib, tmp_path):
    say("case: _dedispered_fil on a descending-band file vs the array call")
    x = data_u8(6)
    p = tmp_path / "beam.fil"
    # on disk: highest frequency first
    write_sigproc(str(p), x[:, ::-1], fch1=float(FCH1 + FOFF * (NCH - 1)), foff=-float(FOFF),
                  tsamp=float(TSAMP))
    rfi = core.RFIconfig(False, True, True, core.IQRMconfig(1, 0.03, 3.0, 1.5, 0.0, 0.0, True))
    a = core._dedispered_fil(str(p), 0.0, 80.0, 1200.0, 1498.0, 1.0, 1, 0.3, 1, 0, "none", rfi)
    f = core.Filterbank(str(p))
    b = core._dedisperse_spec(f.data.reshape(f.ndata, f.nchans), f.header, 0.0, 80.0, 1200.0,
                              1498.0, 1.0, 1, 0.3, 0, "none", rfi)
    assert len(a.dm_times) == len(b.dm_times) > 0
    for u, v in zip(a.dm_times, b.dm_times):
        assert np.array_equal(u, v)


def test_out_of_band_range_is_value_error(gpu_lib):
    with pytest.raises(ValueError, match="out of spectrum range"):
        core._dedisperse_spec_raw(data_u8(), header(), 0.0, 10.0, 1100.0, 1400.0)


def spectrum_oracle(x, tstart, tend, dm, fs, fe, rfi, bad):
    """cpucal::dedisperse_spec_with_dm with the oracle's pieces (FP64 delays)."""
    freq = O.freq_table_f32(float(FCH1), float(FOFF), NCH)
    ff = freq.astype(np.float32)
    t0 = int(np.float32(tstart) / TSAMP)
    t1 = int(np.float32(tend) / TSAMP)
    f0, fN = ff[0], ff[-1]

    def idx(f):
        raw = float(np.float32(f) - f0) / float(fN - f0) * (NCH - 1)
        raw = min(max(raw, 0.0), NCH - 1)
        return int(np.floor(raw + 1e-9))
    cs, ce = sorted((idx(fs), idx(fe)))
    cx = min(ce + 1, NCH)
    d = O.delay_table(freq, cs, cx, max(freq[cs], freq[cx - 1]), float(TSAMP), 1,
                      np.array([float(np.float32(dm))]))[0]
    dmax = int(d.max())
    tl = min(t1 - t0, NSAMP - t0 - dmax)
    r0 = max(t0 - dmax, 0)
    off = t0 - r0
    slen = min(tl + off + 2 * dmax, NSAMP - r0)
    sl = x[r0:r0 + slen].copy()
    if rfi.use_iqrm:
        W = int(np.floor(rfi.iqrm_cfg.win_sec / float(TSAMP))) if rfi.iqrm_cfg.win_sec > 0 else 0
        H = int(np.floor(rfi.iqrm_cfg.hop_sec / float(TSAMP))) if rfi.iqrm_cfg.hop_sec > 0 else W
        m, a0, a1 = O.iqrm(sl, cs, cx, rfi.iqrm_cfg.mode, rfi.iqrm_cfg.radius_frac,
                           rfi.iqrm_cfg.nsigma, rfi.iqrm_cfg.geofactor, 0 if W >= slen else W, H,
                           rfi.iqrm_cfg.include_tail, O.IQRM_GPU)
        O.apply_window_masks(sl, m, a0, a1)
    if rfi.use_mask and bad:
        O.static_mask(sl, bad)
    out = np.empty((tl, cx - cs), x.dtype)
    for c in range(cs, cx):
        out[:, c - cs] = sl[off + d[c - cs]: off + d[c - cs] + tl, c]
    return out, cs, cx


def test_spectrum_at_dm_matches_oracle(gpu_lib, tmp_path):
    say("case: _dedisperse_spec_with_dm u8, IQRM + mask file, DM 80")
    x = data_u8(8)
    mask = tmp_path / "mask.txt"
    mask.write_text("20 21 22")
    rfi = core.RFIconfig(True, False, True, core.IQRMconfig(1, 0.03, 3.0, 1.5, 0.0, 0.0, True))
    args = (0.05, 0.45, 80.0, 1230.0, 1480.0)
    sp = core._dedisperse_spec_with_dm(x, header(), *args, str(mask), rfi)
    ref, cs, cx = spectrum_oracle(x, *args, rfi, [20, 21, 22])
    assert sp.data.shape == ref.shape and sp.data.dtype == np.uint8
    assert (sp.ntimes, sp.nchans) == ref.shape
    assert np.array_equal(sp.data, ref)
    # the input array is left as it was
    assert np.array_equal(x, data_u8(8))


def test_spectrum_at_dm_zero_is_the_input(gpu_lib):
    # python/tests/test_dedispered.py:94-101: at DM 0 without RFI the
    # spectrum is the input's selected channels
    say("case: _dedisperse_spec_with_dm at DM 0")
    x = data_u8(9)
    sp = core._dedisperse_spec_with_dm(x, header(), 0.0, float(NSAMP * TSAMP), 0.0, 1210.0,
                                       1490.0, "none", core.RFIconfig())
    _, cs, cx = spectrum_oracle(x, 0.0, float(NSAMP * TSAMP), 0.0, 1210.0, 1490.0,
                                core.RFIconfig(), [])
    assert sp.data.shape == (NSAMP, cx - cs)
    assert np.array_equal(sp.data, x[:, cs:cx])


def test_repeated_calls_reuse_the_plan(gpu_lib):
    say("case: plan cache of the module: repeated and new parameter sets")
    x = data_u8()
    q = core.IQRMconfig(1, 0.03, 3.0, 1.5, 0.12, 0.06, True)
    rfi = core.RFIconfig(False, True, True, q)
    core._plan_cache_clear()
    assert core._plan_cache_size() == 0
    args = (0.0, 120.0, 1210.0, 1490.0, 1.0, 2)
    a, ma = core._dedisperse_spec_raw(x, header(), *args, 0, "", rfi, True)
    assert core._plan_cache_size() == 1
    b, mb = core._dedisperse_spec_raw(x, header(), *args, 0, "", rfi, True)
    assert core._plan_cache_size() == 1
    assert np.array_equal(a, b) and np.array_equal(ma, mb)
    ref, rmask, *_ = oracle_for_call(x, *args, rfi, [])
    assert np.array_equal(b, ref) and np.array_equal(mb, rmask)
    # other parameters: another plan, the oldest dropped beyond two
    c = core._dedisperse_spec_raw(x, header(), 5.0, 60.0, 1210.0, 1490.0, 1.0, 1, 0, "", rfi)
    assert core._plan_cache_size() == 2
    core._dedisperse_spec_raw(x, header(), 0.0, 30.0, 1210.0, 1490.0, 1.0, 1, 0, "", rfi)
    assert core._plan_cache_size() == 2
    ref_c, *_ = oracle_for_call(x, 5.0, 60.0, 1210.0, 1490.0, 1.0, 1, rfi, [])
    assert np.array_equal(c, ref_c)
    # a failing call keeps no plan
    with pytest.raises(ValueError):
        core._dedisperse_spec_raw(x, header(), 0.0, 30.0, 900.0, 1000.0, 1.0, 1, 0, "", rfi)
    assert core._plan_cache_size() == 2
    core._plan_cache_clear()
    assert core._plan_cache_size() == 0

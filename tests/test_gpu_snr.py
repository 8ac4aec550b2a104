"""Boxcar S/N on the device (af_frb_snr, af_snr.hip) against the host module
(`_calculate_frb_snr`, af_snr.h, itself pinned to closed forms and to the
Python oracle oracle/snr.py in test_snr_cpu.py): the whole result tuple is
equal, value for value.

  - a rectangular pulse on an alternating baseline (the closed form of
    test_snr_cpu.py) through the device path;
  - the 40 seeded spectra of test_snr_cpu.py (TOA, fitting window,
    decimation, noise ranges, unusable noise ranges);
  - long series (2^16 samples x 64 channels, 2^18 x 4) with a pulse;
  - a u8 spectrum from _dedisperse_spec_with_dm, summed as float32 values;
  - a series with a NaN (searched on the host) and the empty spectrum."""
import math
import sys

import numpy as np
import pytest

from astroflow_amd import _astroflow_core as core
from astroflow_amd import capi

pytestmark = pytest.mark.gpu

CODES = {np.dtype(np.uint8): capi.AF_U8, np.dtype(np.uint16): capi.AF_U16,
         np.dtype(np.uint32): capi.AF_U32, np.dtype(np.float32): capi.AF_F32}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def device_snr(x: np.ndarray, **kw):
    x = np.ascontiguousarray(x)
    nt, nf = x.shape
    d = capi.DeviceBuffer(max(1, x.nbytes))
    try:
        if x.nbytes:
            d.upload(x)
        return capi.frb_snr(d.ptr, CODES[x.dtype], nt, nf, **kw)
    finally:
        d.free()


def test_rectangular_pulse_closed_form(gpu_lib):
    say("case: S/N closed form, width-8 pulse on +-1 baseline")
    x = np.where(np.arange(2000) % 2 == 0, 1.0, -1.0).astype(np.float32)[:, None]
    x[1000:1008] += 50
    snr, width, peak, (mu, sd, info) = device_snr(x)
    m = -1.0 / 1967
    s = math.sqrt(1.0 - m * m)
    assert (width, peak) == (8, 1004)
    assert mu == pytest.approx(m, rel=1e-12) and sd == pytest.approx(s, rel=1e-12)
    assert snr == pytest.approx((400 - 8 * m) / (s * math.sqrt(8)), rel=1e-12)
    assert info["method"] == "boxcar" and info["fit_converged"]


def test_seeded_spectra_equal_the_host_module(gpu_lib):
    say("case: S/N device vs host on 40 seeded spectra")
    rng = np.random.default_rng(21)
    for k in range(40):
        nt = int(rng.integers(2, 700))
        nf = int(rng.integers(1, 9))
        x = rng.normal(0, 1, (nt, nf)).astype(np.float32)
        if nt > 40:
            a = int(rng.integers(0, nt - 20))
            x[a:a + int(rng.integers(1, 20))] += float(rng.uniform(0, 6))
        kw = {}
        if k % 3 == 1:
            kw["toa_sample_idx"] = int(rng.integers(-5, nt + 5))
        if k % 4 == 2:
            kw["fitting_window_samples"] = int(rng.integers(1, 200))
        if k % 5 == 3:
            kw["tsamp"] = 64e-6
            kw["target_time_us"] = float(rng.choice([64.0, 128.0, 256.0, 1000.0]))
        if k % 6 == 4:
            kw["noise_range"] = [(0, int(rng.integers(1, nt + 1))), (int(nt * 0.8), nt + 10)]
        if k % 7 == 5:
            kw["noise_range"] = [(5,), (nt, nt + 3)]
        kw["threshold_sigma"] = float(rng.choice([3.0, 5.0]))
        assert device_snr(x, **kw) == core._calculate_frb_snr(x, **kw), (k, kw)


@pytest.mark.parametrize("nt,nf", [(1 << 16, 64), (1 << 18, 4)])
def test_long_series_equal_the_host_module(gpu_lib, nt, nf):
    say(f"case: S/N device vs host, {nt} x {nf}")
    rng = np.random.default_rng(nt + nf)
    x = rng.normal(0, 1, (nt, nf)).astype(np.float32)
    x[nt // 3: nt // 3 + 37] += 1.5
    for kw in ({}, {"toa_sample_idx": nt // 3 + 10, "fitting_window_samples": 4000},
               {"tsamp": 64e-6, "target_time_us": 256.0}):
        got = device_snr(x, **kw)
        assert got == core._calculate_frb_snr(x, **kw), kw
        assert got[0] > 10


def test_u8_spectrum_from_the_dm_call(gpu_lib):
    say("case: S/N device vs host on a u8 spectrum at one DM")
    rng = np.random.default_rng(4)
    nsamp, nch = 4096, 256
    x = rng.integers(90, 160, (nsamp, nch)).astype(np.uint8)
    x[2000:2006] += 60
    h = core.Header(60000.0, 64e-6, 1200.0, 1.0, 8, nch, nsamp, "synthetic")
    sp = core._dedisperse_spec_with_dm(x, h, 0.0, nsamp * 64e-6, 0.0, 1200.0, 1455.0, "none",
                                       core.RFIconfig())
    spec = np.ascontiguousarray(sp.data)
    assert spec.dtype == np.uint8
    got = device_snr(spec, toa_sample_idx=2002)
    assert got == core._calculate_frb_snr(spec.astype(np.float32), toa_sample_idx=2002)
    assert got[2] == pytest.approx(2002, abs=4)


def test_non_finite_series_and_empty_spectrum(gpu_lib):
    say("case: S/N with a NaN sample, and an empty spectrum")
    rng = np.random.default_rng(3)
    x = rng.normal(0, 1, (500, 3)).astype(np.float32)
    x[250, 1] = np.nan
    assert repr(device_snr(x)) == repr(core._calculate_frb_snr(x))
    e = np.zeros((0, 4), np.float32)
    assert device_snr(e) == core._calculate_frb_snr(e)

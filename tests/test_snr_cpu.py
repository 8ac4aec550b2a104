"""Boxcar S/N (`_calculate_frb_snr`, src/bindpy/frb_snr.cpp:144-398): host C++
in the drop-in module, checked against closed forms and the Python oracle.
No GPU."""
import math

import numpy as np
import pytest

from astroflow_amd import _astroflow_core as core
from oracle import snr as S


def alternating(n, nf=1):
    """+1 at even samples, -1 at odd ones (summed over nf equal channels)."""
    x = np.where(np.arange(n) % 2 == 0, 1.0, -1.0).astype(np.float32) / nf
    return np.repeat(x[:, None], nf, axis=1)


# ---------------------------------------------------- pins of the oracle --

def test_width_list_closed_form():
    assert S.widths(0) == [1] and S.widths(1) == [1]
    assert S.widths(9) == list(range(1, 10))
    w = S.widths(512)
    assert w[:16] == list(range(1, 17)) and w[16] == 17 and w[-1] == 512
    assert all(b > a for a, b in zip(w, w[1:])) and len(w) <= 48
    # 32 log-spaced values from 17 to 40 round to every integer in [17, 40]
    assert S.widths(40) == list(range(1, 41))


def test_clipped_statistics_hand_sets():
    assert S.robust([], 5.0) == (0.0, 1.0)
    assert S.robust([float("nan"), 3.0], 5.0) == (3.0, 1.0)   # one value: spread 0 -> 1
    # 20 values of +-1 and two far outliers: the outliers are clipped, the
    # rest has mean 0 and std 1 exactly
    v = [1.0, -1.0] * 10 + [100.0, -90.0]
    assert S.robust(v, 5.0) == (0.0, 1.0)
    # a clip that would keep fewer than 10 samples is not applied
    w = [0.0] * 6 + [1.0, 1.0, 50.0]
    mu, sd = S.robust(w, 5.0)
    assert mu == pytest.approx(sum(w) / 9)


@pytest.mark.parametrize("impl", ["oracle", "module"])
def test_rectangular_pulse_on_alternating_baseline(impl):
    # 2000 samples of +-1 and a pulse of amplitude 50 over samples
    # [1000, 1008).  The best boxcar is that pulse (width 8, centre 1004,
    # sum 400).  The noise is re-estimated without [988, 1021): 983 samples of
    # +1 and 984 of -1, so mean -1/1967 and std sqrt(1 - mean^2).
    x = alternating(2000)
    x[1000:1008] += 50
    fn = S.frb_snr if impl == "oracle" else core._calculate_frb_snr
    arg = x.tolist() if impl == "oracle" else x
    snr, width, peak, (mu, sd, info) = fn(arg)
    m = -1.0 / 1967
    s = math.sqrt(1.0 - m * m)
    assert (width, peak) == (8, 1004)
    assert mu == pytest.approx(m, rel=1e-12) and sd == pytest.approx(s, rel=1e-12)
    assert snr == pytest.approx((400 - 8 * m) / (s * math.sqrt(8)), rel=1e-12)
    assert info == {"fit_converged": True, "method": "boxcar", "width_samples": 8,
                    "search_start": 0, "search_end": 2000, "downsample": 1}


@pytest.mark.parametrize("impl", ["oracle", "module"])
def test_toa_restricts_the_search(impl):
    # a strong pulse at 300 and a weaker one at 1500; a TOA at 1500 searches
    # 1500 +- max(fitting window, 100) = 1500 +- min(2000 / 4, 512) = [1000,
    # 2000) only and finds the weaker one
    x = alternating(2000)
    x[300:304] += 80
    x[1500:1504] += 40
    fn = S.frb_snr if impl == "oracle" else core._calculate_frb_snr
    arg = x.tolist() if impl == "oracle" else x
    snr, width, peak, (_, _, info) = fn(arg, None, 5.0, 1500)
    assert (width, peak) == (4, 1502)
    assert (info["search_start"], info["search_end"]) == (1000, 2000)
    snr0, _, peak0, _ = fn(arg)
    assert peak0 == 302 and snr0 > snr > 50


@pytest.mark.parametrize("impl", ["oracle", "module"])
def test_empty_and_fallback_returns(impl):
    fn = S.frb_snr if impl == "oracle" else core._calculate_frb_snr
    empty = np.zeros((0, 4), np.float32)
    r = fn(empty.tolist() if impl == "oracle" else empty)
    assert r == (-1.0, 0, 0, (0.0, 1.0, {"fit_converged": False}))
    # one sample: no boxcar narrower than the series -> fallback on the peak
    one = np.full((1, 3), 2.0, np.float32)
    r = fn(one.tolist() if impl == "oracle" else one)
    assert r == (0.0, 1, 0, (6.0, 1.0, {"fit_converged": False, "method": "fallback"}))


def test_module_equals_oracle_on_seeded_spectra():
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
            kw["noise_range"] = [(5,), (nt, nt + 3)]       # nothing usable: whole series
        kw["threshold_sigma"] = float(rng.choice([3.0, 5.0]))
        got = core._calculate_frb_snr(x, **kw)
        want = S.frb_snr(x.tolist(), **kw)
        assert got == want, (k, kw)


def test_spec_must_be_two_dimensional():
    with pytest.raises(RuntimeError):
        core._calculate_frb_snr(np.zeros(10, np.float32))

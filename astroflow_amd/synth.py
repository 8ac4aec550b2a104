"""Seeded synthetic filterbanks and the workloads they are measured on.

Frequency convention: the filterbank is recorded highest frequency first
(fch1 = top, foff < 0) and flipped to ascending at ingest, as the reference's
readers do (src/data/filterbank.cpp:793-842, python/astroflow/io/filterbank.py:117-121).
The resulting header has fch1 = lowest channel and foff > 0, and the channel
frequencies are formed in float32 from those header fields
(src/core/dedispered.cu:1003-1005).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

KDM = 4148.808
TSAMP = np.float32(49.152e-6)


def ascending_header(f_lo_mhz: float, f_hi_mhz: float, nchan: int) -> Tuple[np.float32, np.float32]:
    """(fch1, foff) after the flip for a band [f_lo, f_hi] recorded top first
    with nchan channels of width (f_hi - f_lo)/nchan."""
    foff_top = -(f_hi_mhz - f_lo_mhz) / nchan
    fch1 = f_hi_mhz + (nchan - 1) * foff_top     # lowest channel (float64, as Python does)
    return np.float32(fch1), np.float32(-foff_top)


def freq_table(fch1: np.float32, foff: np.float32, nchan: int) -> np.ndarray:
    """fch1 + i*foff in float32 arithmetic, returned as float64."""
    i = np.arange(nchan, dtype=np.float32)
    return (np.float32(fch1) + i * np.float32(foff)).astype(np.float64)


def delay_samples(freq: np.ndarray, f_ref: float, dm: float, tsamp: float) -> np.ndarray:
    inv = 1.0 / (freq * freq) - 1.0 / (f_ref * f_ref)
    return np.round(KDM * dm * inv / tsamp)


def noise_u8(rng: np.random.Generator, nsamp: int, nchan: int, mean=128.0, sigma=16.0) -> np.ndarray:
    x = rng.normal(mean, sigma, size=(nsamp, nchan))
    return np.clip(np.round(x), 0, 255).astype(np.uint8)


def inject_pulse(x: np.ndarray, freq: np.ndarray, f_ref: float, tsamp: float, dm: float,
                 t0: int, fwhm_s: float, amp: float):
    """Add a dispersed Gaussian pulse (arrival at t0 at f_ref) to x in place."""
    sigma = fwhm_s / tsamp / (2.0 * math.sqrt(2.0 * math.log(2.0)))
    half = int(math.ceil(4 * sigma))
    d = delay_samples(freq, f_ref, dm, tsamp).astype(np.int64)
    offs = np.arange(-half, half + 1)
    prof = amp * np.exp(-0.5 * (offs / sigma) ** 2)
    n, nc = x.shape
    for c in range(nc):
        t = t0 + d[c] + offs
        ok = (t >= 0) & (t < n)
        if not ok.any():
            continue
        col = x[t[ok], c].astype(np.float64) + prof[ok]
        if x.dtype == np.float32:
            x[t[ok], c] = col.astype(np.float32)
        else:
            x[t[ok], c] = np.clip(np.round(col), 0, np.iinfo(x.dtype).max).astype(x.dtype)


def inject_narrowband_rfi(x: np.ndarray, rng: np.random.Generator, frac=0.05, amp=40.0,
                          n_stuck=10, pmin=8.0, pmax=4096.0) -> List[int]:
    n, nc = x.shape
    k = max(1, int(round(frac * nc)))
    chans = rng.choice(nc, size=k + n_stuck, replace=False)
    t = np.arange(n, dtype=np.float64)
    for c in chans[:k]:
        period = math.exp(rng.uniform(math.log(pmin), math.log(pmax)))
        phase = rng.uniform(0, 2 * math.pi)
        col = x[:, c].astype(np.float64) + amp * np.sin(2 * math.pi * t / period + phase)
        x[:, c] = np.clip(np.round(col), 0, 255).astype(x.dtype)
    for c in chans[k:]:
        x[:, c] = 255
    return sorted(int(c) for c in chans)


@dataclass
class Band:
    nchan: int
    fch1: np.float32
    foff: np.float32
    tsamp: np.float32

    @property
    def freq(self):
        return freq_table(self.fch1, self.foff, self.nchan)

    @property
    def f_top(self):
        return float(self.freq[-1])


def band_1000_1500(nchan: int) -> Band:
    fch1, foff = ascending_header(1000.0, 1500.0, nchan)
    return Band(nchan, fch1, foff, TSAMP)


def max_delay(band: Band, dm_max: float) -> int:
    f = band.freq
    return int(delay_samples(f[:1], band.f_top, dm_max, float(band.tsamp))[0])


def noise_u8_threaded(seed: int, nsamp: int, nchan: int, mean=128.0, sigma=16.0,
                      threads: int = 8, rows_per_block: int = 8192) -> np.ndarray:
    """clip(round(N(mean, sigma)), 0, 255) as uint8, generated in fixed blocks of
    rows from spawned seeds (the result does not depend on the thread count)."""
    from concurrent.futures import ThreadPoolExecutor

    out = np.empty((nsamp, nchan), np.uint8)
    nblk = (nsamp + rows_per_block - 1) // rows_per_block
    seeds = np.random.SeedSequence(seed).spawn(nblk)

    def work(b):
        r0 = b * rows_per_block
        r1 = min(nsamp, r0 + rows_per_block)
        g = np.random.default_rng(seeds[b])
        v = g.standard_normal((r1 - r0, nchan), dtype=np.float32)
        v *= np.float32(sigma)
        v += np.float32(mean)
        np.rint(v, out=v)
        np.clip(v, 0, 255, out=v)
        out[r0:r1] = v.astype(np.uint8)

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        list(ex.map(work, range(nblk)))
    return out


def chunk_4096(seed: int = 2, threads: int = 8):
    """The 4096-channel u8 chunk of the bench: 0.5 s of output (10173 samples)
    plus the DM 0-3000 overlap tail, with narrowband RFI (5% of channels,
    amplitude 40, periods 8-4096 samples), 10 channels stuck at 255 and pulses
    at DM 150/1200/2800.  Returns (x [nsamp][4096] u8, band, nout, dmax)."""
    band = band_1000_1500(4096)
    tsamp = float(band.tsamp)
    nout = int(math.ceil(0.5 / tsamp))
    dmax = max_delay(band, 3000.0)
    nsamp = nout + dmax
    x = noise_u8_threaded(seed, nsamp, 4096, threads=threads)
    rng = np.random.default_rng(seed + 1)
    inject_narrowband_rfi(x, rng, frac=0.05, amp=40.0, n_stuck=10, pmin=8.0, pmax=4096.0)
    f = band.freq
    for dm, t0 in ((150.0, 2000), (1200.0, 5000), (2800.0, 8000)):
        inject_pulse(x, f, band.f_top, tsamp, dm, t0, 1e-3, 48.0)
    return x, band, nout, dmax


def direct_1024(seed: int = 1, threads: int = 8):
    """The plain direct-method case: 1024 channels over 1000-1500 MHz (top
    first, flipped to ascending), tsamp 49.152 us, 2^18 samples of u8
    N(128,16) noise and one dispersed Gaussian pulse (DM 500, FWHM 1 ms,
    +3 sigma per channel).  Returns (x [2^18][1024] u8, band, pulse t0)."""
    band = band_1000_1500(1024)
    nsamp = 1 << 18
    x = noise_u8_threaded(seed, nsamp, 1024, threads=threads)
    t0 = 60000
    inject_pulse(x, band.freq, band.f_top, float(band.tsamp), 500.0, t0, 1e-3, 48.0)
    return x, band, t0


# The three-step DM plan of the large float32 plane: (dm_low, dm_step, ndm,
# time downsample) -- DM 0-1000 step 0.5 at full rate, 1000-4000 step 1 at 2x
# downsampling, 4000-10000 step 2 at 4x.
C5_STEPS = ((0.0, 0.5, 2001, 1), (1000.0, 1.0, 3001, 2), (4000.0, 2.0, 3001, 4))


def inject_impulsive_rfi(x: np.ndarray, rng: np.random.Generator, frac=0.02, smin=5.0,
                         smax=20.0, lmin=1, lmax=50):
    """Broadband bursts (every channel) of 1-50 samples at 5-20 sigma, placed
    until about `frac` of the time samples are hit.  Returns the burst list."""
    n = x.shape[0]
    hit = 0
    bursts = []
    while hit < frac * n:
        ln = int(rng.integers(lmin, lmax + 1))
        t0 = int(rng.integers(0, max(1, n - ln)))
        amp = float(rng.uniform(smin, smax))
        x[t0:t0 + ln] += np.float32(amp)
        bursts.append((t0, ln, amp))
        hit += ln
    return bursts


def c3_chunks(seed: int = 4, n: int = 48):
    """48 independent beam-chunks of the batched workload: nchan in {1024,
    2048, 4096}, nsamp log-uniform in [2^14, 2^20], dtype uint8 or float32,
    DM upper limit in {500, 2000, 5000} with step 1 or 0.5, time downsample
    in {1, 2, 4}.  Returns a list of dicts (sizes only; data is filled on
    the device)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        out.append(dict(
            nchan=int(rng.choice([1024, 2048, 4096])),
            nsamp=int(round(2.0 ** rng.uniform(14, 20))),
            dtype=np.uint8 if rng.random() < 0.5 else np.float32,
            dm_max=float(rng.choice([500.0, 2000.0, 5000.0])),
            dm_step=float(rng.choice([1.0, 0.5])),
            ds=int(rng.choice([1, 2, 4]))))
    return out


def noise_f32_threaded(seed: int, nsamp: int, nchan: int, threads: int = 8,
                       rows_per_block: int = 8192) -> np.ndarray:
    """N(0,1) float32, generated in fixed blocks of rows from spawned seeds
    (the result does not depend on the thread count)."""
    from concurrent.futures import ThreadPoolExecutor

    out = np.empty((nsamp, nchan), np.float32)
    nblk = (nsamp + rows_per_block - 1) // rows_per_block
    seeds = np.random.SeedSequence(seed).spawn(nblk)

    def work(b):
        r0 = b * rows_per_block
        r1 = min(nsamp, r0 + rows_per_block)
        np.random.default_rng(seeds[b]).standard_normal((r1 - r0, nchan), dtype=np.float32,
                                                         out=out[r0:r1])

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        list(ex.map(work, range(nblk)))
    return out


def _normal_lut(dtype) -> np.ndarray:
    """2^16 equal-probability levels of N(0,1) (the midpoint quantiles), as
    float32, or as uint8 N(128,16) rounded and clipped to 0..255."""
    from scipy.special import ndtri
    q = ndtri((np.arange(65536, dtype=np.float64) + 0.5) / 65536.0)
    if dtype == np.uint8:
        return np.clip(np.rint(128.0 + 16.0 * q), 0, 255).astype(np.uint8)
    return q.astype(np.float32)


def noise_lut_threaded(seed: int, nsamp: int, nchan: int, dtype, threads: int = 8,
                       rows_per_block: int = 8192) -> np.ndarray:
    """Normal noise by table look-up of uniform 16-bit draws (float32 N(0,1),
    or uint8 N(128,16)): several times faster than drawing normals, in fixed
    blocks of rows from spawned seeds (independent of the thread count)."""
    from concurrent.futures import ThreadPoolExecutor

    lut = _normal_lut(dtype)
    out = np.empty((nsamp, nchan), dtype)
    nblk = (nsamp + rows_per_block - 1) // rows_per_block
    seeds = np.random.SeedSequence(seed).spawn(nblk)

    def work(b):
        r0 = b * rows_per_block
        r1 = min(nsamp, r0 + rows_per_block)
        u = np.random.default_rng(seeds[b]).integers(0, 65536, size=(r1 - r0, nchan),
                                                     dtype=np.uint16)
        np.take(lut, u, out=out[r0:r1])

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        list(ex.map(work, range(nblk)))
    return out


def c3_chunk_data(c: dict, seed: int, threads: int = 8) -> np.ndarray:
    """Seeded input of one batched beam-chunk (a dict from c3_chunks): uint8
    N(128,16) or float32 N(0,1) noise (noise_lut_threaded), plus narrowband RFI on 5% of the
    channels (sinusoids of amplitude 40 for uint8 and 2.5 sigma for float32,
    periods log-uniform in 8-4096 samples) and 3 channels stuck at the
    maximum (255 / 8.0), so that IQRM has channels to flag."""
    nsamp, nchan = c["nsamp"], c["nchan"]
    x = noise_lut_threaded(seed, nsamp, nchan, c["dtype"], threads=threads)
    if c["dtype"] == np.uint8:
        amp, top = 40.0, 255.0
    else:
        amp, top = 2.5, 8.0
    rng = np.random.default_rng(seed + 7919)
    k = max(1, int(round(0.05 * nchan)))
    chans = rng.choice(nchan, size=k + 3, replace=False)
    periods = np.exp(rng.uniform(math.log(8.0), math.log(4096.0), size=k))
    phases = rng.uniform(0, 2 * math.pi, size=k)
    from concurrent.futures import ThreadPoolExecutor
    blk = 16384
    sel = chans[:k]

    def work(b):
        r0 = b * blk
        r1 = min(nsamp, r0 + blk)
        t = np.arange(r0, r1, dtype=np.float64)[:, None]
        v = x[r0:r1, sel].astype(np.float64)
        v += amp * np.sin(2 * math.pi * t / periods[None, :] + phases[None, :])
        if x.dtype == np.uint8:
            x[r0:r1, sel] = np.clip(np.round(v), 0, 255).astype(np.uint8)
        else:
            x[r0:r1, sel] = v.astype(np.float32)
        x[r0:r1, chans[k:]] = top

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        list(ex.map(work, range((nsamp + blk - 1) // blk)))
    return x

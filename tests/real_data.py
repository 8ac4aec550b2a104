"""The reference's own test fixture, tests/FRB180417.fil (kept unchanged under
tests/golden/): 336 channels of 1 MHz from 1130 to 1465 MHz, 8-bit, 5120
samples of 1.266 ms, holding the burst FRB 180417 (published DM 474.8
pc cm^-3)."""
import os

import numpy as np

FIL = os.path.join(os.path.dirname(__file__), "golden", "FRB180417.fil")
FRB_DM = 474.8


def load():
    """(samples [5120][336] u8, ascending float32 frequency table as float64, tsamp, header)."""
    from astroflow_amd import _astroflow_core as core
    f = core.Filterbank(FIL)
    x = np.ascontiguousarray(f.data.reshape(f.ndata, f.nchans))
    freq = (np.float32(f.fch1) + np.arange(f.nchans, dtype=np.float32) * np.float32(f.foff))
    return x, freq.astype(np.float64), float(np.float32(f.tsamp)), f


def oracle_spec(freq, tsamp, method, ndm=1001):
    """DM 0-1000 step 1 over all channels, IQRM (std, radius 0.1) and zero-DM."""
    from oracle import oracle as O
    n = len(freq)
    return O.PlanSpec(nchan=n, freq=freq, tsamp=tsamp, dm_low=0.0, dm_step=1.0, ndm=ndm, dd_c0=0,
                      dd_c1=n, f_ref=float(freq[-1]), method=method, sbw=16, ndm0=8,
                      use_iqrm=True, iqrm_c0=0, iqrm_c1=n - 1, use_zero_dm=True, zdm_c0=0,
                      zdm_c1=n, bad_channels=[])


def best_dm(plane):
    """(row of the largest peak S/N, that S/N, the peak's sample): per row,
    (max - median) / std over time."""
    p = plane.astype(np.float64)
    med = np.median(p, axis=1, keepdims=True)
    sd = p.std(axis=1, keepdims=True)
    snr = ((p - med) / sd).max(axis=1)
    m = int(np.argmax(snr))
    return m, float(snr[m]), int(np.argmax(p[m]))

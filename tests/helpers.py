"""Shared helpers: build matching product specs and oracle specs, run both."""
from __future__ import annotations

import dataclasses

import numpy as np

from astroflow_amd import capi, synth
from oracle import oracle as O


def oracle_spec(spec: capi.DedispSpec) -> O.PlanSpec:
    fields = {f.name for f in dataclasses.fields(O.PlanSpec)}
    kw = {k: v for k, v in dataclasses.asdict(spec).items() if k in fields}
    kw["freq"] = np.asarray(spec.freq, np.float64)
    return O.PlanSpec(**kw)


def make_spec(x: np.ndarray, band: synth.Band, dm_low, dm_step, ndm, **kw) -> capi.DedispSpec:
    nsamp, nchan = x.shape
    f = band.freq
    base = dict(dtype=x.dtype, nchan=nchan, nsamp=nsamp, freq=f, tsamp=float(band.tsamp),
                dm_low=dm_low, dm_step=dm_step, ndm=ndm, dd_c0=0, dd_c1=nchan,
                f_ref=float(f[-1]))
    base.update(kw)
    return capi.DedispSpec(**base)


def small_band(nchan, f_lo=1000.0, f_hi=1500.0, tsamp=49.152e-6):
    fch1, foff = synth.ascending_header(f_lo, f_hi, nchan)
    return synth.Band(nchan, fch1, foff, np.float32(tsamp))

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


def write_sigproc(path, data: np.ndarray, fch1: float, foff: float, tsamp: float,
                  tstart: float = 60000.0, source="synthetic"):
    """A SIGPROC filterbank file: keyword header, then [nsamp][nchan] samples."""
    import struct

    def s(x):
        b = x.encode()
        return struct.pack("<i", len(b)) + b

    nbits = data.dtype.itemsize * 8
    hdr = s("HEADER_START") + s("source_name") + s(source)
    hdr += s("machine_id") + struct.pack("<i", 0) + s("telescope_id") + struct.pack("<i", 0)
    hdr += s("data_type") + struct.pack("<i", 1)
    hdr += s("fch1") + struct.pack("<d", fch1) + s("foff") + struct.pack("<d", foff)
    hdr += s("nchans") + struct.pack("<i", data.shape[1]) + s("nbits") + struct.pack("<i", nbits)
    hdr += s("tstart") + struct.pack("<d", tstart) + s("tsamp") + struct.pack("<d", tsamp)
    hdr += s("nifs") + struct.pack("<i", 1) + s("HEADER_END")
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(np.ascontiguousarray(data).tobytes())

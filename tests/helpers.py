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


def write_psrfits(path, data: np.ndarray, obsfreq: float, obsbw: float, chan_bw: float,
                  tbin: float, nsblk: int, imjd=60000, smjd=3600, offs=0.25, extra_cols=True):
    """A PSRFITS search-mode file: primary HDU (OBSFREQ, OBSBW, STT_*) and a
    SUBINT binary table whose DATA column holds [NSBLK][NPOL][NCHAN] samples
    per row (data: [nsamp][npol][nchan], nsamp a multiple of nsblk; uint8
    'B', uint16 as 'I' with TZERO 32768, or float32 'E'), big-endian."""
    nsamp, npol, nchan = data.shape
    assert nsamp % nsblk == 0
    nrow = nsamp // nsblk

    def card(k, v):
        if isinstance(v, bool):
            val = ("T" if v else "F").rjust(20)
        elif isinstance(v, str):
            val = ("'" + v.ljust(8) + "'").ljust(20)
        elif isinstance(v, int):
            val = str(v).rjust(20)
        else:
            val = repr(float(v)).upper().rjust(20)
        c = k.ljust(8) + "= " + val
        return c.ljust(80)[:80].encode()

    def block(cards):
        b = b"".join(cards) + b"END".ljust(80)
        return b + b" " * (-len(b) % 2880)

    prim = block([card("SIMPLE", True), card("BITPIX", 8), card("NAXIS", 0), card("EXTEND", True),
                  card("OBSFREQ", obsfreq), card("OBSBW", obsbw), card("STT_IMJD", imjd),
                  card("STT_SMJD", smjd), card("STT_OFFS", offs)])
    if data.dtype == np.uint8:
        form, raw, tz = f"{nsblk * npol * nchan}B", data, None
        nbits = 8
    elif data.dtype == np.uint16:
        form = f"{nsblk * npol * nchan}I"
        raw = (data.astype(np.int32) - 32768).astype(">i2")
        tz, nbits = 32768, 16
    else:
        form, raw, tz = f"{nsblk * npol * nchan}E", data.astype(">f4"), None
        nbits = 32
    cols = []
    if extra_cols:
        cols += [("TSUBINT", "1D", 8), ("DAT_SCL", f"{nchan * npol}E", 4 * nchan * npol)]
    cols.append(("DATA", form, raw.dtype.itemsize * nsblk * npol * nchan))
    rowbytes = sum(c[2] for c in cols)
    cards = [card("XTENSION", "BINTABLE"), card("BITPIX", 8), card("NAXIS", 2),
             card("NAXIS1", rowbytes), card("NAXIS2", nrow), card("PCOUNT", 0),
             card("GCOUNT", 1), card("TFIELDS", len(cols))]
    for i, (name, f, _) in enumerate(cols, 1):
        cards += [card(f"TTYPE{i}", name), card(f"TFORM{i}", f)]
        if name == "DATA" and tz is not None:
            cards.append(card(f"TZERO{i}", tz))
    cards += [card("EXTNAME", "SUBINT"), card("NPOL", npol), card("POL_TYPE", "AABB"),
              card("NCHAN", nchan), card("NSBLK", nsblk), card("NBITS", nbits),
              card("TBIN", tbin), card("CHAN_BW", chan_bw)]
    hdr = block(cards)
    rows = bytearray()
    raw = raw.reshape(nrow, nsblk * npol * nchan)
    for r in range(nrow):
        if extra_cols:
            rows += np.array([tbin * nsblk], ">f8").tobytes()
            rows += np.ones(nchan * npol, ">f4").tobytes()
        rows += raw[r].tobytes()
    rows += b"\0" * (-len(rows) % 2880)
    with open(path, "wb") as f:
        f.write(prim + hdr + bytes(rows))

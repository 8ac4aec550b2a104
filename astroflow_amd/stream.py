"""Served real-time stream driver: beams of 4096-channel u8 filterbank data
whose chunks arrive by a Poisson process, fed to the chunk server.

The reference's real-time use is a loop of `_dedisperse_spec` calls over
chunk + DM-overlap tail read from a file (python/astroflow/search.py:365,464).
Here the beams' windows stay resident in HBM inside `capi.Server`; this module
builds the beams, replays arrivals at a multiple of real time and reports the
server's latency percentiles and the sustained throughput.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import List

import numpy as np

from . import capi, synth


@dataclass
class StreamBeam:
    spec: capi.DedispSpec      # one window: nsamp = chunk + tail rows, nout = chunk rows
    chunk_rows: int
    chunk_s: float
    window_rows: int


def window_spec(band: synth.Band, chunk_rows: int, dm_max: float, dm_step: float,
                sbw: int = 128, ndm0: int = 32) -> capi.DedispSpec:
    """Subband + IQRM (std, radius 5 channels, 3 sigma, one window) + zero-DM
    over one served window of a 4096-channel-style band."""
    nchan = band.nchan
    f = band.freq
    dmax = synth.max_delay(band, dm_max)
    ndm = int(round(dm_max / dm_step)) + 1
    return capi.DedispSpec(
        dtype=np.uint8, nchan=nchan, nsamp=chunk_rows + dmax, freq=f, tsamp=float(band.tsamp),
        dm_low=0.0, dm_step=dm_step, ndm=ndm, dd_c0=0, dd_c1=nchan, f_ref=float(f[-1]),
        ds=1, nout=chunk_rows, method="subband", sbw=sbw, ndm0=ndm0,
        use_iqrm=True, iqrm_c0=0, iqrm_c1=nchan - 1, iqrm_mode=1,
        iqrm_radius_frac=float(np.float32(5.5 / (nchan - 1))), iqrm_nsigma=3.0,
        iqrm_geofactor=1.5, iqrm_W=0, iqrm_H=0, iqrm_include_tail=True,
        use_zero_dm=True, zdm_c0=0, zdm_c1=nchan - 1)


def make_beams(nbeams: int = 19, nchan: int = 4096, dm_max: float = 5000.0,
               dm_step: float = 1.0, durations=(0.25, 0.5, 1.0), seed: int = 5,
               sbw: int = 128) -> List[StreamBeam]:
    """19 beams x 4096 channels over 1000-1500 MHz, tsamp 49.152 us; each
    beam's chunk duration drawn from `durations` (seeded) and rounded up to
    whole 256-sample blocks (0.25/0.5/1.0 s -> 5120/10240/20480 samples),
    the chunk shape the server's incremental RFI path takes."""
    band = synth.band_1000_1500(nchan)
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nbeams):
        dur = float(rng.choice(durations))
        c = int(math.ceil(dur / float(band.tsamp) / 256)) * 256
        dur = c * float(band.tsamp)
        spec = window_spec(band, c, dm_max, dm_step, sbw=sbw)
        out.append(StreamBeam(spec, c, dur, spec.nsamp))
    return out


class RowPool:
    """Page-locked pool of N(128,16) rows; chunk k of beam b reads a seeded
    offset into it.  10% of chunks (seeded) carry a dispersed pulse in their
    own private copy of the rows."""

    def __init__(self, nchan: int, rows: int, seed: int = 5, threads: int = 16,
                 device: bool = False):
        self.nchan = nchan
        self.rows = rows
        self.device = device
        self.buf = capi.HostBuffer((rows, nchan), np.uint8)
        self.buf.array[:] = synth.noise_u8_threaded(seed, rows, nchan, threads=threads)
        self.dbuf = None
        if device:   # rows already in HBM (a capture path writing to the device)
            self.dbuf = capi.DeviceBuffer(rows * nchan)
            self.dbuf.upload(self.buf.array)
        self.extra = []

    def chunk_ptr(self, rng: np.random.Generator, nrows: int, pulse_dm: float = 0.0,
                  band: synth.Band = None) -> int:
        off = int(rng.integers(0, self.rows - nrows + 1))
        base = self.dbuf.ptr if self.device else self.buf.ptr
        if pulse_dm <= 0:
            return base + off * self.nchan
        hb = capi.HostBuffer((nrows, self.nchan), np.uint8)
        hb.array[:] = self.buf.array[off: off + nrows]
        synth.inject_pulse(hb.array, band.freq, band.f_top, float(band.tsamp), pulse_dm,
                           nrows // 2, 1e-3, 48.0)
        if self.device:
            db = capi.DeviceBuffer(hb.array.nbytes)
            db.upload(hb.array)
            hb.free()
            self.extra.append(db)
            return db.ptr
        self.extra.append(hb)
        return hb.ptr

    def free(self):
        for hb in self.extra:
            hb.free()
        self.extra = []
        if self.dbuf is not None:
            self.dbuf.free()
        self.buf.free()


def prefill(srv: capi.Server, beams: List[StreamBeam], pool: RowPool, rng) -> int:
    """Fills every beam's DM-overlap carry (the stream was already running);
    returns the number of chunks submitted per beam."""
    nfill = []
    for b, sb in enumerate(beams):
        n = (sb.window_rows - 1) // sb.chunk_rows   # the next chunk completes the first window
        for k in range(n):
            srv.submit(b, pool.chunk_ptr(rng, sb.chunk_rows), tag=k)
        nfill.append(n)
    srv.drain()
    while srv.poll(256, 0):
        pass
    return nfill


def run_rate(srv: capi.Server, beams: List[StreamBeam], pool: RowPool, rate: float,
             chunks_per_beam: int, seed: int, band: synth.Band) -> dict:
    """Replays Poisson arrivals at `rate` x real time (aggregate over beams;
    rate 0 = all chunks at once).  Results are released as they come back.
    Returns latency percentiles (submit -> plane ready) and throughput."""
    rng = np.random.default_rng(seed)
    per_beam = np.array([1.0 / b.chunk_s for b in beams])
    agg = per_beam.sum() * max(rate, 1e-9)
    n_total = chunks_per_beam * len(beams)
    # each beam contributes chunks_per_beam arrivals, in a seeded random order
    order = np.repeat(np.arange(len(beams)), chunks_per_beam)
    rng.shuffle(order)
    gaps = rng.exponential(1.0 / agg, n_total) if rate > 0 else np.zeros(n_total)
    t_arr = np.cumsum(gaps)
    # chunk rows are prepared before the clock starts (pulse chunks need
    # their own buffers)
    ptrs = []
    for b in order:
        pulse = float(rng.uniform(50, 5000)) if rng.random() < 0.1 else 0.0
        ptrs.append(pool.chunk_ptr(rng, beams[b].chunk_rows, pulse, band))
    base = srv.stats()
    lat, done, work = [], 0, 0.0
    t0 = time.perf_counter()
    for i, b in enumerate(order):
        while True:
            dt = t_arr[i] - (time.perf_counter() - t0)
            for r in srv.poll(64, 0):
                done += 1
                if r.has_plane:
                    lat.append(r.t_done_ms - r.t_submit_ms)
                    srv.release(r.d_plane)
            if dt <= 0:
                break
            time.sleep(min(dt, 0.0005))
        sb = beams[b]
        srv.submit(int(b), ptrs[i], tag=i)
        work += sb.spec.ndm * sb.chunk_rows
    while done < n_total:
        for r in srv.poll(64, 1000):
            done += 1
            if r.has_plane:
                lat.append(r.t_done_ms - r.t_submit_ms)
                srv.release(r.d_plane)
    wall = time.perf_counter() - t0
    # every chunk is back: the pulse chunks' own buffers can go
    for e in pool.extra:
        e.free()
    pool.extra.clear()
    st = srv.stats()
    lat = np.array(lat)
    offered = sum(b.chunk_s for b in (beams[k] for k in order))
    return {
        "rate_x_realtime": rate if rate > 0 else None,
        "chunks": int(n_total),
        "latency_samples": int(lat.size),
        "p50_ms": float(np.percentile(lat, 50)),
        "p95_ms": float(np.percentile(lat, 95)),
        "p99_ms": float(np.percentile(lat, 99)),
        "max_ms": float(lat.max()),
        "throughput": work / wall,
        "wall_s": wall,
        "stream_s": offered,
        "min_chunk_ms": min(b.chunk_s for b in beams) * 1e3,
        "inflight_peak": st["inflight_peak"],
        "hbm_peak_GB": st["hbm_peak"] / 1e9,
        "completed": int(st["completed"] - base["completed"]),
    }


def feed_file(srv: capi.Server, beam: int, reader: "capi.FileReader", chunk_rows: int,
              nbuf: int = 4, poll_ms: int = 5000):
    """Streams a file through the server: chunk k's rows are read (positioned
    reads, no intermediate copy) into page-locked slot k mod nbuf and
    submitted; a slot is reused only after its chunk's result came back, as
    the server borrows submitted rows until then.  Yields the server results
    (planes still to be released by the caller) in completion order; results
    of other beams polled meanwhile are yielded too."""
    nchunks = reader.ndata // chunk_rows
    ring = [capi.HostBuffer((chunk_rows, reader.nchans), reader.dtype) for _ in range(nbuf)]
    busy = {}                  # tag -> slot
    free = list(range(nbuf))
    k = 0
    try:
        while k < nchunks or busy:
            while k < nchunks and free:
                slot = free.pop()
                reader.read_into(k * chunk_rows, chunk_rows, ring[slot].ptr)
                srv.submit(beam, ring[slot].ptr, tag=k)
                busy[k] = slot
                k += 1
            rs = srv.poll(64, poll_ms)
            if not rs:
                raise RuntimeError("server returned no result")
            for r in rs:
                if r.beam == beam and r.tag in busy:   # results of other beams pass through
                    free.append(busy.pop(r.tag))
                yield r
    finally:
        srv.drain()
        for b in ring:
            b.free()

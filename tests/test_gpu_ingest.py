"""A SIGPROC file streamed into the server through page-locked buffers
(stream.feed_file over af_fil_read): every served window equals the plain
call on the same rows, and the rows the reader delivered are the file's,
flipped to ascending frequency.  A PSRFITS file's pol-averaged rows go
through the plain call as well."""
import sys

import numpy as np
import pytest

from astroflow_amd import capi, stream, synth
from tests.helpers import write_psrfits, write_sigproc

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def fetch(ptr, shape, dtype):
    a = np.empty(shape, dtype)
    capi.check(capi.lib().af_memcpy_d2h(capi._ptr(a), capi.C.c_void_p(ptr), a.nbytes, None))
    return a


def test_filterbank_file_feeds_the_server(gpu_lib, tmp_path):
    say("case: descending-band .fil streamed through pinned buffers into the server")
    nchan, C = 512, 1024
    band = synth.band_1000_1500(nchan)
    spec = stream.window_spec(band, C, 300.0, 1.0, sbw=32)
    W = spec.nsamp
    nchunks = -(-W // C) + 3
    x = synth.c3_chunk_data(dict(nchan=nchan, nsamp=nchunks * C, dtype=np.uint8), seed=77,
                            threads=8)
    synth.inject_pulse(x, band.freq, band.f_top, float(band.tsamp), 150.0, W + 300, 1e-3, 48.0)
    p = tmp_path / "beam.fil"
    f = band.freq
    write_sigproc(str(p), x[:, ::-1], fch1=float(f[-1]), foff=-float(f[1] - f[0]),
                  tsamp=float(band.tsamp))
    reader = capi.FileReader(str(p))
    assert reader.info.flipped == 1 and reader.ndata == nchunks * C
    assert np.array_equal(reader.read(), x)
    srv = capi.Server()
    beam = srv.add_beam(spec, C)
    planes = {}
    for r in stream.feed_file(srv, beam, reader, C, nbuf=3):
        assert r.status == 0
        if r.has_plane:
            planes[int(r.window)] = fetch(r.d_plane, (spec.ndm, C), np.uint32)
            srv.release(r.d_plane)
    srv.close()
    reader.close()
    nwin = (nchunks * C - W) // C + 1
    assert sorted(planes) == list(range(nwin))
    plan = capi.Plan(spec)
    for j in range(nwin):
        alone, _ = plan.run_host(x[j * C: j * C + W])
        assert np.array_equal(planes[j], alone), j
    plan.close()


def test_psrfits_rows_dedisperse_like_the_array(gpu_lib, tmp_path):
    say("case: two-pol PSRFITS, pol mean through the plain call")
    nchan, nsblk, nsub = 256, 512, 8
    band = synth.band_1000_1500(nchan)
    rng = np.random.default_rng(12)
    data = rng.integers(60, 200, (nsblk * nsub, 2, nchan), dtype=np.int32).astype(np.uint8)
    p = tmp_path / "obs.sf"
    f = band.freq
    foff = float(f[1] - f[0])
    write_psrfits(str(p), data[:, :, ::-1], float(f[0] + f[-1]) / 2, -foff * nchan, -foff,
                  float(band.tsamp), nsblk)
    r = capi.FileReader(str(p), kind="psrfits")
    rows = r.read()
    r.close()
    want = ((data[:, 0, :].astype(np.int32) + data[:, 1, :]) // 2).astype(np.uint8)
    assert np.array_equal(rows, want)
    spec = stream.window_spec(band, 1024, 40.0, 1.0, sbw=32)
    assert spec.nsamp <= rows.shape[0]
    plan = capi.Plan(spec)
    a, _ = plan.run_host(rows[:spec.nsamp])
    b, _ = plan.run_host(np.ascontiguousarray(want[:spec.nsamp]))
    assert np.array_equal(a, b)
    plan.close()

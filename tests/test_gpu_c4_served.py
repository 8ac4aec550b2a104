"""The served stream at its real shape: beams of 4096 channels u8 over
1000-1500 MHz (tsamp 49.152 us), DM 0-5000 step 1 (5001 trials), 32 subbands
of 128 channels, IQRM (std, radius 5 channels, 3 sigma, one window) +
zero-DM, chunks of 0.25, 0.5 and 1.0 s (one beam each), the DM-overlap carry
(234,363 rows) resident on the device.  Chunks of the three beams are
submitted interleaved, so windows of different beams are in flight together.

  - every served window equals the plain call (af_execute) on the same rows,
    plane and IQRM masks, bit for bit ("alone" = "in the mix");
  - the first and last window of every beam equal the oracle: IQRM masks,
    the DM-0 row against the per-sample channel sums of the cleaned rows, and
    16 spread DM rows bit for bit;
  - one channel turns noisy part way through each beam, so the windows'
    masks change and the server's re-clean path runs as well as the
    reuse path."""
import sys

import numpy as np
import pytest

from astroflow_amd import capi, stream, synth
from oracle import oracle as O
from tests.helpers import oracle_spec

pytestmark = pytest.mark.gpu

NCHAN = 4096
NWIN = 3          # windows per beam
DURATIONS = (0.25, 0.5, 1.0)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def fetch(ptr, shape, dtype):
    a = np.empty(shape, dtype)
    capi.check(capi.lib().af_memcpy_d2h(capi._ptr(a), capi.C.c_void_p(ptr), a.nbytes, None))
    return a


def make_beam(k, dur, band):
    C = int(np.ceil(dur / float(band.tsamp) / 256)) * 256
    spec = stream.window_spec(band, C, 5000.0, 1.0)
    W = spec.nsamp
    nchunks = -(-W // C) + NWIN - 1
    nrows = nchunks * C
    seed = 500 + k
    x = synth.c3_chunk_data(dict(nchan=NCHAN, nsamp=nrows, dtype=np.uint8), seed=seed,
                            threads=16)
    # a clean channel (not one of the generator's RFI channels) that turns
    # noisy after the first window's rows: window 0 does not flag it, the
    # windows holding those rows do
    rfi = set(np.random.default_rng(seed + 7919).choice(NCHAN, size=205 + 3,
                                                         replace=False).tolist())
    ch = next(c for c in range(1000 + 37 * k, NCHAN) if c not in rfi)
    t = np.arange(W, nrows)
    x[W:, ch] = np.clip(128 + 90 * np.sin(t * 0.37), 0, 255).astype(np.uint8)
    synth.inject_pulse(x, band.freq, band.f_top, float(band.tsamp), 700.0 + 900.0 * k,
                       W + C // 2, 1e-3, 48.0)
    return dict(spec=spec, C=C, W=W, x=x, nrows=nrows, noisy=ch,
                nwin=(nrows - W) // C + 1)


@pytest.mark.timeout(600)
def test_c4_served_windows_equal_plain_calls_and_oracle(gpu_lib):
    say("case: c4 served stream, 3 beams x 4096 ch, DM 0-5000/1, 0.25/0.5/1.0 s chunks")
    band = synth.band_1000_1500(NCHAN)
    beams = [make_beam(k, d, band) for k, d in enumerate(DURATIONS)]
    srv = capi.Server()
    host = []
    for b in beams:
        srv.add_beam(b["spec"], b["C"])
        hb = capi.HostBuffer(b["x"].shape, np.uint8)
        hb.array[:] = b["x"]
        host.append(hb)
        p = capi.Plan(b["spec"])
        b["info"] = p.info
        p.close()
    say(f"beams: chunk rows {[b['C'] for b in beams]}, window rows {[b['W'] for b in beams]}")
    order = []
    nmax = max(b["nrows"] // b["C"] for b in beams)
    for k in range(nmax):
        for i, b in enumerate(beams):
            if k < b["nrows"] // b["C"]:
                order.append((i, k))
    for i, k in order:
        b = beams[i]
        srv.submit(i, host[i].ptr + k * b["C"] * NCHAN, tag=i * 100000 + k)
    got = {}
    expected = len(order)
    while expected:
        rs = srv.poll(64, 5000)
        assert rs, "server stopped producing results"
        for r in rs:
            expected -= 1
            assert r.status == 0
            if not r.has_plane:
                continue
            info = beams[r.beam]["info"]
            plane = fetch(r.d_plane, (info.ndm, info.nout), np.uint32)
            masks = fetch(r.d_masks, (info.nwin, NCHAN), np.uint8)
            srv.release(r.d_plane)
            got[(r.beam, r.window)] = (plane, masks)
    st = srv.stats()
    srv.close()
    for hb in host:
        hb.free()
    say(f"served: {st['planes']} windows, p50 {st['p50_ms']:.2f} ms, p99 {st['p99_ms']:.2f} ms")
    assert st["planes"] == sum(b["nwin"] for b in beams)
    for i, b in enumerate(beams):
        plan = capi.Plan(b["spec"])
        wmasks = []
        for j in range(b["nwin"]):
            xw = b["x"][j * b["C"]: j * b["C"] + b["W"]]
            alone, amask = plan.run_host(xw, want_masks=True)
            plane, masks = got[(i, j)]
            assert np.array_equal(masks, amask), f"beam {i} window {j}: masks"
            assert np.array_equal(plane, alone), f"beam {i} window {j}: plane"
            wmasks.append(masks)
            if j in (0, b["nwin"] - 1):
                say(f"beam {i} window {j}: oracle")
                s = oracle_spec(b["spec"])
                rows = sorted(set(np.linspace(0, s.ndm - 1, 16).round().astype(int).tolist()))
                ref, rmask, cleaned = O.run_pipeline_rows(xw, s, rows, O.IQRM_GPU)
                assert np.array_equal(masks, rmask), f"beam {i} window {j}: oracle masks"
                assert np.array_equal(plane[0],
                                      cleaned[:b["C"]].astype(np.uint32).sum(1)), (i, j)
                for k, m in enumerate(rows):
                    assert np.array_equal(plane[m], ref[k]), f"beam {i} window {j} row {m}"
        plan.close()
        # the channel that turns noisy is flagged in the later windows only
        ch = b["noisy"]
        assert wmasks[-1][0, ch] == 1 and wmasks[0][0, ch] == 0, (i, [m[0, ch] for m in wmasks])


def test_release_of_a_plane_twice_is_rejected(gpu_lib):
    say("case: server rejects a second release of the same plane")
    band = synth.band_1000_1500(256)
    spec = stream.window_spec(band, 512, 50.0, 1.0, sbw=32)
    W = spec.nsamp
    C = 512
    nrows = (-(-W // C)) * C
    x = synth.noise_u8(np.random.default_rng(3), nrows, 256)
    hb = capi.HostBuffer(x.shape, np.uint8)
    hb.array[:] = x
    srv = capi.Server()
    srv.add_beam(spec, C)
    for k in range(nrows // C):
        srv.submit(0, hb.ptr + k * C * 256, tag=k)
    srv.drain()
    planes = [r for r in srv.poll(1024, 0) if r.has_plane]
    assert len(planes) == 1
    srv.release(planes[0].d_plane)
    with pytest.raises(ValueError):
        srv.release(planes[0].d_plane)
    srv.close()
    hb.free()

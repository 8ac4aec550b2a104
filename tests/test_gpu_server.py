"""The chunk server: every served result equals the plain call on the same
window of rows ("alone") and the oracle, whatever else was in flight ("in
the mix"); admission under a tight HBM budget; queue and budget errors."""
import dataclasses
import sys

import numpy as np
import pytest

from astroflow_amd import capi, synth
from oracle import oracle as O
from tests.helpers import make_spec, oracle_spec, small_band

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def beam_stream(seed, nchan, nrows, band):
    rng = np.random.default_rng(seed)
    x = synth.noise_u8(rng, nrows, nchan)
    synth.inject_narrowband_rfi(x, rng, frac=0.05, amp=40, n_stuck=2, pmax=256)
    synth.inject_pulse(x, band.freq, band.f_top, float(band.tsamp), 60.0, nrows // 2, 2e-3, 40.0)
    # a channel that turns noisy half way: windows after it flag it, so the
    # served masks change from window to window at least once
    t = np.arange(nrows // 2, nrows)
    x[nrows // 2:, nchan // 3 + 1] = np.clip(128 + 90 * np.sin(t * 0.7), 0, 255).astype(np.uint8)
    return x


# (seed, nchan, chunk rows, DM step, ndm, method, extra spec fields).  Chunks
# of a multiple of 256 rows take the server's incremental RFI path (beams 0
# and 2); beam 1 runs the plan's full execute on every window.
BEAMS = [
    (11, 256, 512, 0.5, 161, "subband",
     dict(sbw=32, ndm0=16, use_iqrm=True, iqrm_c0=0, iqrm_c1=255, iqrm_mode=1,
          iqrm_radius_frac=np.float32(5.5 / 255), iqrm_nsigma=3.0,
          use_zero_dm=True, zdm_c0=0, zdm_c1=255)),
    (12, 128, 1000, 1.0, 90, "direct",
     dict(use_zero_dm=True, zdm_c0=0, zdm_c1=128, bad_channels=[3, 77])),
    (13, 512, 256, 0.25, 240, "subband",
     dict(sbw=64, ndm0=32, use_iqrm=True, iqrm_c0=0, iqrm_c1=511, iqrm_mode=0,
          iqrm_radius_frac=np.float32(0.02), iqrm_nsigma=3.0)),
]


def window_spec(nchan, C, dm_step, ndm, method, extra):
    band = small_band(nchan, 1200.0, 1500.0, 1e-4)
    dmax = synth.max_delay(band, dm_step * (ndm - 1))
    W = C + dmax
    x0 = np.zeros((W, nchan), np.uint8)
    spec = make_spec(x0, band, 0.0, dm_step, ndm, method=method, nout=C, **extra)
    return spec, band, W


def fetch(ptr, shape, dtype):
    a = np.empty(shape, dtype)
    capi.check(capi.lib().af_memcpy_d2h(capi._ptr(a), capi.C.c_void_p(ptr), a.nbytes, None))
    return a


def make_beams(nchunks):
    beams = []
    for seed, nchan, C, dm_step, ndm, method, extra in BEAMS:
        spec, band, W = window_spec(nchan, C, dm_step, ndm, method, extra)
        nrows = (nchunks + (W + C - 1) // C) * C
        x = beam_stream(seed, nchan, nrows, band)
        beams.append(dict(spec=spec, C=C, W=W, x=x, nrows=nrows))
    return beams


def serve_all(srv, beams, host, release_as_we_go=False, interleave=True):
    """Submits every chunk of every beam (interleaved across beams), collects
    all results; returns {(beam, window): (plane, masks)}."""
    order = []
    nmax = max(b["nrows"] // b["C"] for b in beams)
    for k in range(nmax):
        for i, b in enumerate(beams):
            if k < b["nrows"] // b["C"]:
                order.append((i, k))
    if not interleave:
        order.sort()
    out, expected = {}, 0
    for i, k in order:
        srv.submit(i, host[i].ptr + k * beams[i]["C"] * beams[i]["spec"].nchan, tag=i * 100000 + k)
        expected += 1
        if release_as_we_go:
            for r in srv.poll(64, 0):
                expected -= 1
                _collect(srv, beams, r, out)
    while expected:
        rs = srv.poll(64, 2000)
        assert rs, "server stopped producing results"
        for r in rs:
            expected -= 1
            _collect(srv, beams, r, out)
    return out


def _collect(srv, beams, r, out):
    assert r.status == 0
    if not r.has_plane:
        return
    b = beams[r.beam]
    plan_info = b["info"]
    plane = fetch(r.d_plane, (plan_info.ndm, plan_info.nout), np.uint32)
    masks = None
    if plan_info.nwin:
        masks = fetch(r.d_masks, (plan_info.nwin, b["spec"].nchan), np.uint8)
    srv.release(r.d_plane)
    out[(r.beam, r.window)] = (plane, masks)


@pytest.mark.parametrize("full_rfi", [False, True])
def test_served_results_equal_plain_calls(gpu_lib, full_rfi):
    say(f"case: server (full_rfi={full_rfi}), 3 beams interleaved, every window vs the plain "
        "call and the oracle")
    beams = make_beams(nchunks=6)
    srv = capi.Server(full_rfi=full_rfi)
    host = []
    for b in beams:
        srv.add_beam(b["spec"], b["C"])
        hb = capi.HostBuffer(b["x"].shape, np.uint8)
        hb.array[:] = b["x"]
        host.append(hb)
        p = capi.Plan(b["spec"])
        b["info"] = p.info
        p.close()
    got = serve_all(srv, beams, host)
    st = srv.stats()
    srv.close()
    for i, b in enumerate(beams):
        nwin = (b["nrows"] - b["W"]) // b["C"] + 1
        plan = capi.Plan(b["spec"])
        for j in range(nwin):
            xw = b["x"][j * b["C"]: j * b["C"] + b["W"]]
            alone, amask = plan.run_host(xw, want_masks=True)
            plane, masks = got[(i, j)]
            assert np.array_equal(plane, alone), (i, j)
            if masks is not None:
                assert np.array_equal(masks, amask), (i, j)
            if j in (0, nwin - 1):
                ref, rmask, *_ = O.run_pipeline(xw, oracle_spec(b["spec"]), O.IQRM_GPU)
                assert np.array_equal(plane, ref), (i, j)
                if masks is not None:
                    assert np.array_equal(masks, rmask), (i, j)
        plan.close()
        assert sum(1 for (bi, _) in got if bi == i) == nwin
        if i == 0:
            # both incremental branches ran: reused overlap (same mask as the
            # previous window) and a full re-clean (mask changed); beam 2's
            # mean-mode masks change at every window
            same = [np.array_equal(got[(i, j)][1], got[(i, j - 1)][1]) for j in range(1, nwin)]
            assert any(same) and not all(same), (i, same)
    assert st["planes"] == len(got) and st["p99_ms"] >= st["p50_ms"] > 0
    for hb in host:
        hb.free()


def test_tight_budget_admits_as_planes_are_released(gpu_lib):
    say("case: server with room for one result plane")
    beams = make_beams(nchunks=3)[:2]
    probe = capi.Server()
    host = []
    for b in beams:
        probe.add_beam(b["spec"], b["C"])
        p = capi.Plan(b["spec"])
        b["info"] = p.info
        p.close()
        hb = capi.HostBuffer(b["x"].shape, np.uint8)
        hb.array[:] = b["x"]
        host.append(hb)
    resident = probe.stats()["hbm_used"]
    probe.close()
    largest = max(b["info"].ndm * b["info"].nout * 4 + b["info"].nwin * b["spec"].nchan
                  for b in beams)
    srv = capi.Server(hbm_budget=resident + largest + 4096)
    for b in beams:
        srv.add_beam(b["spec"], b["C"])
    got = serve_all(srv, beams, host, release_as_we_go=True)
    st = srv.stats()
    srv.close()
    assert st["hbm_peak"] <= st["hbm_budget"]
    for i, b in enumerate(beams):
        plan = capi.Plan(b["spec"])
        j = max(w for (bi, w) in got if bi == i)
        xw = b["x"][j * b["C"]: j * b["C"] + b["W"]]
        alone, _ = plan.run_host(xw)
        assert np.array_equal(got[(i, j)][0], alone)
        plan.close()
    for hb in host:
        hb.free()


def test_server_limits(gpu_lib):
    say("case: server budget and queue limits")
    spec, band, W = window_spec(*BEAMS[1][1:])
    srv = capi.Server(hbm_budget=4096)              # far too small for the beam buffers
    with pytest.raises(MemoryError):
        srv.add_beam(spec, BEAMS[1][2])
    srv.close()
    srv = capi.Server(max_queue=1, hbm_budget=0)
    b = srv.add_beam(spec, BEAMS[1][2])
    with pytest.raises(ValueError):                 # served beams run at ds = 1
        srv.add_beam(dataclasses.replace(spec, ds=2, nout=BEAMS[1][2] // 2), BEAMS[1][2] // 2)
    hb = capi.HostBuffer((BEAMS[1][2] * 4, 128), np.uint8)
    hb.array[:] = 7
    n_ok, busy = 0, 0
    for k in range(4):
        try:
            srv.submit(b, hb.ptr + k * BEAMS[1][2] * 128, k)
            n_ok += 1
        except capi.AFError:
            busy += 1
    srv.drain()
    done = []
    while True:
        rs = srv.poll(16, 0)
        if not rs:
            break
        done += rs
    srv.close()
    assert n_ok + busy == 4 and n_ok >= 1 and len(done) == n_ok
    hb.free()

"""Full-size parity on the bench workload (4096 ch u8, 150791 input rows,
6001 DMs, 32 subbands, IQRM + zero-DM) and on the direct-method case (1024 ch
x 2^18, DM 0-1000): IQRM masks bit-exact, and every 16th DM row (plus the
coarse-group edges and pulse DMs) bit-exact against the oracle; plus
row-independent properties."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import oracle_spec

pytestmark = pytest.mark.gpu


def test_bench_workload_full_size(gpu_lib):
    print("case: full-size 4096ch x 150791 subband + IQRM + zero-DM", flush=True)
    import bench
    from astroflow_amd import capi
    xh, spec, info = bench.make_workload(threads=8)
    plan = capi.Plan(spec)
    d_in = capi.DeviceBuffer(xh.nbytes)
    d_in.upload(xh)
    d_out = capi.DeviceBuffer(info["ndm"] * info["nout"] * 4)
    d_m = capi.DeviceBuffer(plan.nwin * info["nchan"])
    plan.execute(d_in.ptr, d_out.ptr, d_m.ptr)
    capi.sync()
    plane = d_out.download(np.empty((info["ndm"], info["nout"]), np.uint32))
    gmask = d_m.download(np.empty((plan.nwin, info["nchan"]), np.uint8))
    for b in (d_in, d_out, d_m):
        b.free()
    print("gpu done; oracle RFI stages", flush=True)

    s = oracle_spec(spec)
    cleaned = xh.copy()
    rmask, t0, t1 = O.iqrm(cleaned, s.iqrm_c0, s.iqrm_c1, s.iqrm_mode, s.iqrm_radius_frac,
                           s.iqrm_nsigma, s.iqrm_geofactor, s.iqrm_W, s.iqrm_H,
                           s.iqrm_include_tail, O.IQRM_GPU)
    assert np.array_equal(gmask, rmask)
    flagged = int(rmask.sum())
    assert flagged > 50, flagged               # the RFI channels are caught
    O.apply_window_masks(cleaned, rmask, t0, t1)
    O.zero_dm(cleaned, s.zdm_c0, s.zdm_c1)

    f = s.freq
    nnom = (s.ndm + s.ndm0 - 1) // s.ndm0
    d1 = O.delay_table(f, 0, 4096, f[-1], s.tsamp, 1,
                       s.dm_low + np.arange(nnom) * s.ndm0 * s.dm_step)
    res = O.residual_table(f, 0, 4096, s.sbw, f[-1], s.tsamp, 1, s.dm_low, s.dm_step,
                           s.ndm, s.ndm0)
    # every 16th DM row, plus coarse-group edges and the pulse DMs
    rows = sorted(set(range(0, s.ndm, 16)) |
                  {1, 31, 32, 300, 2399, 2400, 4447, 5600, 5998, 6000})
    print(f"oracle: {len(rows)} rows", flush=True)
    ref = O.dedisperse_subband_rows(cleaned, 0, 4096, s.sbw, d1, res, s.ndm0, info["nout"], rows)
    for i, m in enumerate(rows):
        assert np.array_equal(plane[m], ref[i]), f"row {m} differs"
    # DM 0 is the per-sample channel sum of the cleaned data
    assert np.array_equal(plane[0], cleaned[: info["nout"]].astype(np.uint32).sum(1))
    # the injected pulses stand out at their DMs (DM 150/1200/2800 -> rows 300/2400/5600)
    for m in (300, 2400, 5600):
        r = plane[m].astype(np.float64)
        assert (r.max() - np.median(r)) > 8 * r.std(), m
    plan.close()


def test_direct_1024ch_full_size(gpu_lib):
    # 1024 ch x 2^18 u8 samples, DM 0-1000 step 1 (1001 trials), direct
    # method, zero-DM over [0, 1023), full-length u32 plane
    print("case: full-size 1024ch x 2^18 direct + zero-DM", flush=True)
    from astroflow_amd import capi, synth
    from tests.helpers import make_spec
    x, band, t0 = synth.direct_1024(seed=1)
    nsamp, nchan = x.shape
    spec = make_spec(x, band, 0.0, 1.0, 1001, method="direct",
                     use_zero_dm=True, zdm_c0=0, zdm_c1=nchan - 1)
    plan = capi.Plan(spec)
    d_in = capi.DeviceBuffer(x.nbytes)
    d_in.upload(x)
    d_out = capi.DeviceBuffer(1001 * nsamp * 4)
    plan.execute(d_in.ptr, d_out.ptr)
    capi.sync()
    plane = d_out.download(np.empty((1001, nsamp), np.uint32))
    d_in.free()
    d_out.free()
    plan.close()
    print("gpu done; oracle rows", flush=True)
    s = oracle_spec(spec)
    cleaned = x.copy()
    O.zero_dm(cleaned, s.zdm_c0, s.zdm_c1)
    f = s.freq
    d = O.delay_table(f, 0, nchan, f[-1], s.tsamp, 1, np.arange(1001, dtype=np.float64))
    rows = sorted(set(range(0, 1001, 16)) | {1, 250, 499, 500, 501, 999, 1000})
    print(f"oracle: {len(rows)} rows", flush=True)
    ref = O.dedisperse_direct_rows(cleaned, 0, nchan, d, nsamp, rows)
    for i, m in enumerate(rows):
        assert np.array_equal(plane[m], ref[i]), f"row {m} differs"
    assert np.array_equal(plane[0], cleaned.astype(np.uint32).sum(1))
    # the pulse peaks at DM 500 at its arrival time at the top channel
    r = plane[500].astype(np.float64)
    assert abs(int(np.argmax(r)) - t0) <= 2
    assert (r.max() - np.median(r)) > 8 * r.std()

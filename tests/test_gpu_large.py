"""The large float32 plane: the three-step DM plan (DM 0-1000/0.5 at full
rate, 1000-4000/1 at 2x, 4000-10000/2 at 4x downsampling), 64 subbands, IQRM.

* At reduced channel count (512 ch x 2^19 samples, N(0,1) + impulsive
  broadband RFI, IQRM + zero-DM): rows within 1e-6 * sum|x| of the FP64
  oracle and masks bit-exact, for every step.
* At full size (8192 ch x 2^22 float32 = 137 GB of input in HBM): the whole
  cleaned buffer does not fit beside the input and the plane, so the first
  step cleans per super-tile.  The input is zero with seeded unit impulses
  (exact in float32 at any accumulation order), so every row is known in
  closed form from the delay and residual tables: an impulse at binned row r
  of channel c lands at r - delay1(c, coarse DM) - residual(DM, subband).
"""
import time

import numpy as np
import pytest

from astroflow_amd import capi, synth
from oracle import oracle as O
from tests.helpers import oracle_spec

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def step_spec(nchan, nsamp, band, step, use_zero_dm, nsub=64):
    dm_low, dm_step, ndm, ds = step
    f = band.freq
    return capi.DedispSpec(
        dtype=np.float32, nchan=nchan, nsamp=nsamp, freq=f, tsamp=float(band.tsamp),
        dm_low=dm_low, dm_step=dm_step, ndm=ndm, dd_c0=0, dd_c1=nchan, f_ref=float(f[-1]),
        ds=ds, nout=0, method="subband", sbw=nchan // nsub, ndm0=32,
        use_iqrm=True, iqrm_c0=0, iqrm_c1=nchan - 1, iqrm_mode=1,
        iqrm_radius_frac=float(np.float32(5.5 / (nchan - 1))), iqrm_nsigma=3.0,
        iqrm_geofactor=1.5, iqrm_W=0, iqrm_H=0, iqrm_include_tail=True,
        use_zero_dm=use_zero_dm, zdm_c0=0, zdm_c1=nchan - 1)


def test_three_step_plan_reduced_channels(gpu_lib):
    nchan, nsamp = 512, 1 << 19
    say(f"case: three-step f32 plan, {nchan} ch x {nsamp}, N(0,1) + impulsive RFI")
    band = synth.band_1000_1500(nchan)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((nsamp, nchan), dtype=np.float32)
    synth.inject_impulsive_rfi(x, rng)
    d_in = capi.DeviceBuffer(x.nbytes)
    d_in.upload(x)
    for step in synth.C5_STEPS:
        spec = step_spec(nchan, nsamp, band, step, True)
        plan = capi.Plan(spec)
        ndm, nout = plan.info.ndm, plan.info.nout
        d_out = capi.DeviceBuffer(ndm * nout * 4)
        d_m = capi.DeviceBuffer(max(1, plan.nwin) * nchan)
        plan.execute(d_in.ptr, d_out.ptr, d_m.ptr)
        capi.sync()
        rows = sorted({0, 1, 31, 32, ndm // 2, ndm - 2, ndm - 1})
        got = np.stack([d_out.download(np.empty(nout, np.float32), m * nout * 4) for m in rows])
        masks = d_m.download(np.empty((plan.nwin, nchan), np.uint8))
        plan.close()
        d_out.free()
        d_m.free()
        say(f"step {step}: gpu done, oracle rows {rows}")
        s = oracle_spec(spec)
        ref, rmask, cleaned = O.run_pipeline_rows(x, s, rows)
        assert np.array_equal(masks, rmask), step
        # tolerance: 1e-6 * the same rows over |cleaned| (+1 keeps it positive)
        d1 = O.delay_table(s.freq, 0, nchan, s.f_ref, s.tsamp, s.ds, s.coarse_dms())
        res = O.residual_table(s.freq, 0, nchan, s.sbw, s.f_ref, s.tsamp, s.ds, s.dm_low,
                               s.dm_step, s.ndm, s.ndm0)
        bound = O.dedisperse_subband_rows(np.abs(cleaned) + 1.0, 0, nchan, s.sbw, d1, res,
                                          s.ndm0, nout, rows)
        err = np.abs(got.astype(np.float64) - ref)
        assert np.all(err <= 1e-6 * bound), (step, float((err / bound).max()))
    d_in.free()


def test_three_step_plan_full_size_8192ch(gpu_lib):
    nchan, nsamp = 8192, 1 << 22
    band = synth.band_1000_1500(nchan)
    f = band.freq.astype(np.float64)
    in_bytes = nsamp * nchan * 4
    say(f"case: three-step f32 plan at full size, {nchan} ch x {nsamp} ({in_bytes / 1e9:.1f} GB)")
    d_in = capi.DeviceBuffer(in_bytes)
    d_in.zero()
    rng = np.random.default_rng(60)
    nimp = 64
    imp_t = np.sort(rng.choice(nsamp, size=nimp, replace=False))
    imp_c = rng.integers(0, nchan - 1, size=nimp)   # the top channel is outside IQRM
    one = np.ones(1, np.float32)
    for t, c in zip(imp_t, imp_c):
        d_in.upload_at(one, (int(t) * nchan + int(c)) * 4)
    capi.sync()
    for si, step in enumerate(synth.C5_STEPS):
        dm_low, dm_step, ndm, ds = step
        nout = (nsamp + ds - 1) // ds
        d_out = capi.DeviceBuffer(ndm * nout * 4)    # before the plan: it sizes itself to what is left
        spec = step_spec(nchan, nsamp, band, step, False)
        t0 = time.perf_counter()
        plan = capi.Plan(spec)
        t1 = time.perf_counter()
        if si == 0:
            assert plan.info.clean_passes > 1, "the whole cleaned buffer should not fit"
        d_m = capi.DeviceBuffer(plan.nwin * nchan)
        plan.execute(d_in.ptr, d_out.ptr, d_m.ptr)
        capi.sync()
        t2 = time.perf_counter()
        say(f"step {step}: plan {t1 - t0:.1f} s, execute {t2 - t1:.2f} s, "
            f"clean passes {plan.info.clean_passes}")
        masks = d_m.download(np.empty((plan.nwin, nchan), np.uint8))
        assert masks.sum() == 0        # a zero channel spread: no lag is informative
        # closed form of the rows: impulses binned (sum of ds samples), then
        # placed by the plan's own FP64 tables, which equal the oracle's
        sbw = nchan // 64
        nnom = (ndm + 31) // 32
        d1 = O.delay_table(f, 0, nchan, f[-1], float(band.tsamp), ds,
                           dm_low + np.arange(nnom) * 32 * dm_step)
        res = O.residual_table(f, 0, nchan, sbw, f[-1], float(band.tsamp), ds, dm_low, dm_step,
                               ndm, 32)
        d1p, resp = plan.tables()
        assert np.array_equal(d1p, d1) and np.array_equal(resp, res)
        rows = sorted({0, 1, 33, ndm // 3, ndm // 2, ndm - 1} |
                      set(int(v) for v in rng.integers(0, ndm, 4)))
        rb = imp_t // ds
        for m in rows:
            want = np.zeros(nout, np.float32)
            tt = rb - d1[m // 32, imp_c] - res[m, imp_c // sbw]
            ok = (tt >= 0) & (tt < nout)
            np.add.at(want, tt[ok], 1.0)
            got = d_out.download(np.empty(nout, np.float32), m * nout * 4)
            assert np.array_equal(got, want), (step, m)
        plan.close()
        d_out.free()
        d_m.free()
    d_in.free()

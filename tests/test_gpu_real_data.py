"""The HIP path on real data: the reference's fixture tests/FRB180417.fil
(336 ch, 8-bit, 5120 samples; FRB 180417, published DM 474.8).  Planes and
IQRM masks equal the oracle's bit for bit, for both methods, and the
drop-in `_dedisperse_spec_raw` finds the burst at its DM.  The float32 path
(the same samples as float32) is within 1e-6 * sum|x| of the FP64 oracle, and
2x time binning is bit-exact."""
import sys

import numpy as np
import pytest

from astroflow_amd import capi
from oracle import oracle as O
from tests import real_data as R

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("method", ["direct", "subband"])
def test_real_data_planes_and_masks_equal_the_oracle(gpu_lib, method):
    say(f"case: FRB180417.fil, {method}, DM 0-1000/1, IQRM + zero-DM")
    x, freq, tsamp, _ = R.load()
    os_ = R.oracle_spec(freq, tsamp, method)
    spec = capi.DedispSpec(dtype=np.uint8, nchan=336, nsamp=5120, freq=freq, tsamp=tsamp,
                           dm_low=0.0, dm_step=1.0, ndm=1001, dd_c0=0, dd_c1=336,
                           f_ref=float(freq[-1]), method=method, sbw=16, ndm0=8, use_iqrm=True,
                           iqrm_c0=0, iqrm_c1=335, iqrm_mode=os_.iqrm_mode,
                           iqrm_radius_frac=os_.iqrm_radius_frac, iqrm_nsigma=os_.iqrm_nsigma,
                           iqrm_geofactor=os_.iqrm_geofactor, use_zero_dm=True, zdm_c0=0,
                           zdm_c1=336)
    plan = capi.Plan(spec)
    got, masks = plan.run_host(x, want_masks=True)
    plan.close()
    ref, rmask, *_ = O.run_pipeline(x, os_, O.IQRM_GPU)
    assert np.array_equal(masks, rmask) and rmask.any()
    assert np.array_equal(got, ref)
    m, snr, t = R.best_dm(got)
    assert abs(m - R.FRB_DM) <= 2.5 and abs(t - 1602) <= 3


def _spec(freq, tsamp, method, dtype=np.uint8, ds=1):
    return capi.DedispSpec(dtype=dtype, nchan=336, nsamp=5120, freq=freq, tsamp=tsamp,
                           dm_low=0.0, dm_step=1.0, ndm=1001, dd_c0=0, dd_c1=336,
                           f_ref=float(freq[-1]), method=method, sbw=16, ndm0=8, ds=ds,
                           use_iqrm=True, iqrm_c0=0, iqrm_c1=335, use_zero_dm=True, zdm_c0=0,
                           zdm_c1=336)


def test_real_data_float32_within_tolerance(gpu_lib):
    say("case: FRB180417.fil as float32, subband, IQRM + zero-DM")
    from tests.test_gpu_parity import assert_f32_close
    from tests.helpers import oracle_spec
    x, freq, tsamp, _ = R.load()
    xf = x.astype(np.float32)
    spec = _spec(freq, tsamp, "subband", np.float32)
    plan = capi.Plan(spec)
    got, masks = plan.run_host(xf, want_masks=True)
    plan.close()
    ref, rmask, *_ = O.run_pipeline(xf, oracle_spec(spec), O.IQRM_GPU)
    assert np.array_equal(masks, rmask)
    assert_f32_close(got, ref, xf, spec)
    m, snr, t = R.best_dm(got)
    assert abs(m - R.FRB_DM) <= 2.5 and abs(t - 1602) <= 3


def test_real_data_downsampled_by_two(gpu_lib):
    say("case: FRB180417.fil, direct, ds 2")
    from tests.helpers import oracle_spec
    x, freq, tsamp, _ = R.load()
    spec = _spec(freq, tsamp, "direct", ds=2)
    plan = capi.Plan(spec)
    got, masks = plan.run_host(x, want_masks=True)
    plan.close()
    ref, rmask, *_ = O.run_pipeline(x, oracle_spec(spec), O.IQRM_GPU)
    assert got.shape == (1001, 2560)
    assert np.array_equal(masks, rmask)
    assert np.array_equal(got, ref)
    m, snr, t = R.best_dm(got)
    assert abs(m - R.FRB_DM) <= 2.5 and abs(t - 801) <= 2


def test_drop_in_call_finds_the_burst(gpu_lib):
    say("case: _dedisperse_spec_raw on FRB180417.fil")
    from astroflow_amd import _astroflow_core as core
    x, _, _, f = R.load()
    h = core.Header(float(f.tstart), float(f.tsamp), float(f.fch1), float(f.foff), int(f.nbits),
                    int(f.nchans), int(f.ndata), "FRB180417")
    q = core.IQRMconfig(1, 0.1, 3.0, 1.5, 0.0, 0.0, True)
    rfi = core.RFIconfig(False, True, True, q)
    plane, masks = core._dedisperse_spec_raw(x, h, 0.0, 1000.0, 1130.0, 1465.0, 1.0, 1, 0, "",
                                             rfi, True)
    assert plane.shape == (1001, 5120) and plane.dtype == np.uint32
    # the oracle with the parameters the reference derives for this call:
    # float32 frequency table, truncating channel range, the top selected
    # channel as reference, 32-channel subbands with 32 fine DMs per coarse DM
    freq = O.freq_table_f32(float(f.fch1), float(f.foff), int(f.nchans))
    cs, ce = O.chan_range(freq, 1130.0, 1465.0)
    spec = O.PlanSpec(nchan=336, freq=freq, tsamp=float(np.float32(f.tsamp)), dm_low=0.0,
                      dm_step=1.0, ndm=O.dm_steps(0.0, 1000.0, 1.0), dd_c0=cs, dd_c1=ce + 1,
                      f_ref=freq[ce], method="subband", sbw=32, ndm0=32, use_iqrm=True,
                      iqrm_c0=cs, iqrm_c1=ce, iqrm_mode=1, iqrm_radius_frac=0.1, iqrm_nsigma=3.0,
                      iqrm_geofactor=1.5, use_zero_dm=True, zdm_c0=cs, zdm_c1=ce,
                      bad_channels=[])
    ref, rmask, *_ = O.run_pipeline(x, spec, O.IQRM_GPU)
    assert np.array_equal(masks, rmask)
    assert np.array_equal(plane, ref)
    # coarser than the per-test plans above (32 fine DMs share one coarse
    # delay per subband), so the peak is lower, but at the same DM and time
    m, snr, t = R.best_dm(plane)
    assert abs(m - R.FRB_DM) <= 2.5 and snr > 5.0 and abs(t - 1602) <= 3

"""Edge shapes and error paths of the HIP path, through the C-ABI.

* An input shorter than the DM sweep: samples past the end of the input are
  zero (the oracle's zero padding, `include/cpucal.hpp:167-186`), for the
  direct and the subband method and every integer kernel family.
* One channel: every DM row is the input column (its delay is 0 at the
  reference frequency).
* One DM trial, and a short output (`nout` below the input length).
* A constant input: the IQRM statistic is flat, so no lag flags a channel
  and the mask is empty; zero-DM then subtracts the constant and the plane
  is all zero.
* Invalid plan descriptions fail with the validator's message (ValueError),
  not a fault.
"""
import sys

import numpy as np
import pytest

from astroflow_amd import capi
from tests.helpers import make_spec, small_band
from tests.test_gpu_parity import noise, run_both

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("method,dtype", [("direct", np.uint8), ("subband", np.uint8),
                                          ("subband", np.uint16), ("direct", np.uint32)])
def test_input_shorter_than_the_dm_sweep(gpu_lib, method, dtype):
    say(f"case: {method} {np.dtype(dtype).name} 128ch x 400 samples, largest delay ~1000 samples")
    rng = np.random.default_rng(41)
    b = small_band(128, 1200.0, 1500.0, 1e-4)
    x = noise(rng, 400, 128, dtype)
    kw = dict(method=method)
    if method == "subband":
        kw.update(sbw=32, ndm0=8)
    spec = make_spec(x, b, 0.0, 2.0, 51, **kw)
    plan = capi.Plan(spec)
    assert plan.info.max_delay > 2 * x.shape[0]
    plan.close()
    got, ref, *_ = run_both(x, spec)
    assert np.array_equal(got, ref)
    # the high-DM rows see only the channels whose delayed window still
    # overlaps the input: row sums are below the DM-0 row's
    assert got[-1].sum() < got[0].sum()


def test_one_channel_rows_are_the_input(gpu_lib):
    say("case: direct u8, one channel, 9 DMs")
    rng = np.random.default_rng(42)
    b = small_band(1, 1400.0, 1400.0, 1e-4)
    x = noise(rng, 3000, 1)
    spec = make_spec(x, b, 0.0, 50.0, 9, method="direct")
    got, ref, *_ = run_both(x, spec)
    assert np.array_equal(got, ref)
    for m in range(9):
        assert np.array_equal(got[m], x[:, 0].astype(np.uint32))


@pytest.mark.parametrize("method", ["direct", "subband"])
def test_one_dm_and_a_short_output(gpu_lib, method):
    say(f"case: {method} u8 256ch, one DM trial, nout 777 of 5000")
    rng = np.random.default_rng(43)
    b = small_band(256, 1100.0, 1500.0, 1e-4)
    x = noise(rng, 5000, 256)
    kw = dict(method=method, nout=777)
    if method == "subband":
        kw.update(sbw=32, ndm0=4)
    spec = make_spec(x, b, 123.0, 1.0, 1, **kw)
    got, ref, *_ = run_both(x, spec)
    assert got.shape == (1, 777)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_constant_input_no_flags_and_a_zero_plane(gpu_lib, dtype):
    say(f"case: constant {np.dtype(dtype).name} input, IQRM + zero-DM, subband")
    b = small_band(256, 1200.0, 1500.0, 1e-4)
    x = np.full((4000, 256), 100, dtype)
    spec = make_spec(x, b, 0.0, 1.0, 40, method="subband", sbw=32, ndm0=8,
                     use_iqrm=True, iqrm_c0=0, iqrm_c1=255, iqrm_mode=1,
                     use_zero_dm=True, zdm_c0=0, zdm_c1=256)
    got, ref, masks, rmask = run_both(x, spec)
    assert not masks.any() and not rmask.any()
    assert not got.any()
    assert np.array_equal(got, ref)


def test_invalid_descriptions_are_rejected(gpu_lib):
    say("case: plan validation messages")
    b = small_band(64, 1200.0, 1500.0, 1e-4)
    x = np.zeros((1000, 64), np.uint8)
    bad = [
        (dict(ndm=0), "ndm must be >= 1"),
        (dict(dd_c0=10, dd_c1=10), "invalid dedispersion channel range"),
        (dict(method="subband", sbw=0, ndm0=8), "subband method needs"),
        (dict(nout=5000), "nout exceeds"),
        (dict(use_iqrm=True, iqrm_c0=0, iqrm_c1=63, iqrm_mode=2), "IQRM mode"),
        (dict(f_ref=1300.0), "negative delay"),
    ]
    for kw, msg in bad:
        base = dict(method="direct")
        base.update(kw)
        ndm = base.pop("ndm", 10)
        spec = make_spec(x, b, 0.0, 1.0, ndm, **base)
        with pytest.raises(ValueError, match=msg):
            capi.Plan(spec)
    # a valid description still works afterwards
    capi.Plan(make_spec(x, b, 0.0, 1.0, 10, method="direct")).close()


def test_profiling_modes(gpu_lib):
    say("case: stage timing modes 1 (every stage) and 2 (the dedispersion stage only)")
    rng = np.random.default_rng(44)
    b = small_band(256, 1200.0, 1500.0, 1e-4)
    x = noise(rng, 4000, 256)
    spec = make_spec(x, b, 0.0, 1.0, 40, method="subband", sbw=32, ndm0=8,
                     use_iqrm=True, iqrm_c0=0, iqrm_c1=255, use_zero_dm=True, zdm_c0=0, zdm_c1=256)
    plan = capi.Plan(spec)
    d_in = capi.DeviceBuffer(x.nbytes)
    d_in.upload(x)
    d_out = capi.DeviceBuffer(int(np.prod(plan.out_shape)) * 4)
    plan.execute(d_in.ptr, d_out.ptr)
    capi.sync()
    ref = d_out.download(np.empty(plan.out_shape, np.uint32))
    for mode in (1, 2):
        plan.set_profiling(mode)
        for _ in range(3):
            plan.execute(d_in.ptr, d_out.ptr)
        t = plan.stage_times()
        assert t["n"] == 3 and t["dedisp"] > 0
        if mode == 1:
            assert t["stats"] > 0 and t["clean"] > 0 and t["stage2"] > 0 and t["total"] > 0
        else:
            assert t["stats"] == 0 and t["clean"] == 0 and t["stage2"] == 0
        # timing does not change the results
        assert np.array_equal(d_out.download(np.empty(plan.out_shape, np.uint32)), ref)
    with pytest.raises(ValueError, match="profiling mode"):
        plan.set_profiling(3)
    plan.set_profiling(0)
    plan.close()

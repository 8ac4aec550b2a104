"""Host-buffer calls whose rows (and plane) lie in page-locked memory from
af_host_alloc go to and from the device without staging; the results must
be those of the staged path and of the oracle, also for a row range inside a
larger page-locked block."""
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


def test_page_locked_rows_equal_staged_and_oracle(gpu_lib):
    say("case: subband u8 1024ch, 12000 rows, page-locked vs pageable host buffers")
    rng = np.random.default_rng(91)
    b = small_band(1024, 1100.0, 1500.0)
    x = np.clip(np.round(rng.normal(128, 16, (12000, 1024))), 0, 255).astype(np.uint8)
    synth.inject_narrowband_rfi(x, rng, frac=0.05, amp=40, n_stuck=4, pmax=512)
    spec = make_spec(x, b, 0.0, 1.0, 300, method="subband", sbw=64, ndm0=16,
                     use_iqrm=True, iqrm_c0=0, iqrm_c1=1023, iqrm_mode=1,
                     use_zero_dm=True, zdm_c0=0, zdm_c1=1024)
    ref, rmask, *_ = O.run_pipeline(x, oracle_spec(spec), O.IQRM_GPU)
    plan = capi.Plan(spec)
    got, masks = plan.run_host(x, want_masks=True)
    assert np.array_equal(got, ref) and np.array_equal(masks, rmask)
    # the rows at an offset inside a larger page-locked block
    hb = capi.HostBuffer((x.shape[0] + 700, x.shape[1]), np.uint8)
    hb.array[:] = 7
    rows = hb.array[300:300 + x.shape[0]]
    rows[:] = x
    got_p, masks_p = plan.run_host(rows, want_masks=True)
    assert np.array_equal(got_p, ref) and np.array_equal(masks_p, rmask)
    # page-locked plane as well (both transfers direct)
    ob = capi.HostBuffer(plan.out_shape, plan.out_dtype)
    capi.check(capi.lib().af_execute_host(plan._h, capi._ptr(rows), capi._ptr(ob.array), None))
    assert np.array_equal(ob.array, ref)
    # images made on the device from either buffer
    sps = 1024
    assert np.array_equal(plan.run_host_images(x, sps), plan.run_host_images(rows, sps))
    # the block's other rows are untouched
    assert (hb.array[:300] == 7).all() and (hb.array[300 + x.shape[0]:] == 7).all()
    plan.close()
    hb.free()
    ob.free()

"""af_execute_noalloc (no counterpart in the reference): the whole call on
the library's static device scratch, no allocation.  Through the C-ABI, two
calls in a row (the second on a smaller input, so stale scratch contents of
the first must not leak into it) equal af_execute of a plan created for
each description: planes and IQRM masks bit for bit.  A description that
needs more than the scratch holds fails with MemoryError."""
import sys

import numpy as np
import pytest

from astroflow_amd import capi, synth
from tests.helpers import make_spec, small_band
from tests.test_gpu_parity import noise

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def cases():
    rng = np.random.default_rng(71)
    b1 = small_band(512, 1100.0, 1500.0, 1e-4)
    x1 = noise(rng, 6000, 512)
    synth.inject_narrowband_rfi(x1, rng, frac=0.05, amp=40, n_stuck=3, pmax=512)
    s1 = make_spec(x1, b1, 0.0, 1.0, 130, method="subband", sbw=64, ndm0=16,
                   use_iqrm=True, iqrm_c0=0, iqrm_c1=511, iqrm_mode=1, iqrm_radius_frac=0.02,
                   iqrm_W=2500, iqrm_H=1500, use_zero_dm=True, zdm_c0=0, zdm_c1=511,
                   bad_channels=[7, 100])
    b2 = small_band(256, 1200.0, 1500.0, 1e-4)
    x2 = noise(rng, 3000, 256)
    synth.inject_narrowband_rfi(x2, rng, frac=0.05, amp=40, n_stuck=2, pmax=256)
    s2 = make_spec(x2, b2, 2.0, 0.5, 70, method="subband", sbw=32, ndm0=8,
                   use_iqrm=True, iqrm_c0=0, iqrm_c1=255, iqrm_mode=1,
                   use_zero_dm=True, zdm_c0=0, zdm_c1=255)
    b3 = small_band(300, 1200.0, 1500.0, 1e-4)
    x3 = noise(rng, 5000, 300, np.float32)
    s3 = make_spec(x3, b3, 0.0, 3.0, 40, method="direct", ds=2, use_zero_dm=True,
                   zdm_c0=0, zdm_c1=300)
    x4 = noise(rng, 2000, 300, np.float32)
    s4 = make_spec(x4, b3, 0.0, 5.0, 20, method="direct", ds=2, use_zero_dm=True,
                   zdm_c0=0, zdm_c1=300)
    return [((x1, s1), (x2, s2)), ((x3, s3), (x4, s4))]


def run_pair(pair, stream):
    results = []
    for x, spec in pair:
        plan = capi.Plan(spec)
        di = capi.DeviceBuffer(x.nbytes)
        di.upload(x)
        nb = plan.info.ndm * plan.info.nout * plan.info.out_elem_bytes
        d_ref, d_got = capi.DeviceBuffer(nb), capi.DeviceBuffer(nb)
        m_ref = capi.DeviceBuffer(max(1, plan.nwin * spec.nchan))
        m_got = capi.DeviceBuffer(max(1, plan.nwin * spec.nchan))
        plan.execute(di.ptr, d_ref.ptr, m_ref.ptr if plan.nwin else 0)
        capi.sync()
        ref = d_ref.download(np.empty(plan.out_shape, plan.out_dtype))
        rmask = m_ref.download(np.empty((plan.nwin, spec.nchan), np.uint8)) if plan.nwin else None
        results.append((spec, plan, di, d_got, m_got, ref, rmask, d_ref, m_ref))
    # the two no-allocation calls back to back on one stream
    for spec, plan, di, d_got, m_got, *_ in results:
        capi.execute_noalloc(spec, di.ptr, d_got.ptr, m_got.ptr if plan.nwin else 0, stream)
    capi.check(capi.lib().af_stream_sync(capi.C.c_void_p(stream) if stream else None))
    for k, (spec, plan, di, d_got, m_got, ref, rmask, d_ref, m_ref) in enumerate(results):
        got = d_got.download(np.empty(plan.out_shape, plan.out_dtype))
        assert np.array_equal(got, ref), f"call {k}: plane differs from af_execute"
        if rmask is not None:
            gm = m_got.download(np.empty(rmask.shape, np.uint8))
            assert rmask.sum() > 0
            assert np.array_equal(gm, rmask), f"call {k}: masks differ from af_execute"
        plan.close()
        for b in (di, d_got, m_got, d_ref, m_ref):
            b.free()


@pytest.mark.parametrize("which", [0, 1])
def test_noalloc_equals_execute(gpu_lib, which):
    say(f"case: af_execute_noalloc pair {which}, null stream, then a created stream")
    pair = cases()[which]
    run_pair(pair, 0)
    st = capi.C.c_void_p()
    capi.check(capi.lib().af_stream_create(capi.C.byref(st)))
    run_pair(pair, st.value)
    capi.check(capi.lib().af_stream_destroy(st))


def test_noalloc_rejects_oversize(gpu_lib):
    say("case: af_execute_noalloc beyond its scratch")
    b = small_band(2048, 1000.0, 1500.0, 1e-4)
    x = np.zeros((40000, 2048), np.uint8)
    spec = make_spec(x, b, 0.0, 1.0, 200, method="subband", sbw=64, ndm0=16)
    di = capi.DeviceBuffer(x.nbytes)
    do = capi.DeviceBuffer(4)
    with pytest.raises(MemoryError, match="no|scratch"):
        capi.execute_noalloc(spec, di.ptr, do.ptr)
    di.free()
    do.free()

"""Batched execution: plans of different nchan / nsamp / DM range / dtype /
method run with one merged launch per stage for their small launches (a
launch that fills the GPU runs alone), replayed from a hipGraph, or each
with its own launches spread over a few streams (also replayed from a
graph); every output equals the plan's own execute bit for bit; the
autotuner picks only result-invariant settings unless asked to tune the
subband width."""
import numpy as np
import pytest

from astroflow_amd import capi
from tests.helpers import make_spec, small_band
from tests.test_gpu_parity import noise

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def batch_cases():
    rng = np.random.default_rng(31)
    cases = []
    b1 = small_band(256, 1200.0, 1500.0, 1e-4)
    x = noise(rng, 6000, 256)
    cases.append((x, make_spec(x, b1, 0.0, 1.0, 90, method="subband", sbw=32, ndm0=16,
                               use_iqrm=True, iqrm_c0=0, iqrm_c1=255, iqrm_mode=1,
                               iqrm_radius_frac=0.03, iqrm_W=2000, iqrm_H=1000,
                               use_zero_dm=True, zdm_c0=0, zdm_c1=255, bad_channels=[5])))
    b2 = small_band(128, 1100.0, 1500.0, 2e-4)
    x = noise(rng, 3000, 128)
    cases.append((x, make_spec(x, b2, 2.0, 3.0, 40, method="direct", use_zero_dm=True,
                               zdm_c0=0, zdm_c1=128)))
    b3 = small_band(512, 1000.0, 1500.0, 1e-4)
    x = noise(rng, 9000, 512)
    cases.append((x, make_spec(x, b3, 0.0, 0.5, 120, method="subband", sbw=64, ndm0=32, ds=2,
                               use_iqrm=True, iqrm_c0=0, iqrm_c1=511, iqrm_mode=0,
                               use_zero_dm=True, zdm_c0=0, zdm_c1=511)))
    x = noise(rng, 4000, 256, np.float32)
    cases.append((x, make_spec(x, b1, 0.0, 2.0, 50, method="subband", sbw=32, ndm0=8,
                               use_zero_dm=True, zdm_c0=0, zdm_c1=256)))
    x = noise(rng, 5000, 256, np.uint16)
    cases.append((x, make_spec(x, b1, 0.0, 1.0, 60, method="subband", sbw=32, ndm0=16,
                               use_iqrm=True, iqrm_c0=0, iqrm_c1=255, iqrm_mode=1)))
    # time super-tiles: several stage-1/stage-2 launches of one plan
    x = noise(rng, 20000, 384)
    b4 = small_band(384, 1100.0, 1500.0, 1e-4)
    cases.append((x, make_spec(x, b4, 0.0, 1.0, 150, method="subband", sbw=64, ndm0=16,
                               nout=17000, use_zero_dm=True, zdm_c0=0, zdm_c1=384,
                               inter_budget=10 * 6 * 2 * 4096)))
    # a plan whose stage-1 launch fills the GPU on its own (4096 channels in
    # 32-channel subbands: over 2048 workgroups), so it runs unmerged
    b5 = small_band(4096, 1000.0, 1500.0, 1e-4)
    x = noise(rng, 12000, 4096)
    cases.append((x, make_spec(x, b5, 0.0, 1.0, 256, method="subband", sbw=32, ndm0=16,
                               use_zero_dm=True, zdm_c0=0, zdm_c1=4095)))
    # a second plan of the first shape (same kernels, merged with it)
    x = noise(rng, 6000, 256)
    cases.append((x, make_spec(x, b1, 0.0, 1.0, 90, method="subband", sbw=32, ndm0=16,
                               use_iqrm=True, iqrm_c0=0, iqrm_c1=255, iqrm_mode=1,
                               iqrm_radius_frac=0.03, use_zero_dm=True, zdm_c0=0, zdm_c1=255)))
    return cases


def test_batched_outputs_equal_single_executes(gpu_lib):
    say("case: batch of 7 heterogeneous plans, merged launches and graph replays")
    cases = batch_cases()
    plans, d_in, d_out, d_m, ref = [], [], [], [], []
    for x, spec in cases:
        p = capi.Plan(spec)
        plans.append(p)
        di = capi.DeviceBuffer(x.nbytes)
        di.upload(x)
        shape = p.out_shape
        do = capi.DeviceBuffer(int(np.prod(shape)) * np.dtype(p.out_dtype).itemsize)
        dm = capi.DeviceBuffer(max(1, p.nwin * spec.nchan))
        p.execute(di.ptr, do.ptr, dm.ptr if p.nwin else 0)
        capi.sync()
        r = do.download(np.empty(shape, p.out_dtype))
        rm = dm.download(np.empty((p.nwin, spec.nchan), np.uint8)) if p.nwin else None
        ref.append((r, rm))
        d_in.append(di)
        d_out.append(do)
        d_m.append(dm)
    batch = capi.Batch(plans, [b.ptr for b in d_in], [b.ptr for b in d_out],
                       [b.ptr if p.nwin else 0 for b, p in zip(d_m, plans)])
    info = batch.info()
    say(f"launches {info}")
    assert info["launches"] < info["recorded"]
    # merged launches, their graph replayed three times; then the plans'
    # own launches on 4 streams, that fork/join graph replayed three times,
    # recaptured for 2 streams, and on 1 stream
    modes = [("merged", False, 0), ("graph", True, 0), ("graph", True, 0), ("graph", True, 0),
             ("streams 4", False, 4), ("streams 4 graph", True, 4), ("streams 4 graph", True, 4),
             ("streams 4 graph", True, 4), ("streams 2 graph", True, 2), ("streams 1", False, 1)]
    for mode, use_graph, nst in modes:
        for b in d_out:
            capi.check(capi.lib().af_memset(capi.C.c_void_p(b.ptr), 0, b.nbytes, None))
        for b, pl in zip(d_m, plans):
            if pl.nwin:
                capi.check(capi.lib().af_memset(capi.C.c_void_p(b.ptr), 0, b.nbytes, None))
        capi.sync()
        batch.execute(use_graph=use_graph, streams=nst)
        capi.sync()
        for i, (p, (r, rm)) in enumerate(zip(plans, ref)):
            got = d_out[i].download(np.empty(p.out_shape, p.out_dtype))
            assert np.array_equal(got, r), (mode, i)
            if rm is not None:
                gm = d_m[i].download(np.empty(rm.shape, np.uint8))
                assert np.array_equal(gm, rm), (mode, i)
        # host work between replays
        _ = sum(range(10000))
    with pytest.raises(ValueError, match="nstreams"):
        batch.execute(streams=9)
    batch.close()
    for p in plans:
        p.close()


def test_autotune_keeps_results(gpu_lib):
    say("case: autotuner on a u8 subband plan")
    x, spec = batch_cases()[0]
    p = capi.Plan(spec)
    di = capi.DeviceBuffer(x.nbytes)
    di.upload(x)
    do = capi.DeviceBuffer(int(np.prod(p.out_shape)) * 4)
    p.execute(di.ptr, do.ptr)
    capi.sync()
    ref = do.download(np.empty(p.out_shape, np.uint32))
    p.close()
    tuned, ms = capi.autotune(spec, di.ptr, do.ptr, tune_subbands=False, reps=2)
    say(f"tuned tile_t={tuned.tile_t} dm_per_block={tuned.dm_per_block} sbw={tuned.sbw} {ms:.3f} ms")
    assert ms > 0 and tuned.sbw == spec.sbw
    assert tuned.dm_per_block in (0, 16) and tuned.tile_t in (0, 256, 512, 1024)
    p2 = capi.Plan(tuned)
    p2.execute(di.ptr, do.ptr)
    capi.sync()
    assert np.array_equal(do.download(np.empty(p2.out_shape, np.uint32)), ref)
    p2.close()
    # cached: the same shape returns at once with the same answer
    t2, ms2 = capi.autotune(spec, di.ptr, do.ptr, tune_subbands=False, reps=2)
    assert (t2.tile_t, t2.dm_per_block) == (tuned.tile_t, tuned.dm_per_block) and ms2 == ms
    tsb, _ = capi.autotune(spec, di.ptr, do.ptr, tune_subbands=True, reps=1)
    assert tsb.sbw in (16, 32, 64, 128)

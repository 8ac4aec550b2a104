"""The batched workload at its full sizes: the 48 seeded beam-chunks of
synth.c3_chunks(seed=4) (nchan 1024/2048/4096, nsamp log-uniform 2^14-2^20,
u8 or float32, DM 0-500/2000/5000 step 1 or 0.5, ds 1/2/4, IQRM + zero-DM,
32-channel subbands), each with its own seeded input (noise plus narrowband
RFI), run through af_batch with hipGraph replay, four chunks per batch, both
as merged launches and on per-chunk streams (af_batch_execute_streams); the
two give the same rows and masks.

Per chunk:
  - every IQRM mask equals the oracle's, bit for bit;
  - the DM-0 row equals the per-sample channel sums of the oracle's cleaned
    data (numpy, independent of the oracle's dedispersion);
  - 16 DM rows spread over the range (first and last included) equal the
    oracle's rows bit for bit (u8), or lie within 1e-6 * sum|x| of the FP64
    oracle per output (float32; sum|x| over the cleaned samples each output
    adds).
Each batch is replayed twice from its graph with the planes overwritten in
between; the second replay is the one checked."""
import sys

import numpy as np
import pytest

from astroflow_amd import capi, synth
from oracle import oracle as O
from tests.helpers import oracle_spec

pytestmark = pytest.mark.gpu

PER_BATCH = 4
CHUNKS = synth.c3_chunks(seed=4)
NBATCH = (len(CHUNKS) + PER_BATCH - 1) // PER_BATCH


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def spread_rows(ndm, n=16):
    return sorted(set(np.linspace(0, ndm - 1, n).round().astype(int).tolist()))


def fetch_rows(ptr, rows, nout, dtype):
    es = np.dtype(dtype).itemsize
    out = np.empty((len(rows), nout), dtype)
    for i, m in enumerate(rows):
        capi.check(capi.lib().af_memcpy_d2h(capi._ptr(out[i]), capi.C.c_void_p(ptr + m * nout * es),
                                            nout * es, None))
    return out


@pytest.mark.timeout(600)
@pytest.mark.parametrize("batch", range(NBATCH))
def test_c3_batch_parity(gpu_lib, batch):
    import bench
    idx = list(range(batch * PER_BATCH, min(len(CHUNKS), (batch + 1) * PER_BATCH)))
    items = []
    for i in idx:
        c = CHUNKS[i]
        say(f"case: c3 chunk {i}: {c['nchan']} ch x {c['nsamp']} {np.dtype(c['dtype']).name}, "
            f"DM 0-{c['dm_max']:g}/{c['dm_step']:g}, ds {c['ds']}")
        x = synth.c3_chunk_data(c, seed=1000 + i, threads=16)
        spec = bench.c3_spec(c)
        plan = capi.Plan(spec)
        di = capi.DeviceBuffer(x.nbytes)
        di.upload(x)
        nout, ndm = plan.info.nout, plan.info.ndm
        do = capi.DeviceBuffer(ndm * nout * plan.info.out_elem_bytes)
        dm = capi.DeviceBuffer(max(1, plan.nwin * spec.nchan))
        items.append(dict(i=i, x=x, spec=spec, plan=plan, di=di, do=do, dm=dm))
    batch_h = capi.Batch([it["plan"] for it in items], [it["di"].ptr for it in items],
                         [it["do"].ptr for it in items], [it["dm"].ptr for it in items])
    say(f"batch {batch}: {batch_h.info()}")
    # per-chunk streams first (its graph captured and replayed): its sampled
    # rows and masks must equal the merged replay's, checked below
    batch_h.execute(use_graph=True, streams=4)
    capi.sync()
    by_streams = []
    for it in items:
        rows = spread_rows(it["plan"].info.ndm)
        by_streams.append((fetch_rows(it["do"].ptr, rows, it["plan"].info.nout, it["plan"].out_dtype),
                           it["dm"].download(np.empty((it["plan"].nwin, it["spec"].nchan), np.uint8))))
    for it in items:
        capi.check(capi.lib().af_memset(capi.C.c_void_p(it["do"].ptr), 0xA5, it["do"].nbytes, None))
        capi.check(capi.lib().af_memset(capi.C.c_void_p(it["dm"].ptr), 0xA5, it["dm"].nbytes, None))
    capi.sync()
    batch_h.execute(use_graph=True)          # capture + first replay
    capi.sync()
    for it in items:                          # overwrite: the replay below must rewrite all
        capi.check(capi.lib().af_memset(capi.C.c_void_p(it["do"].ptr), 0xA5, it["do"].nbytes, None))
        capi.check(capi.lib().af_memset(capi.C.c_void_p(it["dm"].ptr), 0xA5, it["dm"].nbytes, None))
    capi.sync()
    batch_h.execute(use_graph=True)
    capi.sync()
    for it, (s_rows, s_mask) in zip(items, by_streams):
        plan, spec, x = it["plan"], it["spec"], it["x"]
        nout, ndm = plan.info.nout, plan.info.ndm
        odt = plan.out_dtype
        rows = spread_rows(ndm)
        got = fetch_rows(it["do"].ptr, rows, nout, odt)
        gmask = it["dm"].download(np.empty((plan.nwin, spec.nchan), np.uint8))
        assert np.array_equal(got, s_rows) and np.array_equal(gmask, s_mask), \
            f"chunk {it['i']}: per-chunk streams differ from the merged replay"
        say(f"chunk {it['i']}: oracle ({len(rows)} rows)")
        s = oracle_spec(spec)
        ref, rmask, cleaned = O.run_pipeline_rows(x, s, rows, O.IQRM_GPU)
        assert np.array_equal(gmask, rmask), f"chunk {it['i']}: IQRM masks differ"
        assert rmask.sum() > 0, f"chunk {it['i']}: nothing flagged"
        sel = cleaned[:nout, s.dd_c0:s.dd_c1]
        if spec.dtype == np.float32:
            # DM 0 = channel sums; FP64 oracle tolerance per output
            d0 = sel.astype(np.float64).sum(1)
            tol0 = 1e-6 * np.abs(sel).astype(np.float64).sum(1)
            assert np.all(np.abs(got[0].astype(np.float64) - d0) <= tol0), f"chunk {it['i']} DM 0"
            s_abs = oracle_spec(spec)
            s_abs.use_iqrm = s_abs.use_zero_dm = False
            s_abs.bad_channels = []
            # sum|x| of the samples each output adds: the same rows over
            # |cleaned| (already at the binned rate)
            bound, _, _ = O.run_pipeline_rows(np.abs(cleaned), s_abs, rows, binned=True)
            err = np.abs(got.astype(np.float64) - ref)
            assert np.all(err <= 1e-6 * bound), \
                f"chunk {it['i']}: max err/bound {float((err / np.maximum(bound, 1e-30)).max()):.3g}"
        else:
            assert np.array_equal(got[0], sel.astype(np.uint64).sum(1).astype(got.dtype)), \
                f"chunk {it['i']}: DM-0 row differs from the channel sums"
            for k, m in enumerate(rows):
                assert np.array_equal(got[k], ref[k]), f"chunk {it['i']}: row {m} differs"
    batch_h.close()
    for it in items:
        it["plan"].close()
        for b in (it["di"], it["do"], it["dm"]):
            b.free()

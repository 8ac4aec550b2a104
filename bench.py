#!/usr/bin/env python3
"""Throughput of the MI355X dedispersion path on the 4096-channel u8 chunk.

Workload (one step = one plain call on one chunk, input resident in HBM):
  8-bit filterbank, 4096 channels over 1000-1500 MHz (top first, flipped to
  ascending), tsamp 49.152 us; a 0.5 s chunk of 10173 output samples plus the
  140618-sample DM-overlap tail (150791 input rows); N(128,16) noise, 5% of the
  channels with sinusoidal RFI (amplitude 40, periods 8-4096 samples), 10
  channels stuck at 255, pulses at DM 150/1200/2800; IQRM (std, radius 5
  channels, 3 sigma, one window), zero-DM; subband dedispersion with 32
  subbands of 128 channels, 32 fine DMs per coarse DM, DM 0-3000 step 0.5
  (6001 trials).

Prints one JSON line (see the repository's bench contract).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DM-trial×samples/s at 4096ch uint8; p99 chunk latency, batched/served vs plain"
UNIT = "DM-trial*samples/s"
HBM_PEAK_GBS = 8000.0
VALU_ADD_PEAK = 256 * 64 * 2.4e9     # 32-bit adds/s (one per lane per clock)
STAGE1_KERNEL = "k_dd_u8_q"
# HBM bytes per dispatch from the L2-to-memory request counters of this
# workload, calibrated on a copy of known size; one
# "traffic_bytes <kernel> <bytes>" line per kernel.
TRAFFIC_FILE = "profiles/hbm_traffic_4096ch_u8_subband_iqrm_zdm.md"


def profiled_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprof summary."""
    path = os.path.join(ROOT, TRAFFIC_FILE)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes" and parts[1] == kernel:
                return float(parts[2])
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--dm-per-block", type=int, default=0,
                    help="DM rows per stage-1 workgroup (0 = library default)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="target CPU work for the baseline sample")
    ap.add_argument("--no-served", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--served-full-rfi", action="store_true",
                    help="served leg without incremental RFI (every window cleaned whole)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched leg")
    ap.add_argument("--no-direct", action="store_true",
                    help="skip the direct-method leg (1024 ch x 2^18, DM 0-1000)")
    ap.add_argument("--batch-only", action="store_true", help="run the batched leg alone")
    ap.add_argument("--direct-only", action="store_true",
                    help="run the direct-method leg alone (1024 ch x 2^18, DM 0-1000), "
                         "GPU only, --steps executes")
    ap.add_argument("--served-only", action="store_true",
                    help="run the served leg alone (no plain-call line)")
    ap.add_argument("--served-chunks", type=int, default=48,
                    help="timed chunks per beam at each arrival rate of the served leg "
                         "(19 beams: 912 latency samples per rate)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one 256 MiB device-to-device copy (a known-size kernel "
                         "for scaling hardware traffic counters)")
    return ap.parse_args()


def c2_spec(band, nsamp, nout):
    from astroflow_amd import capi
    f = band.freq
    nchan = band.nchan
    radius_frac = float(np.float32(5.5 / 4095))   # floor(radius_frac * 4095) = 5 channels
    return capi.DedispSpec(
        dtype=np.uint8, nchan=nchan, nsamp=nsamp, freq=f, tsamp=float(band.tsamp),
        dm_low=0.0, dm_step=0.5, ndm=6001, dd_c0=0, dd_c1=nchan, f_ref=float(f[-1]),
        ds=1, nout=nout, method="subband", sbw=128, ndm0=32,
        use_iqrm=True, iqrm_c0=0, iqrm_c1=nchan - 1, iqrm_mode=1,
        iqrm_radius_frac=radius_frac, iqrm_nsigma=3.0, iqrm_geofactor=1.5,
        iqrm_W=0, iqrm_H=0, iqrm_include_tail=True,
        use_zero_dm=True, zdm_c0=0, zdm_c1=nchan - 1)


def make_workload(threads=8):
    """Seeded synthetic chunk on the host; returns (x, spec, info)."""
    from astroflow_amd import synth
    x, band, nout, dmax = synth.chunk_4096(seed=2, threads=threads)
    spec = c2_spec(band, x.shape[0], nout)
    info = dict(nchan=band.nchan, nout=nout, dmax=dmax, nsamp=x.shape[0], ndm=6001, ds=1)
    return x, spec, info


def host_cpu():
    """Threads for the CPU baseline (every core this process may run on, as
    nproc counts them, within the cgroup's CPU quota when one is set) and the
    CPU model."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = max(1, int(math.ceil(int(q) / int(per))))
    except (OSError, ValueError):
        pass
    threads = min(cores, quota) if quota else cores
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"threads": int(threads), "nproc": int(cores), "cgroup_quota": quota, "model": model}


def cpu_baseline(x_host: np.ndarray, info, target_s: float):
    """The reference CPU path (cpucal direct loop nest, OpenMP) timed on a DM
    subset of the same chunk; returns the rate in DM-trial*samples/s."""
    from astroflow_amd import synth
    from oracle import oracle as O   # CPU baseline leg: the oracle's cpucal restatement

    band = synth.band_1000_1500(info["nchan"])
    f = band.freq
    nchan, nout = info["nchan"], info["nout"]
    hc = host_cpu()
    threads = hc["threads"]
    native = O.cpucal_native()     # built on this host: -Ofast -march=native -fopenmp
    all_dms = np.arange(info["ndm"]) * 0.5

    def run(k):
        idx = np.linspace(0, info["ndm"] - 1, k).round().astype(np.int64)
        # cpucal excludes the top channel: channels [chan_start, chan_end)
        d = O.delay_table(f, 0, nchan - 1, f[-1], float(band.tsamp), 1, all_dms[idx])
        t0 = time.perf_counter()
        O.cpucal_direct_u8(x_host, 0, nchan - 1, d, 0, nout, 1, nthreads=threads, lib=native)
        return time.perf_counter() - t0

    k = 8
    dt = run(k)
    k2 = int(min(info["ndm"], max(8, round(k * target_s / max(dt, 1e-3)))))
    reps = []
    for _ in range(3):
        reps.append(run(k2))
        if sum(reps) > 2.5 * target_s:
            break
    med = float(np.median(reps))
    rate = k2 * nout / med
    return {"value": rate, "unit": UNIT, "cores": threads, "kind": "port",
            "cpu_model": hc["model"], "nproc": hc["nproc"], "cgroup_cpu_quota": hc["cgroup_quota"],
            "build": "gcc -Ofast -march=native -fopenmp on this host (CMakeLists.txt:118-119 flags)",
            "sample": f"{k2} of {info['ndm']} DM trials (evenly spaced over 0-3000) x {nout} "
                      f"samples x {nchan - 1} channels, cpucal direct loop nest "
                      f"(include/cpucal.hpp:167-186), u64 accumulation, {threads} OpenMP threads, "
                      f"median of {len(reps)} runs, {med:.2f} s each; rate scaled linearly to "
                      f"all {info['ndm']} trials"}


def served_leg(threads, chunks_per_beam, full_rfi=False):
    """19 beams x 4096 ch u8, chunk durations 0.25/0.5/1.0 s per beam, DM
    0-5000 step 1 (5001 trials), 32 subbands, IQRM + zero-DM; Poisson
    arrivals at 1x, 4x and 16x real time, then all at once (backlogged)."""
    from astroflow_amd import capi, stream, synth
    t_setup = time.perf_counter()
    beams = stream.make_beams()
    band = synth.band_1000_1500(4096)
    srv = capi.Server(full_rfi=full_rfi)
    for b in beams:
        srv.add_beam(b.spec, b.chunk_rows)
    pool = stream.RowPool(4096, 65536, seed=5, threads=threads)
    rng = np.random.default_rng(7)
    stream.prefill(srv, beams, pool, rng)
    log(f"served leg set up in {time.perf_counter() - t_setup:.1f} s")
    rates = {}
    for i, r in enumerate((1.0, 4.0, 16.0, 0.0)):
        res = stream.run_rate(srv, beams, pool, r, chunks_per_beam, seed=100 + i, band=band)
        rates["backlogged" if r == 0 else f"{r:g}x"] = res
        log(f"served rate {r:g}: p50 {res['p50_ms']:.2f} ms p99 {res['p99_ms']:.2f} ms "
            f"throughput {res['throughput']:.3e}")
    # the same backlog with the new rows already in HBM: the device-side
    # ceiling of the served path, without the PCIe upload
    dpool = stream.RowPool(4096, 65536, seed=6, threads=threads, device=True)
    res = stream.run_rate(srv, beams, dpool, 0.0, chunks_per_beam, seed=200, band=band)
    rates["backlogged_rows_in_hbm"] = res
    log(f"served backlog from HBM: p50 {res['p50_ms']:.2f} ms p99 {res['p99_ms']:.2f} ms "
        f"throughput {res['throughput']:.3e}")
    dpool.free()
    st = srv.stats()
    srv.close()
    # host-to-device rate of the new rows (page-locked), for the PCIe ceiling
    db = capi.DeviceBuffer(pool.buf.array.nbytes)
    db.upload(pool.buf.array)
    capi.sync()
    t0 = time.perf_counter()
    for _ in range(3):
        db.upload(pool.buf.array)
    capi.sync()
    h2d_gbs = 3 * pool.buf.array.nbytes / (time.perf_counter() - t0) / 1e9
    db.free()
    pool.free()
    # the plain single-call GPU path on one 0.5 s window of the same stream:
    # whole window uploaded from host memory, plane downloaded (the reference's
    # per-chunk pattern), and the same call on a window already in HBM
    b = min(beams, key=lambda bb: abs(bb.chunk_s - 0.5))
    plan = capi.Plan(b.spec)
    xw = synth.noise_u8_threaded(9, b.window_rows, 4096, threads=threads)
    out = np.empty((plan.info.ndm, plan.info.nout), np.uint32)
    def median_ms(fn, reps=3):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(ts))

    host_ms = median_ms(lambda: plan.run_host(xw))
    # the drop-in call of the search loop (_dedisperse_spec): host rows in,
    # DM-time images of t_sample = one chunk out, the plane imaged on the
    # device and never downloaded
    images_ms = median_ms(lambda: plan.run_host_images(xw, b.chunk_rows))
    d_in = capi.DeviceBuffer(xw.nbytes)
    d_in.upload(xw)
    d_out = capi.DeviceBuffer(out.nbytes)
    plan.execute(d_in.ptr, d_out.ptr)
    capi.sync()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.execute(d_in.ptr, d_out.ptr)
        capi.sync()
    dev_ms = (time.perf_counter() - t0) * 1e3 / reps
    plan.close()
    d_in.free()
    d_out.free()
    work = b.spec.ndm * b.chunk_rows
    host_best = rates["backlogged"]["throughput"]
    dev_best = rates["backlogged_rows_in_hbm"]["throughput"]
    return {
        "workload": "19 beams x 4096 ch u8 1000-1500 MHz, chunks of 0.25/0.5/1.0 s per beam "
                    "(seeded), DM 0-5000 step 1 (5001 trials), 32 subbands x 128 ch, IQRM std "
                    "radius 5 + zero-DM; DM-overlap carry resident in HBM; Poisson arrivals",
        "rates": rates,
        "hbm_peak_GB": st["hbm_peak"] / 1e9,
        "h2d_pinned_GBps": h2d_gbs,
        "plain_single_call": {"chunk_s": b.chunk_s, "host_buffers_ms": host_ms,
                              "host_buffers_images_ms": images_ms,
                              "device_resident_ms": dev_ms,
                              "host_buffers_rate": work / (host_ms / 1e3),
                              "device_resident_rate": work / (dev_ms / 1e3)},
        # served with host rows vs the plain call on host buffers; served with
        # rows in HBM vs the plain call on a window already in HBM
        "speedup_vs_plain_host_call": host_best / (work / (host_ms / 1e3)),
        "speedup_vs_plain_device_call": dev_best / (work / (dev_ms / 1e3)),
    }


def c3_spec(c):
    """One chunk of the batched workload: 1000-1500 MHz, tsamp 49.152 us, DM
    0..dm_max, subband (32-channel subbands, 32 fine DMs per coarse DM, the
    reference defaults), IQRM (std, radius 5 channels, 3 sigma, whole
    buffer) + zero-DM, full output length."""
    from astroflow_amd import capi, synth
    band = synth.band_1000_1500(c["nchan"])
    f = band.freq
    nchan = c["nchan"]
    ndm = int(round(c["dm_max"] / c["dm_step"])) + 1
    return capi.DedispSpec(
        dtype=c["dtype"], nchan=nchan, nsamp=c["nsamp"], freq=f, tsamp=float(band.tsamp),
        dm_low=0.0, dm_step=c["dm_step"], ndm=ndm, dd_c0=0, dd_c1=nchan, f_ref=float(f[-1]),
        ds=c["ds"], nout=0, method="subband", sbw=32, ndm0=32,
        use_iqrm=True, iqrm_c0=0, iqrm_c1=nchan - 1, iqrm_mode=1,
        iqrm_radius_frac=float(np.float32(5.5 / (nchan - 1))), iqrm_nsigma=3.0,
        iqrm_geofactor=1.5, iqrm_W=0, iqrm_H=0, iqrm_include_tail=True,
        use_zero_dm=True, zdm_c0=0, zdm_c1=nchan - 1)


BATCH_STREAMS = 4   # side streams of the streams mode (the device's hardware queues per process)


def batch_leg(threads, budget_gb=40.0, reps=5):
    """The 48 seeded beam-chunks (uneven nchan, nsamp, DM range, dtype, ds),
    each with IQRM + zero-DM and its own seeded input (synth.c3_chunk_data,
    seed 1000 + i, the inputs of the batched parity test), resident in HBM.
    Chunks are grouped under an HBM budget.  Per group, each mode runs after
    a warm-up `reps` times and its median counts:
      plain        one execute per chunk, default settings (the reference's
                   per-call pattern);
      plain_tuned  the same with each chunk's settings from af_autotune
                   (stage-1 form and stage-2 tile; the plane is unchanged);
      merged       the tuned plans as one af_batch: one launch per stage and
                   kernel for the whole group;
      graph        that batch replayed from its hipGraph;
      streams      the same batch with each plan's own launches, the plans
                   spread over 4 streams (af_batch_execute_streams), so that
                   different chunks' stages overlap; replayed from its graph.
    device_ms is the sum of the chunks' per-stage HIP-event times (the GPU's
    busy time); wall time above it is what merging launches can remove."""
    from astroflow_amd import capi, synth
    chunks = synth.c3_chunks(seed=4)
    budget = budget_gb * 1e9
    modes = ("plain", "plain_tuned", "merged", "graph", "streams")
    tot = {m: 0.0 for m in modes}
    spread = {m: [] for m in modes}
    device_s = 0.0
    work = 0
    launches = recorded = 0
    ngroups = 0
    tuned_settings = []
    t_gen = 0.0
    i = 0
    t_setup = time.perf_counter()

    def timed(fn):
        fn()
        capi.sync()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            capi.sync()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), ts

    while i < len(chunks):
        group = []
        used = 0.0
        while i < len(chunks):
            spec = c3_spec(chunks[i])
            plan = capi.Plan(spec)
            es = np.dtype(spec.dtype).itemsize
            need = (spec.nsamp * spec.nchan * es + plan.info.scratch_bytes +
                    plan.info.ndm * plan.info.nout * plan.info.out_elem_bytes)
            if group and used + need > budget:
                plan.close()
                break
            group.append((i, spec, plan))
            used += need
            i += 1
        ngroups += 1
        ins, outs, tplans = [], [], []
        for k, spec, plan in group:
            t0 = time.perf_counter()
            x = synth.c3_chunk_data(chunks[k], seed=1000 + k, threads=threads)
            t_gen += time.perf_counter() - t0
            di = capi.DeviceBuffer(x.nbytes)
            di.upload(x)
            del x
            do = capi.DeviceBuffer(plan.info.ndm * plan.info.nout * plan.info.out_elem_bytes)
            ins.append(di)
            outs.append(do)
            work += plan.info.ndm * plan.info.nout * spec.ds
            tspec, _ = capi.autotune(spec, di.ptr, do.ptr, tune_subbands=False, reps=2)
            tuned_settings.append((tspec.dm_per_block, tspec.tile_t))
            tplans.append(capi.Plan(tspec))
        capi.sync()
        plans = [p for _, _, p in group]

        def run(ps):
            def f():
                for p, di, do in zip(ps, ins, outs):
                    p.execute(di.ptr, do.ptr)
            return f
        med = {}
        med["plain"], ts = timed(run(plans))
        spread["plain"] += ts
        med["plain_tuned"], ts = timed(run(tplans))
        spread["plain_tuned"] += ts
        batch = capi.Batch(tplans, [b.ptr for b in ins], [b.ptr for b in outs])
        inf = batch.info()
        launches += inf["launches"]
        recorded += inf["recorded"]
        med["merged"], ts = timed(lambda: batch.execute(use_graph=False))
        spread["merged"] += ts
        med["graph"], ts = timed(lambda: batch.execute(use_graph=True))
        spread["graph"] += ts
        med["streams"], ts = timed(lambda: batch.execute(use_graph=True, streams=BATCH_STREAMS))
        spread["streams"] += ts
        batch.close()
        # GPU busy time of the tuned plans, from their stage events
        for p, di, do in zip(tplans, ins, outs):
            p.set_profiling(True)
            for _ in range(2):
                p.execute(di.ptr, do.ptr)
            capi.sync()
            device_s += p.stage_times()["total"] / 1e3
            p.set_profiling(False)
        for m in modes:
            tot[m] += med[m]
        for p in plans + tplans:
            p.close()
        for b in ins + outs:
            b.free()
        log(f"batch group {ngroups}: {len(group)} chunks, {used / 1e9:.1f} GB; medians "
            + " ".join(f"{m} {med[m] * 1e3:.1f} ms" for m in modes))
    res = {
        "workload": "48 seeded beam-chunks: nchan 1024/2048/4096, nsamp log-uniform 2^14-2^20, "
                    "u8 or f32, DM 0-500/2000/5000 step 1 or 0.5, ds 1/2/4; IQRM + zero-DM; "
                    "32-channel subbands; full-length planes; each chunk its own seeded input "
                    "(noise + narrowband RFI), resident in HBM",
        "chunks": len(chunks), "groups": ngroups, "hbm_budget_GB": budget_gb, "reps": reps,
        "work": work, "launches_batched": launches, "launches_plain": recorded,
        "tuned_settings": {f"dm_per_block {a} tile {b}": tuned_settings.count((a, b))
                           for a, b in sorted(set(tuned_settings))},
        "device_s": device_s,
        "input_generation_s": t_gen,
        "setup_s": time.perf_counter() - t_setup - sum(sum(v) for v in spread.values()),
    }
    for m in modes:
        res[f"{m}_s"] = tot[m]
        res[f"{m}_rate"] = work / tot[m]
        res[f"{m}_rep_s_min_max"] = [min(spread[m]), max(spread[m])]
    res["speedup_tuned_vs_plain"] = tot["plain"] / tot["plain_tuned"]
    res["speedup_merged_vs_plain"] = tot["plain"] / tot["merged"]
    res["speedup_graph_vs_plain"] = tot["plain"] / tot["graph"]
    res["speedup_graph_vs_plain_tuned"] = tot["plain_tuned"] / tot["graph"]
    res["speedup_streams_vs_plain"] = tot["plain"] / tot["streams"]
    res["streams"] = BATCH_STREAMS
    # the most any launch merging could give: plain-tuned wall time over GPU
    # busy time
    res["ceiling_launch_merging"] = tot["plain_tuned"] / device_s if device_s > 0 else None
    return res


def direct_leg(threads, cpu_target_s, reps=10, with_cpu=True):
    """The plain direct-method case, on the GPU and with the reference's CPU
    loop nest on the same input: 1024 ch u8 x 2^18 samples over 1000-1500
    MHz, one dispersed pulse (DM 500), DM 0-1000 step 1 (1001 trials),
    zero-DM, full-length u32 plane."""
    from astroflow_amd import capi, synth
    x, band, _ = synth.direct_1024(seed=1, threads=threads)
    nsamp, nchan = x.shape
    f = band.freq
    spec = capi.DedispSpec(dtype=np.uint8, nchan=nchan, nsamp=nsamp, freq=f,
                           tsamp=float(band.tsamp), dm_low=0.0, dm_step=1.0, ndm=1001,
                           dd_c0=0, dd_c1=nchan, f_ref=float(f[-1]), method="direct",
                           use_zero_dm=True, zdm_c0=0, zdm_c1=nchan - 1)
    plan = capi.Plan(spec)
    d_in = capi.DeviceBuffer(x.nbytes)
    d_in.upload(x)
    d_out = capi.DeviceBuffer(1001 * nsamp * 4)
    plan.execute(d_in.ptr, d_out.ptr)
    capi.sync()
    plan.set_profiling(2)   # events around the dedispersion stage only
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.execute(d_in.ptr, d_out.ptr)
    capi.sync()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    st = plan.stage_times()
    plan.close()
    d_in.free()
    d_out.free()
    work = 1001 * nsamp
    res = {"workload": "1024 ch u8 1000-1500 MHz x 2^18 samples, direct method, DM 0-1000 "
                       "step 1 (1001 trials), zero-DM, full-length u32 plane",
           "ms": ms, "rate": work / (ms / 1e3), "dedisp_ms": st["dedisp"],
           "adds_per_s": 1001.0 * nchan * nsamp / (st["dedisp"] / 1e3)}
    if with_cpu:
        from oracle import oracle as O   # CPU baseline leg: the oracle's cpucal restatement
        # cpucal's channel range excludes the top channel: [0, nchan - 1)
        all_dms = np.arange(1001, dtype=np.float64)
        hc = host_cpu()
        native = O.cpucal_native()

        def run(k):
            idx = np.linspace(0, 1000, k).round().astype(np.int64)
            d = O.delay_table(f, 0, nchan - 1, f[-1], float(band.tsamp), 1, all_dms[idx])
            t = time.perf_counter()
            O.cpucal_direct_u8(x, 0, nchan - 1, d, 0, nsamp, 1, nthreads=hc["threads"],
                               lib=native)
            return time.perf_counter() - t

        dt = run(4)
        k = int(min(1001, max(4, round(4 * cpu_target_s / max(dt, 1e-3)))))
        el = run(k)
        cpu_rate = k * nsamp / el
        res["cpu_baseline"] = {
            "value": cpu_rate, "unit": UNIT, "cores": hc["threads"], "kind": "port",
            "cpu_model": hc["model"],
            "sample": f"{k} of 1001 DM trials (evenly spaced) x {nsamp} samples x {nchan - 1} "
                      f"channels, cpucal direct loop nest (include/cpucal.hpp:167-186), "
                      f"-Ofast -march=native -fopenmp, {el:.2f} s"}
        res["speedup_vs_cpu"] = res["rate"] / cpu_rate
    return res


def small_batch_leg(threads, nchunk=64, reps=5):
    """Many small beam-chunks in one batch, where merging pays: 64 chunks of
    1024 ch u8 x 16384 samples, DM 0-500 step 1 (501 trials), 32-channel
    subbands, IQRM + zero-DM.  Plain executes one after another vs the
    batched launches vs their hipGraph replay (inputs resident in HBM)."""
    from astroflow_amd import capi
    c = dict(nchan=1024, nsamp=16384, dtype=np.uint8, dm_max=500.0, dm_step=1.0, ds=1)
    spec = c3_spec(c)
    from astroflow_amd import synth
    synth_pool = synth.noise_u8_threaded(43, 16384, 1024, threads=threads)
    plans, ins, outs = [], [], []
    for _ in range(nchunk):
        p = capi.Plan(spec)
        di = capi.DeviceBuffer(synth_pool.nbytes)
        di.upload(synth_pool)
        do = capi.DeviceBuffer(p.info.ndm * p.info.nout * 4)
        plans.append(p)
        ins.append(di)
        outs.append(do)
    capi.sync()

    def plain():
        for p, di, do in zip(plans, ins, outs):
            p.execute(di.ptr, do.ptr)

    batch = capi.Batch(plans, [b.ptr for b in ins], [b.ptr for b in outs])
    inf = batch.info()
    res = {}
    for name, fn in (("plain", plain), ("merged", lambda: batch.execute(use_graph=False)),
                     ("graph", lambda: batch.execute(use_graph=True)),
                     ("streams", lambda: batch.execute(use_graph=True, streams=BATCH_STREAMS))):
        fn()
        capi.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        capi.sync()
        res[name + "_ms"] = (time.perf_counter() - t0) * 1e3 / reps
    batch.close()
    for p in plans:
        p.close()
    for b in ins + outs:
        b.free()
    work = nchunk * 501 * 16384
    return {"workload": f"{nchunk} chunks of 1024 ch u8 x 16384 samples, DM 0-500 step 1 "
                        "(501 trials), 32-channel subbands, IQRM + zero-DM",
            "launches_batched": inf["launches"], "launches_plain": inf["recorded"],
            **res,
            "plain_rate": work / (res["plain_ms"] / 1e3),
            "graph_rate": work / (res["graph_ms"] / 1e3),
            "speedup_merged_vs_plain": res["plain_ms"] / res["merged_ms"],
            "speedup_graph_vs_plain": res["plain_ms"] / res["graph_ms"],
            "speedup_streams_vs_plain": res["plain_ms"] / res["streams_ms"]}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus must be 1 (the path "
                                   "shards by beam or DM range with no exchange; that layout "
                                   "is not run here)"}))
        sys.exit(2)
    from astroflow_amd import capi

    if capi.device_count() < 1:
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    capi.check(capi.lib().af_set_device(0))
    threads = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)))
    if a.direct_only:
        print(json.dumps({"direct_1024": direct_leg(threads, 0.0, reps=max(1, a.steps),
                                                    with_cpu=False)}))
        return
    if a.served_only:
        print(json.dumps({"served": served_leg(threads, a.served_chunks, a.served_full_rfi)}))
        return
    if a.batch_only:
        print(json.dumps({"batched": batch_leg(threads), "batched_small": small_batch_leg(threads)}))
        return
    x, spec, info = make_workload(threads)
    spec.dm_per_block = a.dm_per_block
    plan = capi.Plan(spec)
    d_in = capi.DeviceBuffer(x.nbytes)
    d_in.upload(x)
    d_out = capi.DeviceBuffer(info["ndm"] * info["nout"] * 4)

    if a.calib_copy:
        cb = capi.DeviceBuffer(256 << 20)
        capi.check(capi.lib().af_memcpy_d2d(capi.C.c_void_p(cb.ptr), capi.C.c_void_p(d_in.ptr),
                                            256 << 20, None))
        capi.sync()
        cb.free()
    for _ in range(a.warmup):
        plan.execute(d_in.ptr, d_out.ptr)
    capi.sync()

    # timed region: HIP events only around the dominant stage (two per step;
    # each recorded event idles the GPU for a few microseconds)
    plan.set_profiling(2)
    capi.sync()
    t_start = time.perf_counter()
    for _ in range(a.steps):
        plan.execute(d_in.ptr, d_out.ptr)
    capi.sync()
    t_end = time.perf_counter()
    dom = plan.stage_times()
    # per-stage breakdown from the same number of steps, outside the timed region
    plan.set_profiling(True)
    for _ in range(a.steps):
        plan.execute(d_in.ptr, d_out.ptr)
    capi.sync()
    stages = plan.stage_times()
    plan.set_profiling(False)

    ms = (t_end - t_start) * 1e3 / a.steps
    work = info["ndm"] * info["nout"] * info["ds"]
    value = work / (ms / 1e3)

    # per-call latency of the plain path (synchronous calls)
    call_ms = []
    for _ in range(max(5, min(a.steps, 20))):
        t0 = time.perf_counter()
        plan.execute(d_in.ptr, d_out.ptr)
        capi.sync()
        call_ms.append((time.perf_counter() - t0) * 1e3)

    # B_alg: the input window read once plus the plane written once
    # (nchan * (N_out * ds + D_max) * sizeof(T_in) + N_DM * N_out * sizeof(T_out))
    b_in = info["nchan"] * (info["nout"] * info["ds"] + info["dmax"]) * 1
    b_out = info["ndm"] * info["nout"] * 4
    b_alg = b_in + b_out
    # dominant kernel: subband stage 1; achieved = B_alg / its average time
    dom_ms = dom["dedisp"]
    achieved = b_alg / (dom_ms / 1e3) / 1e9
    adds = plan.info.adds
    traffic = profiled_traffic(STAGE1_KERNEL)

    res = {
        "device": capi.device_props(0),
        "metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8", "data": "synthetic",
        "config": {"workload": "single 0.5 s chunk, subband dedispersion + IQRM + zero-DM: "
                               "4096 ch u8 1000-1500 MHz, 10173 out + 140618 tail samples, "
                               "DM 0-3000 step 0.5 (6001 trials), 32 subbands x 128 ch, "
                               "32 fine DMs per coarse DM",
                   "nchan": info["nchan"], "nout": info["nout"], "nsamp_in": info["nsamp"],
                   "ndm": info["ndm"], "nsub": 32, "ndm0": 32, "path": "plain single call"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": STAGE1_KERNEL + " (subband stage 1)",
                     "kernel_ms": dom_ms, "alg_bytes": b_alg,
                     "alg_bytes_def": "B_alg = nchan*(N_out*ds + D_max)*1 B + N_DM*N_out*4 B "
                                      "(input window read once + plane written once)",
                     "path_frac": b_alg / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                     "traffic_source": TRAFFIC_FILE if traffic else None},
        "path_roofline": {"alg_bytes": b_alg, "GB_s": b_alg / (ms / 1e3) / 1e9,
                          "hbm_frac": b_alg / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                          "adds": adds, "valu_add_frac": adds / (ms / 1e3) / VALU_ADD_PEAK},
        "stage_ms": {k: v for k, v in stages.items() if k != "n"},
        "latency_ms": {"p50": float(np.percentile(call_ms, 50)),
                       "p99": float(np.percentile(call_ms, 99)),
                       "chunk_duration_ms": info["nout"] * 49.152e-3},
    }
    plan.close()
    d_in.free()
    d_out.free()
    if not a.no_served:
        res["served"] = served_leg(threads, a.served_chunks, a.served_full_rfi)
    if not a.no_batch:
        res["batched"] = batch_leg(threads)
        res["batched_small"] = small_batch_leg(threads)
        log(f"small-chunk batch: {json.dumps(res['batched_small'])}")
    if not a.no_direct:
        res["direct_1024"] = direct_leg(threads, a.cpu_seconds / 2, with_cpu=not a.no_cpu)
        log(f"direct leg: {res['direct_1024']['ms']:.2f} ms per call")
    if not a.no_cpu:
        res["cpu_baseline"] = cpu_baseline(x, info, a.cpu_seconds)
        res["speedup_vs_cpu"] = value / res["cpu_baseline"]["value"]
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res))


if __name__ == "__main__":
    main()

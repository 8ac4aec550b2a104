"""CPU oracle for the dedispersion + RFI hot path (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
may import this module; it is the checker, never the product.  It wraps the C
restatement in ``liboracle.so`` (``af_oracle.c``) and chains its stages the
way ``gpucal::dedisperse_spec`` does (src/core/dedispered.cu:1050-1235):
bin -> IQRM on binned data -> apply IQRM windows -> zero-DM -> static mask ->
dedisperse (direct or subband).
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
U8, U16, U32, F32 = 1, 2, 3, 4
IQRM_GPU, IQRM_CPU = 0, 1
KDM = 4148.808

_DT = {np.dtype(np.uint8): U8, np.dtype(np.uint16): U16,
       np.dtype(np.uint32): U32, np.dtype(np.float32): F32}


def dtype_code(a) -> int:
    return _DT[np.dtype(a.dtype if hasattr(a, "dtype") else a)]


def _load(name):
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise OSError(f"{path} is not built (run `make -C oracle`)")
    return C.CDLL(path)


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load("liboracle.so")
        P, I, L, D, F = C.c_void_p, C.c_int, C.c_long, C.c_double, C.c_float
        S, U = C.c_size_t, C.c_uint
        sig = {
            "afo_freq_table_f32": (None, [F, F, I, P]),
            "afo_chan_range": (I, [P, I, F, F, P, P]),
            "afo_dm_steps": (L, [F, F, F]),
            "afo_delay_table": (None, [P, I, I, D, D, I, P, I, P]),
            "afo_residual_table": (None, [P, I, I, I, D, D, I, D, D, I, I, P]),
            "afo_bin": (None, [P, I, S, I, I, P]),
            "afo_iqrm_windows": (I, [S, U, U, I, P, P, I]),
            "afo_iqrm_lags": (I, [I, D, P, I]),
            "afo_iqrm_radius": (I, [F, I]),
            "afo_iqrm_stat": (None, [P, I, I, U, U, I, I, I, I, P]),
            "afo_iqrm_on_stat": (None, [P, I, F, F, D, I, P]),
            "afo_iqrm_quartiles": (S, [P, S, I, P, P]),
            "afo_iqrm": (I, [P, I, S, I, I, I, I, F, F, D, U, U, I, I, P, P, P, I]),
            "afo_apply_window_masks": (None, [P, I, S, I, P, P, P, I]),
            "afo_zero_dm": (None, [P, I, S, I, I, I]),
            "afo_static_mask": (None, [P, I, S, I, P, I]),
            "afo_dedisperse_direct": (None, [P, I, S, I, I, I, P, I, S, P]),
            "afo_dedisperse_subband": (None, [P, I, S, I, I, I, I, P, P, I, I, S, P]),
            "afo_dedisperse_subband_rows": (None, [P, I, S, I, I, I, I, P, P, I, I, S, P, I, P]),
            "afo_cpucal_direct_u8": (None, [P, S, I, I, I, P, I, S, S, I, P, I]),
        }
        for n, (r, a) in sig.items():
            f = getattr(_lib, n)
            f.restype = r
            f.argtypes = a
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ------------------------------------------------------------------ setup --

def freq_table_f32(fch1: float, foff: float, nchan: int) -> np.ndarray:
    out = np.empty(nchan, np.float64)
    lib().afo_freq_table_f32(fch1, foff, nchan, _p(out))
    return out


def chan_range(freq: np.ndarray, fs: float, fe: float):
    cs, ce = C.c_long(), C.c_long()
    rc = lib().afo_chan_range(_p(freq), len(freq), fs, fe, C.byref(cs), C.byref(ce))
    if rc != 0:
        raise ValueError("frequency range outside the spectrum")
    return cs.value, ce.value


def dm_steps(lo: float, hi: float, step: float) -> int:
    return int(lib().afo_dm_steps(lo, hi, step))


def delay_table(freq, c0, c1, f_ref, tsamp, ds, dms) -> np.ndarray:
    dms = np.ascontiguousarray(dms, np.float64)
    out = np.empty((len(dms), c1 - c0), np.int32)
    lib().afo_delay_table(_p(np.ascontiguousarray(freq, np.float64)), c0, c1, f_ref,
                          tsamp, ds, _p(dms), len(dms), _p(out))
    return out


def residual_table(freq, c0, c1, sbw, f_ref, tsamp, ds, dm_low, dm_step, ndm, ndm0):
    nsb = (c1 - c0 + sbw - 1) // sbw
    out = np.empty((ndm, nsb), np.int32)
    lib().afo_residual_table(_p(np.ascontiguousarray(freq, np.float64)), c0, c1, sbw,
                             f_ref, tsamp, ds, dm_low, dm_step, ndm, ndm0, _p(out))
    return out


# ---------------------------------------------------------------- stages --

def bin_time(x: np.ndarray, ds: int) -> np.ndarray:
    x = np.ascontiguousarray(x)
    n, nc = x.shape
    out = np.empty(((n + ds - 1) // ds, nc), x.dtype)
    lib().afo_bin(_p(x), dtype_code(x), n, nc, ds, _p(out))
    return out


def iqrm_windows(nsample, W, H, include_tail):
    cap = 1 + (nsample // max(1, H if H else max(W, 1))) + 2 if W else 1
    t0 = np.zeros(cap, np.uint32)
    t1 = np.zeros(cap, np.uint32)
    n = lib().afo_iqrm_windows(nsample, W, H, int(include_tail), _p(t0), _p(t1), cap)
    assert n <= cap
    return t0[:n].copy(), t1[:n].copy()


def iqrm_lags(radius, geofactor):
    buf = np.zeros(512, np.int32)
    n = lib().afo_iqrm_lags(radius, geofactor, _p(buf), 512)
    return buf[:n].tolist()


def iqrm_radius(radius_frac, C_):
    return int(lib().afo_iqrm_radius(radius_frac, C_))


def iqrm_stat(x, t0, t1, c0, c1, mode, flavour=IQRM_GPU):
    x = np.ascontiguousarray(x)
    out = np.empty(c1 - c0, np.float32)
    lib().afo_iqrm_stat(_p(x), dtype_code(x), x.shape[1], t0, t1, c0, c1, mode, flavour, _p(out))
    return out


def iqrm_quartiles(d, flavour=IQRM_GPU):
    """(q1, median, q3) of lagged differences d, as the IQRM vote takes them."""
    d = np.ascontiguousarray(d, np.float32)
    srt = np.empty(max(1, len(d)), np.float32)
    q = np.zeros(3, np.float32)
    n = lib().afo_iqrm_quartiles(_p(d), len(d), flavour, _p(srt), _p(q))
    return (q if n else None), n


def iqrm_on_stat(stat, radius_frac, nsigma, geofactor, flavour=IQRM_GPU):
    stat = np.ascontiguousarray(stat, np.float32)
    m = np.empty(len(stat), np.uint8)
    lib().afo_iqrm_on_stat(_p(stat), len(stat), radius_frac, nsigma, geofactor, flavour, _p(m))
    return m


def iqrm(x, c0, c1, mode, radius_frac, nsigma, geofactor, W, H, include_tail,
         flavour=IQRM_GPU):
    """Returns (masks[nwin][nchan] u8, t0s, t1s)."""
    x = np.ascontiguousarray(x)
    n, nc = x.shape
    t0, t1 = iqrm_windows(n, W, H, include_tail)
    nw = len(t0)
    masks = np.zeros((max(nw, 1), nc), np.uint8)
    t0b = np.zeros(max(nw, 1), np.uint32)
    t1b = np.zeros(max(nw, 1), np.uint32)
    got = lib().afo_iqrm(_p(x), dtype_code(x), n, nc, c0, c1, mode, radius_frac, nsigma,
                         geofactor, W, H, int(include_tail), flavour, _p(masks), _p(t0b),
                         _p(t1b), max(nw, 1))
    assert got == nw
    return masks[:nw].copy(), t0b[:nw].copy(), t1b[:nw].copy()


def apply_window_masks(x, masks, t0s, t1s):
    n, nc = x.shape
    masks = np.ascontiguousarray(masks, np.uint8)
    t0s = np.ascontiguousarray(t0s, np.uint32)
    t1s = np.ascontiguousarray(t1s, np.uint32)
    lib().afo_apply_window_masks(_p(x), dtype_code(x), n, nc, _p(masks), _p(t0s), _p(t1s), len(t0s))


def zero_dm(x, c0, c1):
    n, nc = x.shape
    lib().afo_zero_dm(_p(x), dtype_code(x), n, nc, c0, c1)


def static_mask(x, bad):
    n, nc = x.shape
    bad = np.ascontiguousarray(bad, np.int32)
    lib().afo_static_mask(_p(x), dtype_code(x), n, nc, _p(bad), len(bad))


def _acc_dtype(code):
    return {U8: np.uint32, U16: np.uint32, U32: np.uint64, F32: np.float64}[code]


def dedisperse_direct(x, c0, c1, delays, nout):
    x = np.ascontiguousarray(x)
    n, nc = x.shape
    delays = np.ascontiguousarray(delays, np.int32)
    ndm = delays.shape[0]
    out = np.empty((ndm, nout), _acc_dtype(dtype_code(x)))
    lib().afo_dedisperse_direct(_p(x), dtype_code(x), n, nc, c0, c1, _p(delays), ndm, nout, _p(out))
    return out


def dedisperse_subband(x, c0, c1, sbw, delays1, res, ndm0, nout):
    x = np.ascontiguousarray(x)
    n, nc = x.shape
    delays1 = np.ascontiguousarray(delays1, np.int32)
    res = np.ascontiguousarray(res, np.int32)
    ndm = res.shape[0]
    out = np.empty((ndm, nout), _acc_dtype(dtype_code(x)))
    lib().afo_dedisperse_subband(_p(x), dtype_code(x), n, nc, c0, c1, sbw, _p(delays1),
                                 _p(res), ndm, ndm0, nout, _p(out))
    return out


def dedisperse_subband_rows(x, c0, c1, sbw, delays1, res, ndm0, nout, rows):
    """Subband plane for the fine-DM rows listed in ``rows`` only."""
    x = np.ascontiguousarray(x)
    n, nc = x.shape
    delays1 = np.ascontiguousarray(delays1, np.int32)
    res = np.ascontiguousarray(res, np.int32)
    rows = np.ascontiguousarray(rows, np.int32)
    out = np.empty((len(rows), nout), _acc_dtype(dtype_code(x)))
    lib().afo_dedisperse_subband_rows(_p(x), dtype_code(x), n, nc, c0, c1, sbw, _p(delays1),
                                      _p(res), res.shape[0], ndm0, nout, _p(rows), len(rows),
                                      _p(out))
    return out


def dedisperse_direct_rows(x, c0, c1, delays, nout, rows):
    """Direct plane for the DM rows listed in ``rows`` only."""
    rows = np.asarray(rows, np.int64)
    return dedisperse_direct(x, c0, c1, np.ascontiguousarray(delays[rows]), nout)


_native = None


def cpucal_native():
    """af_cpucal.c compiled on the running host with the reference's host
    flags (-Ofast -march=native -fopenmp, CMakeLists.txt:118-119) into a
    temporary directory, so -march=native means this machine's CPU."""
    global _native
    if _native is None:
        import subprocess
        import tempfile
        src = os.path.join(_HERE, "af_cpucal.c")
        out = os.path.join(tempfile.mkdtemp(prefix="af_cpucal_"), "libcpucal_native.so")
        subprocess.run(["gcc", "-Ofast", "-march=native", "-fopenmp", "-fPIC", "-shared",
                        "-std=c11", f"-I{_HERE}", src, "-o", out], check=True)
        _native = C.CDLL(out)
        f = _native.afo_cpucal_direct_u8
        f.restype = None
        P, I, S = C.c_void_p, C.c_int, C.c_size_t
        f.argtypes = [P, S, I, I, I, P, I, S, S, I, P, I]
    return _native


def cpucal_direct_u8(x, c0, c1, delays, start, length, ds, nthreads=0, lib=None):
    x = np.ascontiguousarray(x, np.uint8)
    n, nc = x.shape
    delays = np.ascontiguousarray(delays, np.int32)
    ndm = delays.shape[0]
    out = np.empty((ndm, (length + ds - 1) // ds), np.uint64)
    (lib or globals()["lib"]()).afo_cpucal_direct_u8(_p(x), n, nc, c0, c1, _p(delays), ndm,
                                                     start, length, ds, _p(out), nthreads)
    return out


# -------------------------------------------------------------- pipeline --

@dataclass
class PlanSpec:
    """Everything that defines one dedispersion call (mirrors af_plan_desc)."""
    nchan: int
    freq: np.ndarray            # [nchan] MHz, ascending
    tsamp: float                # native sample time, s
    dm_low: float
    dm_step: float
    ndm: int
    dd_c0: int
    dd_c1: int                  # exclusive
    f_ref: float
    ds: int = 1
    nout: Optional[int] = None  # binned outputs to produce (default: all)
    method: str = "subband"     # or "direct"
    sbw: int = 32               # channels per subband
    ndm0: int = 32              # fine DMs per coarse group
    use_iqrm: bool = False
    iqrm_c0: int = 0
    iqrm_c1: int = 0
    iqrm_mode: int = 1
    iqrm_radius_frac: float = 0.1
    iqrm_nsigma: float = 3.0
    iqrm_geofactor: float = 1.5
    iqrm_W: int = 0             # window length in binned samples (0 = whole)
    iqrm_H: int = 0
    iqrm_include_tail: bool = True
    use_zero_dm: bool = False
    zdm_c0: int = 0
    zdm_c1: int = 0
    bad_channels: List[int] = field(default_factory=list)

    def dms(self):
        return self.dm_low + np.arange(self.ndm, dtype=np.float64) * self.dm_step

    def coarse_dms(self):
        nnom = (self.ndm + self.ndm0 - 1) // self.ndm0
        return self.dm_low + (np.arange(nnom, dtype=np.float64) * self.ndm0) * self.dm_step


def run_pipeline(x: np.ndarray, spec: PlanSpec, flavour=IQRM_GPU):
    """Oracle for one plain call.  Returns (plane, masks, t0s, t1s, cleaned)."""
    x = np.array(x, copy=True, order="C")
    if spec.ds > 1:
        x = bin_time(x, spec.ds)
    nbin = x.shape[0]
    nout = spec.nout or nbin      # 0 or None: the full binned length
    masks = np.zeros((0, spec.nchan), np.uint8)
    t0s = t1s = np.zeros(0, np.uint32)
    if spec.use_iqrm:
        masks, t0s, t1s = iqrm(x, spec.iqrm_c0, spec.iqrm_c1, spec.iqrm_mode,
                               spec.iqrm_radius_frac, spec.iqrm_nsigma, spec.iqrm_geofactor,
                               spec.iqrm_W, spec.iqrm_H, spec.iqrm_include_tail, flavour)
        apply_window_masks(x, masks, t0s, t1s)
    if spec.use_zero_dm:
        zero_dm(x, spec.zdm_c0, spec.zdm_c1)
    if len(spec.bad_channels):
        static_mask(x, spec.bad_channels)
    if spec.method == "direct":
        d = delay_table(spec.freq, spec.dd_c0, spec.dd_c1, spec.f_ref, spec.tsamp, spec.ds,
                        spec.dms())
        plane = dedisperse_direct(x, spec.dd_c0, spec.dd_c1, d, nout)
    else:
        d1 = delay_table(spec.freq, spec.dd_c0, spec.dd_c1, spec.f_ref, spec.tsamp, spec.ds,
                         spec.coarse_dms())
        res = residual_table(spec.freq, spec.dd_c0, spec.dd_c1, spec.sbw, spec.f_ref,
                             spec.tsamp, spec.ds, spec.dm_low, spec.dm_step, spec.ndm, spec.ndm0)
        plane = dedisperse_subband(x, spec.dd_c0, spec.dd_c1, spec.sbw, d1, res, spec.ndm0, nout)
    return plane, masks, t0s, t1s, x


def run_pipeline_rows(x: np.ndarray, spec: PlanSpec, rows, flavour=IQRM_GPU, binned=False):
    """run_pipeline for the DM rows `rows` only (large inputs).  Returns
    (plane rows, masks, cleaned).  binned=True: x is already at the binned
    rate (the binning step is skipped; the tables still use spec.ds)."""
    if spec.ds > 1 and not binned:
        x = bin_time(x, spec.ds)      # a new array
    else:
        x = np.array(x, copy=True, order="C")
    nbin = x.shape[0]
    nout = spec.nout or nbin      # 0 or None: the full binned length
    masks = np.zeros((0, spec.nchan), np.uint8)
    if spec.use_iqrm:
        masks, t0s, t1s = iqrm(x, spec.iqrm_c0, spec.iqrm_c1, spec.iqrm_mode,
                               spec.iqrm_radius_frac, spec.iqrm_nsigma, spec.iqrm_geofactor,
                               spec.iqrm_W, spec.iqrm_H, spec.iqrm_include_tail, flavour)
        apply_window_masks(x, masks, t0s, t1s)
    if spec.use_zero_dm:
        zero_dm(x, spec.zdm_c0, spec.zdm_c1)
    if len(spec.bad_channels):
        static_mask(x, spec.bad_channels)
    if spec.method == "direct":
        d = delay_table(spec.freq, spec.dd_c0, spec.dd_c1, spec.f_ref, spec.tsamp, spec.ds,
                        spec.dms())
        out = dedisperse_direct_rows(x, spec.dd_c0, spec.dd_c1, d, nout, rows)
    else:
        d1 = delay_table(spec.freq, spec.dd_c0, spec.dd_c1, spec.f_ref, spec.tsamp, spec.ds,
                         spec.coarse_dms())
        res = residual_table(spec.freq, spec.dd_c0, spec.dd_c1, spec.sbw, spec.f_ref,
                             spec.tsamp, spec.ds, spec.dm_low, spec.dm_step, spec.ndm, spec.ndm0)
        out = dedisperse_subband_rows(x, spec.dd_c0, spec.dd_c1, spec.sbw, d1, res, spec.ndm0,
                                      nout, rows)
    return out, masks, x


# ---------------------------------------------------- reference (oracle/_ref) --

_ref = None


def ref_lib():
    """The reference's own CPU IQRM / mask code, built from its sources by
    ``make -C oracle ref``; None when not built (e.g. without the sources)."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libafref.so")
        if not os.path.exists(path):
            return None
        _ref = C.CDLL(path)
        P, I, D, F, S = C.c_void_p, C.c_int, C.c_double, C.c_float, C.c_size_t
        _ref.afref_iqrm_u8.restype = I
        _ref.afref_iqrm_u8.argtypes = [P, S, I, I, I, D, I, F, F, D, D, D, I, P, P, P, I]
        _ref.afref_iqrm_u16.restype = I
        _ref.afref_iqrm_u16.argtypes = [P, S, I, I, I, D, I, F, F, D, D, D, I, P, P, P, I]
        _ref.afref_iqrm_on_stat.restype = None
        _ref.afref_iqrm_on_stat.argtypes = [P, I, F, F, D, P]
        _ref.afref_mask_u8.restype = None
        _ref.afref_mask_u8.argtypes = [P, I, S, P, P, P, I]
        _ref.afref_mark_rfi_u8.restype = I
        _ref.afref_mark_rfi_u8.argtypes = [P, I, S, C.c_char_p, P, I]
    return _ref

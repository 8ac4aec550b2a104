"""ctypes binding of the C-ABI in include/astroflow_amd.h.

The library is loaded from the package directory (built in-tree by
``python -m astroflow_amd.build``).  There is no fallback: a missing library
raises, so nothing silently runs on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libastroflow_amd.so")

AF_OK, AF_ERR_INVALID, AF_ERR_RUNTIME, AF_ERR_NOMEM, AF_ERR_NODEV, AF_ERR_BUSY = range(6)
AF_U8, AF_U16, AF_U32, AF_F32 = 1, 2, 3, 4
AF_U64 = 5   # plane element type of u32 samples (af_image_slices)
METHOD_DIRECT, METHOD_SUBBAND = 0, 1

DTYPE_CODE = {np.dtype(np.uint8): AF_U8, np.dtype(np.uint16): AF_U16,
              np.dtype(np.uint32): AF_U32, np.dtype(np.float32): AF_F32}
CODE_DTYPE = {v: k for k, v in DTYPE_CODE.items()}
OUT_DTYPE = {AF_U8: np.dtype(np.uint32), AF_U16: np.dtype(np.uint32),
             AF_U32: np.dtype(np.uint64), AF_F32: np.dtype(np.float32)}


class AFError(RuntimeError):
    pass


class PlanDesc(C.Structure):
    _fields_ = [
        ("dtype", C.c_int), ("nchan", C.c_int), ("nsamp", C.c_size_t), ("ds", C.c_int),
        ("tsamp", C.c_double), ("freq_mhz", C.POINTER(C.c_double)), ("f_ref_mhz", C.c_double),
        ("dd_c0", C.c_int), ("dd_c1", C.c_int), ("dm_low", C.c_double), ("dm_step", C.c_double),
        ("ndm", C.c_int), ("nout", C.c_size_t), ("method", C.c_int), ("sb_nchan", C.c_int),
        ("ndm0", C.c_int),
        ("use_iqrm", C.c_int), ("iqrm_c0", C.c_int), ("iqrm_c1", C.c_int),
        ("iqrm_mode", C.c_int), ("iqrm_radius_frac", C.c_float), ("iqrm_nsigma", C.c_float),
        ("iqrm_geofactor", C.c_double), ("iqrm_win", C.c_uint), ("iqrm_hop", C.c_uint),
        ("iqrm_include_tail", C.c_int),
        ("use_zero_dm", C.c_int), ("zdm_c0", C.c_int), ("zdm_c1", C.c_int),
        ("bad_chans", C.POINTER(C.c_int)), ("n_bad", C.c_int),
        ("tile_t", C.c_int), ("dm_per_block", C.c_int), ("inter_budget", C.c_size_t),
        ("clean_budget", C.c_size_t),
    ]


class PlanInfo(C.Structure):
    _fields_ = [("nbin", C.c_size_t), ("nout", C.c_size_t), ("ndm", C.c_int),
                ("out_elem_bytes", C.c_int), ("nwin", C.c_int), ("nsb", C.c_int),
                ("ndm_nom", C.c_int), ("max_delay", C.c_int), ("scratch_bytes", C.c_size_t),
                ("adds", C.c_double), ("inter_bytes", C.c_size_t),
                ("clean_passes", C.c_int), ("stage1_form", C.c_int)]


class DeviceProps(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("cu_count", C.c_int),
                ("clock_khz", C.c_int), ("lds_per_block", C.c_int), ("lds_per_cu", C.c_int),
                ("l2_bytes", C.c_int), ("hbm_bytes", C.c_size_t)]


class StageTimes(C.Structure):
    _fields_ = [("bin", C.c_float), ("stats", C.c_float), ("iqrm", C.c_float),
                ("clean", C.c_float), ("dedisp", C.c_float), ("stage2", C.c_float),
                ("total", C.c_float), ("n", C.c_int)]


class ServerDesc(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_inflight", C.c_int), ("max_queue", C.c_int),
                ("full_rfi", C.c_int)]


class ServerResult(C.Structure):
    _fields_ = [("beam", C.c_int), ("has_plane", C.c_int), ("status", C.c_int),
                ("tag", C.c_uint64), ("window", C.c_uint64),
                ("d_plane", C.c_void_p), ("d_masks", C.c_void_p),
                ("t_submit_ms", C.c_double), ("t_admit_ms", C.c_double),
                ("t_done_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("admitted", C.c_uint64),
                ("completed", C.c_uint64), ("planes", C.c_uint64),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("max_ms", C.c_double),
                ("mean_ms", C.c_double), ("first_submit_ms", C.c_double),
                ("last_done_ms", C.c_double), ("hbm_used", C.c_size_t),
                ("hbm_peak", C.c_size_t), ("hbm_budget", C.c_size_t),
                ("inflight_peak", C.c_int)]


class FilInfo(C.Structure):
    _fields_ = [("nchans", C.c_int), ("nbits", C.c_int), ("dtype", C.c_int), ("npol", C.c_int),
                ("flipped", C.c_int), ("ndata", C.c_uint64), ("data_offset", C.c_uint64),
                ("fch1", C.c_double), ("foff", C.c_double), ("tsamp", C.c_double),
                ("tstart", C.c_double)]


class SnrParams(C.Structure):
    _fields_ = [("threshold_sigma", C.c_double), ("has_toa", C.c_int), ("toa", C.c_int),
                ("has_window", C.c_int), ("window", C.c_int), ("has_tsamp", C.c_int),
                ("has_target", C.c_int), ("tsamp", C.c_double), ("target_time_us", C.c_double),
                ("has_noise_ranges", C.c_int), ("n_noise_ranges", C.c_int),
                ("noise_ranges", C.c_void_p)]


class SnrResult(C.Structure):
    _fields_ = [("snr", C.c_double), ("width", C.c_int), ("peak", C.c_int),
                ("noise_mean", C.c_double), ("noise_std", C.c_double), ("method", C.c_int),
                ("converged", C.c_int), ("search_start", C.c_int), ("search_end", C.c_int),
                ("downsample", C.c_int)]


_lib = None

# name -> (restype, argtypes); every symbol declared in include/astroflow_amd.h
_P, _I, _S = C.c_void_p, C.c_int, C.c_size_t
SIGNATURES = {
    "af_plan_create": (_I, [C.POINTER(PlanDesc), C.POINTER(_P)]),
    "af_plan_destroy": (_I, [_P]),
    "af_plan_query": (_I, [_P, C.POINTER(PlanInfo)]),
    "af_plan_tables": (_I, [_P, _P, _P]),
    "af_plan_windows": (_I, [_P, _P, _P]),
    "af_plan_set_profiling": (_I, [_P, _I]),
    "af_plan_stage_times": (_I, [_P, C.POINTER(StageTimes)]),
    "af_execute": (_I, [_P, _P, _P, _P, _P]),
    "af_execute_host": (_I, [_P, _P, _P, _P]),
    "af_execute_host_images": (_I, [_P, _P, _S, _I, _P, _P]),
    "af_image_slices": (_I, [_P, _I, _S, _S, _S, _S, _I, _P, _P]),
    "af_execute_spectrum": (_I, [_P, _P, _P, _S, _S, _P, _P]),
    "af_device_query": (_I, [_I, C.POINTER(DeviceProps)]),
    "af_device_count": (_I, [C.POINTER(_I)]),
    "af_set_device": (_I, [_I]),
    "af_malloc": (_I, [C.POINTER(_P), _S]),
    "af_free": (_I, [_P]),
    "af_memcpy_h2d": (_I, [_P, _P, _S, _P]),
    "af_memcpy_d2h": (_I, [_P, _P, _S, _P]),
    "af_memcpy_d2d": (_I, [_P, _P, _S, _P]),
    "af_memset": (_I, [_P, _I, _S, _P]),
    "af_stream_create": (_I, [C.POINTER(_P)]),
    "af_stream_destroy": (_I, [_P]),
    "af_stream_sync": (_I, [_P]),
    "af_sync": (_I, []),
    "af_time_execute": (_I, [_P, _P, _P, _I, C.POINTER(C.c_float)]),
    "af_execute_noalloc": (_I, [C.POINTER(PlanDesc), _P, _P, _P, _P]),
    "af_batch_create": (_I, [C.POINTER(_P), _I, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P),
                             C.POINTER(_P)]),
    "af_batch_execute": (_I, [_P, _I, _P]),
    "af_batch_execute_streams": (_I, [_P, _I, _I, _P]),
    "af_batch_info": (_I, [_P, C.POINTER(_I), C.POINTER(_I)]),
    "af_batch_destroy": (_I, [_P]),
    "af_autotune": (_I, [C.POINTER(PlanDesc), _P, _P, _I, _I, C.POINTER(C.c_float)]),
    "af_host_alloc": (_I, [C.POINTER(_P), _S]),
    "af_host_free": (_I, [_P]),
    "af_server_create": (_I, [C.POINTER(ServerDesc), C.POINTER(_P)]),
    "af_server_destroy": (_I, [_P]),
    "af_server_add_beam": (_I, [_P, C.POINTER(PlanDesc), _S, C.POINTER(_I)]),
    "af_server_submit": (_I, [_P, _I, _P, C.c_uint64]),
    "af_server_poll": (_I, [_P, C.POINTER(ServerResult), _I, C.POINTER(_I), _I]),
    "af_server_release": (_I, [_P, _P]),
    "af_server_drain": (_I, [_P]),
    "af_server_get_stats": (_I, [_P, C.POINTER(ServerStats)]),
    "af_fil_open": (_I, [C.c_char_p, C.POINTER(_P), C.POINTER(FilInfo)]),
    "af_fil_read": (_I, [_P, C.c_uint64, _S, _P]),
    "af_fil_close": (_I, [_P]),
    "af_psrfits_open": (_I, [C.c_char_p, C.POINTER(_P), C.POINTER(FilInfo)]),
    "af_psrfits_read": (_I, [_P, C.c_uint64, _S, _P]),
    "af_psrfits_close": (_I, [_P]),
    "af_frb_snr": (_I, [_P, _I, _I, _I, C.POINTER(SnrParams), C.POINTER(SnrResult), _P]),
    "af_last_error": (C.c_char_p, []),
    "af_version": (C.c_char_p, []),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AFError(f"{LIB_PATH} is not built; run `python -m astroflow_amd.build`")
        _lib = C.CDLL(LIB_PATH)
        for name, (r, a) in SIGNATURES.items():
            f = getattr(_lib, name)
            f.restype = r
            f.argtypes = a
    return _lib


_EXC = {AF_ERR_INVALID: ValueError, AF_ERR_NOMEM: MemoryError}


def check(rc: int, what: str = ""):
    if rc != AF_OK:
        msg = lib().af_last_error().decode(errors="replace")
        exc = _EXC.get(rc, AFError)
        raise exc(f"{what}: {msg}" if what else msg)


def device_count() -> int:
    n = C.c_int(0)
    check(lib().af_device_count(C.byref(n)), "af_device_count")
    return n.value


def device_props(dev: int = 0) -> dict:
    p = DeviceProps()
    check(lib().af_device_query(dev, C.byref(p)), "af_device_query")
    return {k: (getattr(p, k).decode() if isinstance(getattr(p, k), bytes) else getattr(p, k))
            for k, _ in DeviceProps._fields_}


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


@dataclass
class DedispSpec:
    """One dedispersion call, in the C-ABI's explicit terms (af_plan_desc)."""
    dtype: np.dtype
    nchan: int
    nsamp: int
    freq: np.ndarray
    tsamp: float
    dm_low: float
    dm_step: float
    ndm: int
    dd_c0: int
    dd_c1: int
    f_ref: float
    ds: int = 1
    nout: Optional[int] = None
    method: str = "subband"
    sbw: int = 32
    ndm0: int = 32
    use_iqrm: bool = False
    iqrm_c0: int = 0
    iqrm_c1: int = 0
    iqrm_mode: int = 1
    iqrm_radius_frac: float = 0.1
    iqrm_nsigma: float = 3.0
    iqrm_geofactor: float = 1.5
    iqrm_W: int = 0
    iqrm_H: int = 0
    iqrm_include_tail: bool = True
    use_zero_dm: bool = False
    zdm_c0: int = 0
    zdm_c1: int = 0
    bad_channels: List[int] = field(default_factory=list)
    tile_t: int = 0
    dm_per_block: int = 0
    inter_budget: int = 0
    clean_budget: int = 0

    def to_desc(self):
        """Returns (PlanDesc, keepalive) -- keepalive owns the arrays it points to."""
        freq = np.ascontiguousarray(self.freq, np.float64)
        bad = np.ascontiguousarray(self.bad_channels, np.int32)
        d = PlanDesc()
        d.dtype = DTYPE_CODE[np.dtype(self.dtype)]
        d.nchan = self.nchan
        d.nsamp = self.nsamp
        d.ds = self.ds
        d.tsamp = self.tsamp
        d.freq_mhz = freq.ctypes.data_as(C.POINTER(C.c_double))
        d.f_ref_mhz = self.f_ref
        d.dd_c0, d.dd_c1 = self.dd_c0, self.dd_c1
        d.dm_low, d.dm_step, d.ndm = self.dm_low, self.dm_step, self.ndm
        d.nout = self.nout or 0
        d.method = METHOD_DIRECT if self.method == "direct" else METHOD_SUBBAND
        d.sb_nchan, d.ndm0 = self.sbw, self.ndm0
        d.use_iqrm = int(self.use_iqrm)
        d.iqrm_c0, d.iqrm_c1, d.iqrm_mode = self.iqrm_c0, self.iqrm_c1, self.iqrm_mode
        d.iqrm_radius_frac, d.iqrm_nsigma = self.iqrm_radius_frac, self.iqrm_nsigma
        d.iqrm_geofactor = self.iqrm_geofactor
        d.iqrm_win, d.iqrm_hop = self.iqrm_W, self.iqrm_H
        d.iqrm_include_tail = int(self.iqrm_include_tail)
        d.use_zero_dm, d.zdm_c0, d.zdm_c1 = int(self.use_zero_dm), self.zdm_c0, self.zdm_c1
        d.bad_chans = bad.ctypes.data_as(C.POINTER(C.c_int)) if len(bad) else None
        d.n_bad = len(bad)
        d.tile_t, d.dm_per_block = self.tile_t, self.dm_per_block
        d.inter_budget = self.inter_budget
        d.clean_budget = self.clean_budget
        return d, (freq, bad)


def execute_noalloc(spec: DedispSpec, d_in: int, d_out: int, d_masks: int = 0, stream: int = 0):
    """af_execute_noalloc: plan and execute in one call, every device buffer
    taken from the library's static scratch (no allocation)."""
    desc, keep = spec.to_desc()
    check(lib().af_execute_noalloc(C.byref(desc), C.c_void_p(d_in), C.c_void_p(d_out),
                                   C.c_void_p(d_masks) if d_masks else None,
                                   C.c_void_p(stream) if stream else None), "af_execute_noalloc")
    del keep


class Plan:
    """Owns an af_plan.  Device buffers are raw pointers (ints)."""

    def __init__(self, spec: DedispSpec):
        self.spec = spec
        desc, keep = spec.to_desc()
        h = C.c_void_p()
        check(lib().af_plan_create(C.byref(desc), C.byref(h)), "af_plan_create")
        self._h = h
        info = PlanInfo()
        check(lib().af_plan_query(h, C.byref(info)), "af_plan_query")
        self.info = info
        self.in_dtype = np.dtype(spec.dtype)
        self.out_dtype = OUT_DTYPE[DTYPE_CODE[self.in_dtype]]

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().af_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def out_shape(self):
        return (self.info.ndm, self.info.nout)

    @property
    def nwin(self):
        return self.info.nwin

    def tables(self):
        s = self.spec
        nch = s.dd_c1 - s.dd_c0
        rows = self.info.ndm if s.method == "direct" else self.info.ndm_nom
        d = np.empty((rows, nch), np.int32)
        res = None
        if s.method != "direct":
            res = np.empty((self.info.ndm, self.info.nsb), np.int32)
        check(lib().af_plan_tables(self._h, _ptr(d), _ptr(res) if res is not None else None))
        return d, res

    def windows(self):
        n = self.info.nwin
        t0 = np.empty(max(n, 1), np.uint32)
        t1 = np.empty(max(n, 1), np.uint32)
        check(lib().af_plan_windows(self._h, _ptr(t0), _ptr(t1)))
        return t0[:n].copy(), t1[:n].copy()

    def run_host(self, x: np.ndarray, want_masks: bool = False):
        x = np.ascontiguousarray(x)
        if x.dtype != self.in_dtype:
            raise ValueError(f"input dtype {x.dtype} != plan dtype {self.in_dtype}")
        if x.size != self.spec.nsamp * self.spec.nchan:
            raise ValueError("input size does not match the plan")
        out = np.empty(self.out_shape, self.out_dtype)
        masks = None
        if want_masks and self.nwin:
            masks = np.empty((self.nwin, self.spec.nchan), np.uint8)
        check(lib().af_execute_host(self._h, _ptr(x), _ptr(out),
                                    _ptr(masks) if masks is not None else None),
              "af_execute_host")
        return out, masks

    def run_host_images(self, x: np.ndarray, sps: int, size: int = 512):
        """The DM-time images of the plane (slices of sps columns, size x size
        BGR), made on the device; the plane itself is not downloaded."""
        x = np.ascontiguousarray(x)
        if x.dtype != self.in_dtype:
            raise ValueError(f"input dtype {x.dtype} != plan dtype {self.in_dtype}")
        if x.size != self.spec.nsamp * self.spec.nchan:
            raise ValueError("input size does not match the plan")
        if sps <= 0:
            raise ValueError("sps must be positive")
        nsl = -(-self.info.nout // sps)
        imgs = np.empty((nsl, size, size, 3), np.uint8)
        check(lib().af_execute_host_images(self._h, _ptr(x), sps, size, _ptr(imgs), None),
              "af_execute_host_images")
        return imgs

    def execute(self, d_in: int, d_out: int, d_masks: int = 0, stream: int = 0):
        check(lib().af_execute(self._h, C.c_void_p(d_in), C.c_void_p(d_out),
                               C.c_void_p(d_masks) if d_masks else None,
                               C.c_void_p(stream) if stream else None), "af_execute")

    def time_execute(self, d_in: int, d_out: int, reps: int) -> float:
        ms = C.c_float(0)
        check(lib().af_time_execute(self._h, C.c_void_p(d_in), C.c_void_p(d_out), reps,
                                    C.byref(ms)), "af_time_execute")
        return ms.value

    def set_profiling(self, on):
        """False/0 off, True/1 every stage, 2 the dedispersion stage only."""
        check(lib().af_plan_set_profiling(self._h, int(on)))

    def stage_times(self) -> dict:
        t = StageTimes()
        check(lib().af_plan_stage_times(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in StageTimes._fields_}


class DeviceBuffer:
    """A device allocation made through the C-ABI (no torch needed)."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        check(lib().af_malloc(C.byref(p), nbytes), "af_malloc")
        self.ptr = p.value
        self.nbytes = nbytes

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().af_memcpy_h2d(C.c_void_p(self.ptr), _ptr(a), a.nbytes, None), "h2d")

    def download(self, a: np.ndarray, offset: int = 0):
        """Fills `a` from the bytes at `offset` of this buffer."""
        assert a.flags.c_contiguous and offset >= 0 and offset + a.nbytes <= self.nbytes
        check(lib().af_memcpy_d2h(_ptr(a), C.c_void_p(self.ptr + offset), a.nbytes, None), "d2h")
        return a

    def upload_at(self, a: np.ndarray, offset: int):
        """Copies `a` to the bytes at `offset` of this buffer."""
        a = np.ascontiguousarray(a)
        assert offset >= 0 and offset + a.nbytes <= self.nbytes
        check(lib().af_memcpy_h2d(C.c_void_p(self.ptr + offset), _ptr(a), a.nbytes, None), "h2d")

    def zero(self):
        check(lib().af_memset(C.c_void_p(self.ptr), 0, self.nbytes, None), "memset")

    def free(self):
        if self.ptr:
            lib().af_free(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def sync():
    check(lib().af_sync(), "af_sync")


class HostBuffer:
    """Page-locked host memory (af_host_alloc) viewed as a numpy array."""

    def __init__(self, shape, dtype):
        dtype = np.dtype(dtype)
        n = int(np.prod(shape)) * dtype.itemsize
        p = C.c_void_p()
        check(lib().af_host_alloc(C.byref(p), n), "af_host_alloc")
        self.ptr = p.value
        buf = (C.c_char * n).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=dtype).reshape(shape)

    def free(self):
        if self.ptr:
            self.array = None
            lib().af_host_free(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """Owns an af_server: resident per-beam windows, budgeted admission."""

    def __init__(self, hbm_budget: int = 0, max_inflight: int = 0, max_queue: int = 0,
                 full_rfi: bool = False):
        d = ServerDesc(hbm_budget, max_inflight, max_queue, int(full_rfi))
        h = C.c_void_p()
        check(lib().af_server_create(C.byref(d), C.byref(h)), "af_server_create")
        self._h = h
        self.beams = []

    def add_beam(self, window: DedispSpec, chunk_rows: int) -> int:
        desc, keep = window.to_desc()
        b = C.c_int(-1)
        check(lib().af_server_add_beam(self._h, C.byref(desc), chunk_rows, C.byref(b)),
              "af_server_add_beam")
        self.beams.append((window, chunk_rows))
        return b.value

    def submit(self, beam: int, rows_ptr: int, tag: int = 0):
        check(lib().af_server_submit(self._h, beam, C.c_void_p(rows_ptr), tag), "af_server_submit")

    def poll(self, max_n: int = 64, wait_ms: int = 0):
        arr = (ServerResult * max_n)()
        n = C.c_int(0)
        check(lib().af_server_poll(self._h, arr, max_n, C.byref(n), wait_ms), "af_server_poll")
        return [arr[i] for i in range(n.value)]

    def release(self, d_plane: int):
        check(lib().af_server_release(self._h, C.c_void_p(d_plane)), "af_server_release")

    def drain(self):
        check(lib().af_server_drain(self._h), "af_server_drain")

    def stats(self) -> dict:
        st = ServerStats()
        check(lib().af_server_get_stats(self._h, C.byref(st)), "af_server_get_stats")
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().af_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Owns an af_batch over (plan, input, output[, masks]) device buffers."""

    def __init__(self, plans, d_ins, d_outs, d_masks=None):
        n = len(plans)
        arr = lambda vals: (C.c_void_p * n)(*[C.c_void_p(v or None) for v in vals])
        ph = (C.c_void_p * n)(*[p.handle for p in plans])
        h = C.c_void_p()
        check(lib().af_batch_create(ph, n, arr(d_ins), arr(d_outs),
                                    arr(d_masks) if d_masks is not None else None, C.byref(h)),
              "af_batch_create")
        self._h = h
        self.plans = list(plans)

    def execute(self, use_graph: bool = False, stream: int = 0, streams: int = 0):
        """streams == 0: merged launches (af_batch_execute); streams > 0: each
        plan's own launches, the plans spread over that many streams
        (af_batch_execute_streams)."""
        st = C.c_void_p(stream) if stream else None
        if streams:
            check(lib().af_batch_execute_streams(self._h, int(streams), int(use_graph), st),
                  "af_batch_execute_streams")
        else:
            check(lib().af_batch_execute(self._h, int(use_graph), st), "af_batch_execute")

    def info(self):
        a, b = C.c_int(), C.c_int()
        check(lib().af_batch_info(self._h, C.byref(a), C.byref(b)))
        return {"launches": a.value, "recorded": b.value}

    def close(self):
        if getattr(self, "_h", None):
            lib().af_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(spec: "DedispSpec", d_in: int, d_out: int, tune_subbands: bool = False,
             reps: int = 2):
    """Fastest (tile_t, dm_per_block, sbw) for the spec on these buffers;
    returns (tuned spec copy, ms per execute)."""
    import dataclasses
    desc, keep = spec.to_desc()
    ms = C.c_float(0)
    check(lib().af_autotune(C.byref(desc), C.c_void_p(d_in), C.c_void_p(d_out),
                            int(tune_subbands), reps, C.byref(ms)), "af_autotune")
    tuned = dataclasses.replace(spec, tile_t=desc.tile_t, dm_per_block=desc.dm_per_block,
                                sbw=desc.sb_nchan, ndm0=desc.ndm0)
    return tuned, ms.value


def frb_snr(d_spec: int, dtype: int, ntime: int, nfreq: int, noise_range=None,
            threshold_sigma: float = 5.0, toa_sample_idx=None, fitting_window_samples=None,
            tsamp=None, target_time_us=None, stream: int = 0):
    """af_frb_snr on a device spectrum, returned as _calculate_frb_snr does:
    (snr, width, peak, (noise_mean, noise_std, info))."""
    p = SnrParams()
    p.threshold_sigma = threshold_sigma
    pairs = []
    if noise_range is not None and len(noise_range) > 0:
        p.has_noise_ranges = 1
        pairs = [(int(r[0]), int(r[1])) for r in noise_range if len(r) == 2]
    arr = np.asarray(pairs, np.int32).reshape(-1, 2) if pairs else np.zeros((0, 2), np.int32)
    p.n_noise_ranges = len(pairs)
    p.noise_ranges = arr.ctypes.data if pairs else None
    if toa_sample_idx is not None:
        p.has_toa, p.toa = 1, int(toa_sample_idx)
    if fitting_window_samples is not None:
        p.has_window, p.window = 1, int(fitting_window_samples)
    if tsamp is not None:
        p.has_tsamp, p.tsamp = 1, float(tsamp)
    if target_time_us is not None:
        p.has_target, p.target_time_us = 1, float(target_time_us)
    r = SnrResult()
    check(lib().af_frb_snr(C.c_void_p(d_spec) if d_spec else None, dtype, ntime, nfreq,
                           C.byref(p), C.byref(r), C.c_void_p(stream) if stream else None),
          "af_frb_snr")
    info = {"fit_converged": bool(r.converged)}
    if r.method == 0:
        return (-1.0, 0, 0, (0.0, 1.0, info))
    if r.method == 2:
        info["method"] = "fallback"
        return (r.snr, r.width, r.peak, (r.noise_mean, r.noise_std, info))
    info.update({"method": "boxcar", "width_samples": r.width, "search_start": r.search_start,
                 "search_end": r.search_end, "downsample": r.downsample})
    return (r.snr, r.width, r.peak, (r.noise_mean, r.noise_std, info))


_DT = {AF_U8: np.uint8, AF_U16: np.uint16, AF_U32: np.uint32, AF_F32: np.float32}


class FileReader:
    """A SIGPROC (kind "fil") or PSRFITS (kind "psrfits") file opened through
    the C-ABI readers: rows time-major, channels ascending."""

    def __init__(self, path: str, kind: str = "fil"):
        if kind not in ("fil", "psrfits"):
            raise ValueError("kind is 'fil' or 'psrfits'")
        self._k = kind
        h = C.c_void_p()
        info = FilInfo()
        check(getattr(lib(), f"af_{kind}_open")(path.encode(), C.byref(h), C.byref(info)),
              f"af_{kind}_open")
        self._h = h.value
        self.info = info
        self.dtype = np.dtype(_DT[info.dtype])

    @property
    def nchans(self):
        return self.info.nchans

    @property
    def ndata(self):
        return int(self.info.ndata)

    def read_into(self, row0: int, nrows: int, ptr: int):
        """Rows [row0, row0 + nrows) into memory at ptr (e.g. a HostBuffer)."""
        check(getattr(lib(), f"af_{self._k}_read")(self._h, row0, nrows, C.c_void_p(ptr)),
              f"af_{self._k}_read")

    def read(self, row0: int = 0, nrows: int = None) -> np.ndarray:
        nrows = self.ndata - row0 if nrows is None else nrows
        out = np.empty((nrows, self.nchans), self.dtype)
        if nrows:
            self.read_into(row0, nrows, out.ctypes.data)
        return out

    def close(self):
        if self._h:
            getattr(lib(), f"af_{self._k}_close")(C.c_void_p(self._h))
            self._h = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

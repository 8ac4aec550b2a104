"""C-ABI library checks that need no GPU: it loads, exports every symbol the
header declares, struct layouts match, and calls fail cleanly without a GPU."""
import ctypes as C
import os
import re
import subprocess

import pytest

from astroflow_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))
           if f.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(af_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_functions():
    names = declared_functions()
    assert "af_plan_create" in names and "af_execute" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_python_binding_covers_header():
    assert set(declared_functions()) == set(capi.SIGNATURES)


def _c_sizeof(type_name, tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include "astroflow_amd.h"\n'
                   'int main(void){printf("%%zu\\n", sizeof(%s));return 0;}\n' % type_name)
    exe = tmp_path / "sz"
    subprocess.run(["gcc", f"-I{ROOT}/include", str(src), "-o", str(exe)], check=True)
    return int(subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True, check=True).stdout)


@pytest.mark.parametrize("cname,pytype", [("af_plan_desc", capi.PlanDesc),
                                          ("af_plan_info", capi.PlanInfo),
                                          ("af_stage_times", capi.StageTimes),
                                          ("af_device_props", capi.DeviceProps),
                                          ("af_server_desc", capi.ServerDesc),
                                          ("af_server_result", capi.ServerResult),
                                          ("af_server_stats", capi.ServerStats),
                                          ("af_snr_params", capi.SnrParams),
                                          ("af_snr_result", capi.SnrResult)])
def test_struct_layouts_match_c(cname, pytype, tmp_path):
    assert _c_sizeof(cname, tmp_path) == C.sizeof(pytype)


def test_version_and_error_strings():
    assert b"gfx950" in capi.lib().af_version()
    assert isinstance(capi.lib().af_last_error(), bytes)


@pytest.mark.skipif(capi.device_count() > 0, reason="a GPU is visible")
def test_plan_create_without_gpu_reports_nodev():
    import numpy as np
    from tests.helpers import make_spec, small_band
    x = np.zeros((64, 8), np.uint8)
    spec = make_spec(x, small_band(8), 0.0, 1.0, 4, method="direct")
    with pytest.raises(capi.AFError, match="No HIP devices"):
        capi.Plan(spec)


def test_batch_rejects_the_same_plan_twice():
    # the duplicate check precedes any use of the plans: opaque non-null
    # handles are enough to reach it without a GPU
    lib = capi.lib()
    token = C.create_string_buffer(64)
    h = C.cast(token, C.c_void_p).value
    bufs = (C.c_void_p * 2)(C.c_void_p(h), C.c_void_p(h))
    plans = (C.c_void_p * 2)(C.c_void_p(h), C.c_void_p(h))
    out = C.c_void_p()
    rc = lib.af_batch_create(plans, 2, bufs, bufs, None, C.byref(out))
    assert rc == capi.AF_ERR_INVALID
    assert b"more than once" in lib.af_last_error()
    assert not out.value


def test_image_slices_rejects_bad_arguments_before_any_device_call():
    lib = capi.lib()
    p = C.c_void_p(4096)   # never dereferenced: every case fails its argument checks first
    cases = [
        (p, 9, 10, 10, 10, 5, 512),                 # no such plane element type
        (p, capi.AF_U32, 10, 10, 10, 0, 512),       # zero slice width
        (p, capi.AF_U32, 0, 10, 10, 5, 512),        # no rows
        (p, capi.AF_F32, 10, 20, 10, 5, 512),       # ld < cols
        (p, capi.AF_U64, 10, 10, 10, 5, 0),         # image size 0
    ]
    for plane, dt, rows, cols, ld, sps, size in cases:
        rc = lib.af_image_slices(plane, dt, rows, cols, ld, sps, size, p, None)
        assert rc == capi.AF_ERR_INVALID, (dt, rows, cols, ld, sps, size)
    assert lib.af_image_slices(None, capi.AF_F32, 1, 1, 1, 1, 8, p, None) == capi.AF_ERR_INVALID
    assert lib.af_execute_host_images(None, p, 10, 512, p, None) == capi.AF_ERR_INVALID


def test_frb_snr_without_data_returns_no_method():
    # an empty spectrum answers before any device call
    assert capi.frb_snr(0, capi.AF_F32, 0, 4) == (-1.0, 0, 0, (0.0, 1.0, {"fit_converged": False}))
    with pytest.raises(ValueError):
        capi.frb_snr(0, 9, 10, 4)

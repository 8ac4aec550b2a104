"""Builds the native parts of astroflow_amd in-tree for gfx950.

  libastroflow_amd.so   HIP kernels + C-ABI (hipcc, --offload-arch=gfx950)
  _astroflow_core*.so   pybind11 drop-in module over the C-ABI (g++)

Usage: python -m astroflow_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libastroflow_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("AF_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: floating-point expressions round exactly as written
# (the delay tables, IQRM statistics and zero-DM means must match the oracle
# bit for bit).  Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INC}", f"-I{CSRC}",
]

HIP_SOURCES = ["af_kernels.hip", "af_plan.hip"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "argument unused" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


OBJ_DIR = os.path.join(PKG, "_build")


def build_lib(verbose=False):
    """Compiles every .hip source to an object (in parallel: one hipcc per
    source), then links the shared library."""
    from concurrent.futures import ThreadPoolExecutor
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    extra = [os.path.join(CSRC, s) for s in os.listdir(CSRC)
             if s.endswith(".hip") and s not in HIP_SOURCES]
    srcs += sorted(extra)
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = [os.path.join(OBJ_DIR, os.path.basename(s)[:-4] + ".o") for s in srcs]
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", "0")) or (os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        outs = list(ex.map(lambda so: _run([HIPCC, *HIP_FLAGS, "-c", so[0], "-o", so[1]]),
                           zip(srcs, objs)))
    tmp = LIB + ".tmp"
    outs.append(_run([HIPCC, f"--offload-arch={ARCH}", "-shared", *objs, "-o", tmp]))
    os.replace(tmp, LIB)
    out = "\n".join(o for o in outs if o.strip())
    if verbose and out.strip():
        print(out)
    return LIB


def core_module_path():
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(PKG, "_astroflow_core" + suffix)


def build_core(verbose=False):
    src = os.path.join(CSRC, "af_bindings.cpp")
    if not os.path.exists(src):
        return None
    import pybind11  # noqa: WPS433
    out_path = core_module_path()
    tmp = out_path + ".tmp"
    cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fopenmp",
           "-fvisibility=hidden", f"-I{INC}", f"-I{CSRC}", f"-I{pybind11.get_include()}",
           f"-I{sysconfig.get_paths()['include']}",
           src, "-o", tmp, f"-L{PKG}", "-lastroflow_amd", "-Wl,-rpath,$ORIGIN"]
    out = _run(cmd)
    os.replace(tmp, out_path)
    if verbose and out.strip():
        print(out)
    return out_path


def build_all(verbose=False):
    lib = build_lib(verbose)
    core = build_core(verbose)
    return lib, core


if __name__ == "__main__":
    lib, core = build_all(verbose=True)
    print("built", lib)
    if core:
        print("built", core)
    sys.exit(0)

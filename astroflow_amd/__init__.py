"""astroflow_amd -- MI355X (gfx950) dedispersion + RFI hot path of astroflow.

Native parts: ``libastroflow_amd.so`` (HIP kernels + C-ABI, include/astroflow_amd.h)
and the pybind11 drop-in module ``_astroflow_core``.  Build in-tree with
``python -m astroflow_amd.build``.
"""
__version__ = "0.1.0"

"""openqg_amd: OpenQG's QG timestep (step_qg) as hand-written FP64 HIP kernels for MI355X.

The compute path is liboqg.so (include/oqg.h); this package is its Python host
layer.  See `openqg_amd.qg.QG` and `openqg_amd.core.step_qg`.
"""
from ._lib import FLAG_YSOLVE_THOMAS, OqgError, load  # noqa: F401
from .core import init_dynamic_state, step_qg, suppress_comp_mode  # noqa: F401
from .qg import QG, Box, Topog, init_qg  # noqa: F401

__all__ = ["QG", "Box", "Topog", "init_qg", "step_qg", "suppress_comp_mode", "init_dynamic_state",
           "OqgError", "load", "FLAG_YSOLVE_THOMAS"]

"""MI355X-native SeisCL time-domain finite-difference engine.

The compute path is libseiscl_amd.so (hand-written gfx950 HIP kernels behind
the C-ABI of include/seiscl_amd.h); this package holds its Python bindings and
the SeisCL.py-compatible front end.
"""
from .engine import Batch, Context, Model, make_csts  # noqa: F401

__all__ = ["Batch", "Context", "Model", "make_csts"]

"""MI355X-native finite-element analysis path of the TopOpt3D topology optimizer.

The numerical work runs in libtofe.so (hand-written HIP kernels for gfx950 behind the
C ABI of include/tofe.h); this package binds it (``_lib``), mirrors the reference's
TopOpt/Compliance call surface (``topopt``) and builds the synthetic workloads
(``configs``).
"""
from . import configs  # noqa: F401
from ._lib import Context, TofeError, lib  # noqa: F401

__all__ = ["Context", "TofeError", "lib", "configs"]

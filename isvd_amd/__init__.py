"""MI355X-native iSVD: sketch -> orthogonalize -> integrate -> postprocess on one GPU.

The product is ``lib/libisvd_mi355x.so`` (hand-written HIP for gfx950 behind the C ABI of
``include/isvd.h``).  This package holds its build script and a thin Python layer over the C
ABI (``capi``) plus array conveniences (``pipeline``); there is no CPU code path.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]

"""MI355X-native photon-packet transport and per-cell balance for MOCASSIN.

The compute path is libmocassin_amd.so (HIP kernels for gfx950) behind the C
ABI declared in include/mocassin_amd.h; this package is its Python host
mirror.  Importing the package does not touch the GPU.
"""
from ._capi import load as load_library, McError  # noqa: F401
from .transport import Transport, scale_J, transport_batch  # noqa: F401

__all__ = ["Transport", "scale_J", "transport_batch", "load_library", "McError"]

"""pdekit_amd: MI355X-native residual-distribution cell loop.

A drop-in GPU path for the explicit continuous RD residual assembly of
pdekit (src/solver/rdm): hand-written gfx950 HIP kernels behind the C ABI of
include/pdekit_rd.h, with the host-side FE/quadrature/mesh inputs in C++.
"""
from . import _lib  # noqa: F401
from ._lib import PdrdError, lib  # noqa: F401

__all__ = ["lib", "PdrdError"]

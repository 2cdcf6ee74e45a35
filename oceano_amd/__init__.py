"""oceano_amd: MI355X-native per-stage DG right-hand side of the oceano shallow-water solver.

The compute path is liboceano_dg.so (hand-written gfx950 HIP kernels behind the C ABI in
include/oceano_dg.h).  Importing this package does not touch the GPU; the first call into
the library loads it and raises if it has not been built.
"""
from ._capi import lib, OceanoError  # noqa: F401

__all__ = ["lib", "OceanoError"]

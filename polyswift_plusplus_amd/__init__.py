"""psx-scft: MI355X-native SCFT hot path (split-step propagator, FFT pair,
contour quadrature, Q, Flory-Huggins update and free energy).

The compute path is libpsx_scft.so (HIP, gfx950) behind the C ABI in
include/psx_scft.h; this package is its Python host mirror.
"""
from ._abi import PSX_FP32, PSX_FP64, PsxError, load  # noqa: F401
from .scft import (Block, BlendSCFT, BlockCopolymerSCFT, BatchSCFT, DiblockSCFT, Mixing,  # noqa: F401
                   Polymer, PsFFT)
from . import fields  # noqa: F401

__all__ = ["Block", "BlendSCFT", "BlockCopolymerSCFT", "BatchSCFT", "DiblockSCFT", "Mixing",
           "Polymer", "PsFFT",
           "PsxError", "load", "fields",
           "PSX_FP64", "PSX_FP32"]

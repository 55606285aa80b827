"""The forward / backward op interface of the SCFT evaluation.

``DensityOp.forward`` maps the conjugate fields and the box lengths to the
densities and the partition function, (w_A, w_B, L) -> (phi_A, phi_B, Q):
PsFlexPseudoSpec::reset + solveQ, PsBlock::calcBigQ and
setCalcQQTIntegral + setPhysFields of the reference
(pspolymer/PsFlexPseudoSpec.cpp:159-293, pspolymer/PsBlock.cpp:154-380).
``DensityOp.backward`` is its reverse mode: given the cotangents of phi_A,
phi_B and Q it returns those of w_A, w_B and L, from the hand-written adjoint
sweeps of the engine (psx_scft_backward).  The reference has no derivatives;
this interface is new.
"""
from __future__ import annotations

import numpy as np

from .scft import DiblockSCFT


class DensityOp:
    """(w_A, w_B, box) -> (phi_A, phi_B, Q) on an engine (a diblock or a
    linear multiblock in FP64 with every slice stored)."""

    def __init__(self, engine: DiblockSCFT):
        self.e = engine
        self.shape = engine.shape

    def forward(self, wA, wB, box=None):
        """Densities and Q at the fields (and box lengths, when given)."""
        if box is not None:
            self.e.setCell([float(L) / n for L, n in zip(box, self.shape)])
        self.e.setFields(wA, wB)
        self.e.solveQ()
        Q = self.e.calcBigQ()
        phiA, phiB = self.e.densities()
        return phiA, phiB, Q

    def backward(self, g_phiA=None, g_phiB=None, g_Q: float = 0.0):
        """(dL/dw_A, dL/dw_B, dL/dL_box) for L = <g_phiA, phi_A> +
        <g_phiB, phi_B> + g_Q Q at the last forward point."""
        return self.e.backward(g_phiA, g_phiB, g_Q, box=True)

    def box(self) -> np.ndarray:
        return np.array(self.shape, dtype=float) * np.array(self.e.cell)

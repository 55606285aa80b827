"""Solvers built on the engine: converge an SCFT saddle point, and relax the
box lengths to zero stress (cell optimisation).

The reference has neither (it runs a fixed number of steepest-descent
steps, polyswift/polyswift.cxx:544, and has no stress); both are new.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .scft import DiblockSCFT, Mixing

# Anderson settings that converge the diblock workloads here: Picard with
# lambda0 = 0.2 until the residual RMS is below 0.05, then depth-10 Anderson
# with beta = 1.
DEFAULT_MIXING = Mixing("anderson", lambda0=0.2, depth=10, beta=1.0, anderson_start=0.05)


def converge(e: DiblockSCFT, tol: float = 1e-10, max_iters: int = 5000,
             mixing: Mixing | None = None) -> dict:
    """Iterate to residual RMS < tol, then sweep once more at the final
    fields so that slices, Q and densities belong to them.  The histories
    carry ``converged``: whether the last residual RMS is below tol (False
    when max_iters ran out first)."""
    h = e.iterate(max_iters, mixing or DEFAULT_MIXING, tol=tol)
    h["converged"] = bool(h["iters"] > 0 and h["err"][-1] < tol)
    e.update()
    return h


def tail_mixing(mix: Mixing, depth: int = 30, beta: float = 0.5) -> Mixing:
    """The FP64 tail's mixing: Anderson with a deeper history and a mixing
    parameter of at most 0.5.  From the FP32 phase's end (residual RMS 2e-5
    at 128^3) depth 30 reached 1e-8 in 140 iterations where depth 10 took
    367, and beta 0.5 took 131 and 115 iterations where beta 1 took 149 and
    121 from the same fields; started from scratch, FP64 converges fastest
    at depth 10."""
    if mix.kind != "anderson":
        return mix
    return dataclasses.replace(mix, depth=max(mix.depth, depth), beta=min(mix.beta, beta))


def mixed_precision_solve(e: DiblockSCFT, tol: float = 1e-8, switch_tol: float = 2e-5,
                          max_iters: int = 5000, mixing: Mixing | None = None,
                          tail: Mixing | None = None, keep_history: bool = False) -> dict:
    """FP32 sweeps until the residual RMS falls below switch_tol (the FP32
    noise floor is near 1e-6), then FP64 iterations to tol.  The FP64 tail
    starts a fresh Anderson history (``tail``, by default
    ``tail_mixing(mixing)``), or with ``keep_history`` carries the FP32
    phase's history on (converted exactly to FP64) with the same mixing.  Q,
    the quadrature, the update and F are FP64 throughout."""
    mix = mixing or DEFAULT_MIXING
    e.setPrecision("fp32")
    h32 = e.iterate(max_iters, mix, tol=switch_tol)
    e.setPrecision("fp64")
    if keep_history:
        tail = mix
    else:
        e.resetHistory()
    h64 = e.iterate(max_iters, tail or tail_mixing(mix), tol=tol)
    return dict(fp32_iters=h32["iters"], fp64_iters=h64["iters"], F=h64["F"][-1] if h64["iters"] else None,
                err=h64["err"][-1] if h64["iters"] else None, h32=h32, h64=h64)


def relax_box(e: DiblockSCFT, tol_stress: float = 1e-7, tol_scft: float = 1e-11,
              max_outer: int = 40, first_step: float = 5.0, axes=None,
              mixing: Mixing | None = None, max_iters: int = 5000, log=None) -> dict:
    """Relax box lengths L_d = n_d * cell_d to dF/dL_d = 0 by a secant step
    per axis (the orthorhombic cell keeps its angles).  Each outer step
    converges the fields from the previous ones, then takes the stress from
    the adjoint of the converged sweep.

    The relaxation stops, with ``converged`` False and ``reason`` set, when an
    inner solve misses tol_scft within max_iters (its stress belongs to
    unconverged fields, so no step is taken from it) or when max_outer runs
    out.  ``L`` is always the box of the last evaluated step, and the engine
    is left at that box."""
    n = np.array(e.shape, dtype=float)
    cell = np.array(e.cell, dtype=float)
    active = np.array([s > 1 for s in e.shape]) if axes is None else np.array(axes)
    hist = []
    L_prev = g_prev = None
    reason = "max_outer"
    for outer in range(max_outer):
        h = converge(e, tol=tol_scft, mixing=mixing, max_iters=max_iters)
        g = e.boxStress()
        F = e.getFreeE()
        L = n * cell
        hist.append(dict(L=L.copy(), dFdL=g.copy(), F=F, iters=int(h["iters"]),
                         scft_err=float(h["err"][-1]) if h["iters"] else None,
                         scft_converged=h["converged"]))
        if log:
            log(f"outer {outer}: {h['iters']} iterations (err {hist[-1]['scft_err']}), L = {L}, "
                f"dF/dL = {g}, F = {F:.12f}")
        if not h["converged"]:
            reason = "scft_not_converged"
            break
        if np.max(np.abs(g[active])) < tol_stress:
            reason = "stress"
            break
        step = np.zeros(3)
        for d in np.where(active)[0]:
            if L_prev is None or g[d] == g_prev[d]:
                step[d] = -first_step * g[d]
            else:
                step[d] = -g[d] * (L[d] - L_prev[d]) / (g[d] - g_prev[d])
        if outer == max_outer - 1:
            break          # leave the engine at the last evaluated box
        L_prev, g_prev = L, g
        cell = (L + step) / n
        e.setCell(cell)
    last = hist[-1]
    return dict(history=hist, L=last["L"].copy(), reason=reason,
                converged=bool(reason == "stress" and last["scft_converged"]))

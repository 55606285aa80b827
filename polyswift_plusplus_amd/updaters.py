"""Updaters that run beside the field update: the reference's spectral
filters and Poisson updater, over the HIP FFT plugin.

* ``UpdaterSchedule`` is PsUpdater's ``applyStart`` / ``applyFrequency`` /
  ``applyEnd`` gate (psbase/PsUpdater.cpp:60-118,180-206).
* ``SpecFilterUpdater`` is ``simpleSpecFilter`` (no ``specCellSizes``) or
  ``multiSpecFilter`` (pseffhamil/PsSimpleSpecFilter.cpp:74-144,
  pseffhamil/PsMultiSpecFilter.cpp:148-267) acting on an engine's conjugate
  fields in place (``psx_scft_spec_filter``).
* ``PoissonUpdater`` is ``PsPoissonUpdater::update``
  (pseffhamil/PsPoissonUpdater.cpp:214-289) on a (charge, psi) field pair
  held by the caller (``psx_fft_poisson_update``).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


class UpdaterError(ValueError):
    """An updater configuration the reference rejects."""


@dataclass
class UpdaterSchedule:
    """When an updater applies: at steps t with applyStart <= t <= applyEnd
    and (t - applyStart) % applyFrequency == 0.  applyEnd defaults to
    nsteps + the step the run starts from, as the reference sets it."""
    applyStart: int = 0
    applyFrequency: int = 1
    applyEnd: int | None = None

    def resolve(self, nsteps: int, start_step: int = 0) -> "UpdaterSchedule":
        end = self.applyEnd if self.applyEnd is not None else nsteps + start_step
        s = UpdaterSchedule(int(self.applyStart), int(self.applyFrequency), int(end))
        if s.applyFrequency < 1:
            raise UpdaterError("applyFrequency must be >= 1")
        if s.applyStart >= s.applyEnd:
            raise UpdaterError("applyStart >= applyEnd")
        if s.applyEnd - s.applyStart <= s.applyFrequency:
            raise UpdaterError("applyFrequency too large")
        return s

    def due(self, t: int) -> bool:
        end = self.applyEnd if self.applyEnd is not None else np.iinfo(np.int64).max
        if t < self.applyStart or t > end:
            return False
        return (t - self.applyStart) % self.applyFrequency == 0

    def next_due(self, t: int) -> int | None:
        """The first step > t at which the updater applies, or None."""
        s = max(t + 1, self.applyStart)
        r = (s - self.applyStart) % self.applyFrequency
        if r:
            s += self.applyFrequency - r
        end = self.applyEnd if self.applyEnd is not None else np.iinfo(np.int64).max
        return s if s <= end else None


@dataclass
class SpecFilterUpdater:
    """Spectral filter on the conjugate fields of ``fields`` ("A", "B",
    "AB"): mean removed, then the modes with |FFT(w)|^2 below cutoffFactor
    times the maximum (of the whole spectrum, or of each spectral cell of
    ``specCellSizes`` cells) scaled by filterStrength."""
    fields: str = "AB"
    cutoffFactor: float = 0.0
    filterStrength: float = 0.0
    specCellSizes: tuple | None = None
    schedule: UpdaterSchedule = field(default_factory=UpdaterSchedule)
    name: str = "specFilter"
    # placed before the field update in updaterSequence: it then filters the
    # fields between the step's densities and its field update
    before: bool = False

    def apply(self, engine) -> None:
        if self.cutoffFactor <= 0.0:      # the reference returns before anything
            return
        engine.specFilter(self.fields, self.cutoffFactor, self.filterStrength, self.specCellSizes)

    def update(self, engine, t: int) -> bool:
        """Apply if due at step t; True if it applied."""
        if not self.schedule.due(t):
            return False
        self.apply(engine)
        return True


@dataclass
class PoissonUpdater:
    """The electrostatic potential update of a charged field pair."""
    bjerrumLen: float
    cell: tuple
    schedule: UpdaterSchedule = field(default_factory=UpdaterSchedule)

    def update(self, fft, charge, psi, t: int):
        """(charge, psi) after the update at step t (unchanged if not due);
        ``fft`` is a PsFFT plan of the grid."""
        if not self.schedule.due(t):
            return charge, psi
        return fft.poissonUpdate(charge, psi, self.bjerrumLen, self.cell)

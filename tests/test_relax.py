"""Host logic of the box relaxation (relax.relax_box, relax.converge) on a
stand-in engine whose free energy is a quadratic bowl in the box lengths:
the secant steps reach zero stress, an inner solve that misses its tolerance
stops the relaxation without a step from its stress, and the reported box is
always the last evaluated one."""
import numpy as np

from polyswift_plusplus_amd.relax import converge, relax_box


class BowlEngine:
    """dF/dL_d = k_d (L_d - L0_d); iterate 'converges' unless told to fail."""

    def __init__(self, shape=(8, 8, 1), cell=(1.0, 1.1, 1.0), L0=(9.0, 8.0, 1.0), fail_at=None):
        self.shape, self.cell = shape, tuple(cell)
        self.L0 = np.array(L0)
        self.k = np.array([0.3, 0.5, 0.0])
        self.fail_at = fail_at
        self.solves = 0
        self.cells_set = []

    def iterate(self, n, mix, tol=0.0):
        self.solves += 1
        err = 10 * tol if self.solves == self.fail_at else 0.1 * tol
        return dict(F=np.zeros(3), Q=np.ones(3), err=np.array([1.0, 0.5, err]), iters=3)

    def update(self):
        pass

    def L(self):
        return np.array(self.shape) * np.array(self.cell)

    def boxStress(self):
        return self.k * (self.L() - self.L0)

    def getFreeE(self):
        return float(0.5 * np.sum(self.k * (self.L() - self.L0) ** 2))

    def setCell(self, cell):
        self.cell = tuple(cell)
        self.cells_set.append(tuple(cell))


def test_converge_reports_tolerance():
    e = BowlEngine()
    assert converge(e, tol=1e-8)["converged"]
    e = BowlEngine(fail_at=1)
    assert not converge(e, tol=1e-8)["converged"]


def test_relax_reaches_zero_stress():
    e = BowlEngine()
    r = relax_box(e, tol_stress=1e-9, first_step=1.0)
    assert r["converged"] and r["reason"] == "stress"
    assert np.allclose(r["L"][:2], e.L0[:2], atol=1e-8)
    assert np.array_equal(r["L"], e.L())


def test_relax_stops_when_a_solve_misses_tolerance():
    e = BowlEngine(fail_at=3)
    r = relax_box(e, tol_stress=1e-12, first_step=1.0)
    assert not r["converged"] and r["reason"] == "scft_not_converged"
    assert len(r["history"]) == 3 and not r["history"][-1]["scft_converged"]
    # no step was taken from the unconverged solve's stress
    assert len(e.cells_set) == 2
    assert np.array_equal(r["L"], r["history"][-1]["L"]) and np.array_equal(r["L"], e.L())


def test_relax_out_of_outer_steps_reports_last_evaluated_box():
    e = BowlEngine()
    # a tiny first step; the secant step after it would be exact, but two
    # outer steps allow no third evaluation
    r = relax_box(e, tol_stress=1e-9, max_outer=2, first_step=1e-3)
    assert not r["converged"] and r["reason"] == "max_outer"
    assert len(r["history"]) == 2 and len(e.cells_set) == 1
    assert np.array_equal(r["L"], r["history"][-1]["L"]) and np.array_equal(r["L"], e.L())

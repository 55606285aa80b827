"""Batched parameter sweeps: every instance of a batch follows the same
trajectory as the same system run alone."""
import numpy as np
import pytest

from polyswift_plusplus_amd import BatchSCFT, DiblockSCFT, Mixing, fields

pytestmark = pytest.mark.gpu

SPECS = [dict(shape=(16, 16, 16), cell=(0.3,) * 3, fA=0.30, ds=0.05, chiN=14.0),
         dict(shape=(16, 16, 16), cell=(0.3,) * 3, fA=0.45, ds=0.05, chiN=20.0),
         dict(shape=(32, 16, 8), cell=(0.25, 0.3, 0.35), fA=0.35, ds=0.05, chiN=17.0),
         dict(shape=(12, 12, 12), cell=(0.4,) * 3, fA=0.40, ds=0.05, chiN=12.0),
         dict(shape=(16, 16, 16), cell=(0.28,) * 3, fA=0.25, ds=0.05, chiN=25.0)]
MIX = Mixing("anderson", lambda0=0.2, depth=5, beta=1.0, anderson_start=0.05)


def _fields(i, sp):
    return fields.seeded_fields(sp["shape"], sp["cell"], sigma=0.3, seed=100 + i, seed_kind="lamellar")


def _single(sp, i, prec, n):
    e = DiblockSCFT(sp["shape"], sp["cell"], sp["fA"], sp["ds"], sp["chiN"], prec=prec)
    e.setFields(*_fields(i, sp))
    return e.iterate(n, MIX)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_matches_single_instances(prec, monkeypatch):
    """Like for like: the batch runs its quadrature after the sweeps, so the
    single-instance references do too (PSX_OVERLAP_QUAD=0, the default)."""
    print(f"batch {prec}", flush=True)
    monkeypatch.setenv("PSX_OVERLAP_QUAD", "0")
    b = BatchSCFT(SPECS, prec=prec)
    assert b.nclasses == 3
    for i, sp in enumerate(SPECS):
        b.instances[i].setFields(*_fields(i, sp))
    h = b.iterate(8, MIX)
    for i, sp in enumerate(SPECS):
        ref = _single(sp, i, prec, 8)
        assert np.max(np.abs(h["F"][i] - ref["F"])) <= 1e-14 * np.max(np.abs(ref["F"])), i
        wA, _, _ = b.instances[i].getFields()
    b.close()


def test_batch_autotune_keeps_results():
    b = BatchSCFT(SPECS, prec="fp64")
    for i, sp in enumerate(SPECS):
        b.instances[i].setFields(*_fields(i, sp))
    choice = b.autotune(3)
    print("autotune choices per size class:", choice, flush=True)
    h = b.iterate(6, MIX)
    for i, sp in enumerate(SPECS):
        ref = _single(sp, i, "fp64", 6)
        assert np.max(np.abs(h["F"][i] - ref["F"])) <= 1e-11 * np.max(np.abs(ref["F"])), i
    b.close()


def test_batch_rejects_mixed_step_counts():
    from polyswift_plusplus_amd import PsxError
    specs = [dict(SPECS[0]), dict(SPECS[1], ds=0.025)]
    with pytest.raises(PsxError):
        BatchSCFT(specs)
